"""The reference's own 2D flame presets (assets/presets.fae, the six this
engine renders; fixture tests/golden/ref_presets_2d.fae): each loads, writes
back as a config that reloads to the same scene, and the oracle renders it."""
import os

import numpy as np
import pytest

import pyfae
from oracle import oracle as O

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "ref_presets_2d.fae")
TEXT = open(FIXTURE).read()
NAMES = ["Grand Julian", "Julian Disc", "Plastic", "Spherical3", "Square Tile", "Flower"]


def test_the_six_presets_load():
    assert pyfae.scene_count(TEXT) == 6
    for i, name in enumerate(NAMES):
        sc = pyfae.load(TEXT, i)
        assert sc.name.decode() == name and sc.render_3d == 0


@pytest.mark.parametrize("i", range(6))
def test_preset_round_trips_and_renders_on_the_oracle(i):
    sc = pyfae.load(TEXT, i)
    again = pyfae.load(pyfae.formats.scene_to_json(sc))
    assert pyfae.formats.scene_summary(again) == pyfae.formats.scene_summary(sc)
    job = pyfae.scene_job(sc, 64, 48, iterations=300_000, iterations_per_thread=64)
    job.seed_base = 12345
    rgba, acc, total = O.render(sc.flame, pyfae.scene_stops(sc), job, pal_size=sc.palette_size,
                                palette_xf=sc.palette_xf, threads=8)
    assert total >= 300_000 and acc[..., 3].sum() > 0
