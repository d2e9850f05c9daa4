"""The reference's own 2D flame presets (fixture tests/golden/ref_presets_2d.fae,
copied from assets/presets.fae) rendered on the GPU against the oracle: their
transcendental and random variations (bubble, pre_blur, julian, disc, blur,
noise, ...) keep GPU and oracle walks apart only by last-bit differences, so
block densities differ no more than two oracle seeds differ from each other."""
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402
import pyfae  # noqa: E402

pytestmark = pytest.mark.gpu
TEXT = open(os.path.join(os.path.dirname(__file__), "golden", "ref_presets_2d.fae")).read()


def block_sums(acc, b=8):
    h, w = acc.shape[:2]
    return acc[: h // b * b, : w // b * b, 3].reshape(h // b, b, w // b, b).sum((1, 3))


@pytest.mark.parametrize("i", range(6))
def test_preset_within_noise_floor(engine, i):
    sc = pyfae.load(TEXT, i)
    print(f"preset {i} {sc.name.decode()}", flush=True)
    engine.set_scene(sc)
    stops = pyfae.scene_stops(sc)

    def job(seed):
        j = pyfae.scene_job(sc, 128, 96, iterations=2_000_000, iterations_per_thread=64)
        j.seed_base = seed
        j.de_max_radius = 0.0
        return j

    gacc, stats = engine.render_accum(job(12345))
    _, o1, tot = O.render(sc.flame, stops, job(12345), pal_size=sc.palette_size,
                          palette_xf=sc.palette_xf, threads=8)
    _, o2, _ = O.render(sc.flame, stops, job(777), pal_size=sc.palette_size,
                        palette_xf=sc.palette_xf, threads=8)
    assert stats.total_samples == tot
    bg, b1, b2 = block_sums(gacc), block_sums(o1), block_sums(o2)
    floor = np.abs(b1 - b2).sum() / b1.sum()
    diff = np.abs(bg - b1).sum() / b1.sum()
    print(f"  floor {floor:.4f} gpu-vs-oracle {diff:.4f}", flush=True)
    assert diff <= 1.5 * floor + 2e-3
