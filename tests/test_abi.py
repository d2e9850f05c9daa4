"""CPU checks of the C-ABI boundary: libffm.so loads and exports every entry
point include/ffm.h declares, and the Python mirror builds descriptors
exactly as the header lays them out. No kernel runs here."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import pyfae
from pyfae import _abi, configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ffm.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ffm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(pyfae.LIB_PATH), "build libffm.so first (fflame-rust_amd/build.py)"
    out = subprocess.run(["nm", "-D", "--defined-only", pyfae.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (ffm_[a-z0-9_]+)", out))
    missing = [s for s in declared() if s not in exported]
    assert not missing, missing
    extra = sorted(s for s in exported if s not in declared())
    assert not extra, extra


def test_library_loads_and_binds():
    lib = pyfae.library()
    assert lib.ffm_abi_version() == 10
    bound = {name for name, _, _ in _abi.SIGNATURES}
    assert bound == set(declared())


def test_struct_layout_matches_header():
    # sizes as a C compiler lays out include/ffm.h (checked with gcc here)
    src = r"""
#include <stdio.h>
#include <stddef.h>
#include "ffm.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(ffm_xform), sizeof(ffm_flame_desc),
         sizeof(ffm_view), sizeof(ffm_tonemap_params), sizeof(ffm_job), sizeof(ffm_render_stats),
         offsetof(ffm_job, tonemap), offsetof(ffm_flame_desc, finals));
  printf("%zu %zu %zu %zu %zu\n", sizeof(ffm_palette_xf), sizeof(ffm_scene),
         offsetof(ffm_scene, pan_x), offsetof(ffm_scene, tonemap), offsetof(ffm_scene, name));
  printf("%zu %zu %zu %zu\n", sizeof(ffm_effect), sizeof(ffm_effects),
         offsetof(ffm_effects, pixel_scale), offsetof(ffm_scene, effects));
  return 0;
}
"""
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "l.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "l")
        subprocess.run(["gcc", "-I" + os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        got = list(map(int, subprocess.run([exe], capture_output=True, text=True).stdout.split()))
    want = [C.sizeof(_abi.Xform), C.sizeof(_abi.FlameDesc), C.sizeof(_abi.View),
            C.sizeof(_abi.TonemapParams), C.sizeof(_abi.Job), C.sizeof(_abi.RenderStats),
            _abi.Job.tonemap.offset, _abi.FlameDesc.finals.offset,
            C.sizeof(_abi.PaletteXf), C.sizeof(_abi.Scene), _abi.Scene.pan_x.offset,
            _abi.Scene.tonemap.offset, _abi.Scene.name.offset,
            C.sizeof(_abi.Effect), C.sizeof(_abi.Effects), _abi.Effects.pixel_scale.offset,
            _abi.Scene.effects.offset]
    assert got == want


def test_flame_desc_local_index_order():
    fd = pyfae.flame_desc([pyfae.xform(variations={"spherical": 0.5, "linear": 0.5}),
                           pyfae.xform(variations={"julian": 1.0}, params={"julian.power": 3})],
                          [pyfae.xform(variations={"curl": 1.0})])
    assert [fd.var_id[k] for k in range(fd.n_vars)] == [2, 0, 24, 47]
    assert fd.xforms[1].var_weight[2] == 1.0 and fd.xforms[1].var_param[2][0] == 3.0
    assert fd.xforms[1].var_param[2][1] == 1.0  # julian dist default
    assert fd.finals[0].var_param[3][0] == 1.0  # curl c1 default
    with pytest.raises(pyfae.EngineError):
        pyfae.flame_desc([pyfae.xform(variations={"nope": 1.0})])
    with pytest.raises(pyfae.EngineError):
        pyfae.flame_desc([pyfae.xform()] * 129)


def test_configs_are_deterministic_and_shaped():
    a, pa, ja = configs.plain()
    b, pb, jb = configs.plain()
    assert bytes(a) == bytes(b) and np.array_equal(pa, pb)
    assert ja.walkers == 1 << 16 and ja.ipt == 1000 and ja.view.width == 512
    f1, _, j1 = configs.single4k()
    assert f1.n_xforms == 6 and f1.n_finals == 1 and j1.walkers == 1 << 24
    assert (j1.view.width, j1.view.height) == (3840, 2160)
    for i in range(6):
        ws = [f1.xforms[i].var_weight[k] for k in range(f1.n_vars)]
        assert sum(1 for w in ws if w != 0) == 3 and abs(sum(ws) - 1) < 1e-6
    m = np.array([[f1.xforms[0].a, f1.xforms[0].b], [f1.xforms[0].c, f1.xforms[0].d]])
    assert abs(np.linalg.svd(m, compute_uv=False)[0] - 0.8) < 1e-6
