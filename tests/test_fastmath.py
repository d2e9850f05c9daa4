"""The walker's sin/cos/atan2 (fflame-rust_amd/csrc/ffm_fastmath.h) built
for the host from the same text, swept against double-precision libm.
The host build differs from the device build only in its quotients (IEEE
division here, the refined hardware reciprocal on the device), so this pins
the reductions and polynomials: a wrong constant, a dropped reduction term
or a swapped octant fix-up shows as a large ulp error."""
import ctypes as C
import math
import os
import subprocess
import tempfile

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "fflame-rust_amd", "csrc")

HARNESS = r"""
#include <math.h>
#include <stdint.h>
#include "ffm_fastmath.h"
void fm_sincos(const float* x, float* s, float* c, int n) {
  for (int i = 0; i < n; i++) ffm_sincosf(x[i], &s[i], &c[i]);
}
void fm_atan2(const float* y, const float* x, float* t, int n) {
  for (int i = 0; i < n; i++) t[i] = ffm_atan2f(y[i], x[i]);
}
"""


@pytest.fixture(scope="module")
def fm():
    d = tempfile.mkdtemp()
    src, so = os.path.join(d, "fm.c"), os.path.join(d, "fm.so")
    open(src, "w").write(HARNESS)
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-I", CSRC, src, "-o",
                    so, "-lm"], check=True)
    lib = C.CDLL(so)
    fp = C.POINTER(C.c_float)
    lib.fm_sincos.argtypes = [fp, fp, fp, C.c_int]
    lib.fm_atan2.argtypes = [fp, fp, fp, C.c_int]
    return lib


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def ulps(got, want):
    """|got - want| in ulps of float32(want) (want in float64)."""
    w = np.abs(want.astype(np.float32))
    sp = np.spacing(np.maximum(w, np.float32(1e-38))).astype(np.float64)
    return np.abs(got.astype(np.float64) - want) / sp


def sincos(fm, x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    s, c = np.empty_like(x), np.empty_like(x)
    fm.fm_sincos(_p(x), _p(s), _p(c), len(x))
    return s, c


def atan2(fm, y, x):
    y = np.ascontiguousarray(y, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    t = np.empty_like(x)
    fm.fm_atan2(_p(y), _p(x), _p(t), len(x))
    return t


def test_sincos_within_two_ulp(fm):
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(-4, 4, 1_000_000), rng.uniform(-32768, 32768, 1_000_000),
                        np.linspace(-10, 10, 200_001),
                        # near multiples of pi/2, where the reduction cancels
                        (np.arange(-20000, 20000) * (np.pi / 2))]).astype(np.float32)
    s, c = sincos(fm, x)
    xd = x.astype(np.float64)
    es, ec = ulps(s, np.sin(xd)), ulps(c, np.cos(xd))
    assert es.max() <= 1.6 and ec.max() <= 1.6, (es.max(), ec.max())
    assert np.median(es) <= 0.5 and np.median(ec) <= 0.5


def test_sincos_quadrants_and_wide_arguments(fm):
    k = np.arange(-8, 9)
    s, c = sincos(fm, (k * np.pi / 2).astype(np.float32))
    np.testing.assert_allclose(s, np.round(np.sin(k * np.pi / 2)), atol=1e-6)
    np.testing.assert_allclose(c, np.round(np.cos(k * np.pi / 2)), atol=1e-6)
    # past 32768 the reduction runs in f64 before the same polynomials
    x = np.random.default_rng(4).uniform(-1e6, 1e6, 10000).astype(np.float32)
    s, c = sincos(fm, x)
    assert np.abs(s - np.sin(x.astype(np.float64))).max() < 1e-6
    assert np.abs(c - np.cos(x.astype(np.float64))).max() < 1e-6


def test_atan2_within_two_ulp(fm):
    rng = np.random.default_rng(5)
    n = 1_000_000
    mag = lambda: rng.uniform(-1, 1, n) * 10.0 ** rng.uniform(-4, 4, n)  # noqa: E731
    y, x = mag().astype(np.float32), mag().astype(np.float32)
    # the octant boundaries: |y| = |x|, an axis, a ratio just past 1/8
    ring = rng.uniform(0, 2 * np.pi, 100_000)
    y = np.concatenate([y, np.sin(ring), [1, 1, -1, -1, 0.125, 3, -2]]).astype(np.float32)
    x = np.concatenate([x, np.cos(ring), [1, -1, 1, -1, 1, 3, 2]]).astype(np.float32)
    t = atan2(fm, y, x)
    e = ulps(t, np.arctan2(y.astype(np.float64), x.astype(np.float64)))
    assert e.max() <= 1.6, e.max()
    assert np.median(e) <= 0.5


def test_atan2_signed_zeros_and_octants(fm):
    y = np.array([0.0, -0.0, 0.0, -0.0, 1.0, -1.0, 1.0, -1.0, 2.0, 1e-30], np.float32)
    x = np.array([1.0, 1.0, -1.0, -1.0, 0.0, 0.0, -0.0, -0.0, -2.0, 1.0], np.float32)
    t = atan2(fm, y, x)
    want = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    assert np.all(np.signbit(t) == np.signbit(want))
    np.testing.assert_allclose(t, want, rtol=2e-7, atol=0)
    # a swapped argument order would read atan2(x, y): pi/2 - t in the first octant
    assert abs(atan2(fm, np.float32([0.25]), np.float32([1.0]))[0] - math.atan2(0.25, 1.0)) < 1e-7


def test_effect_sincos_is_the_walker_method(fm):
    # the oracle's sin/cos (the variation bodies' and the colour effects')
    # restates the device's method: bit-identical to the header's, the f64
    # reduction past 32768 included
    rng = np.random.default_rng(6)
    x = np.concatenate([rng.uniform(-32768, 32768, 20000),
                        rng.uniform(-1e9, 1e9, 2000)]).astype(np.float32)
    s, c = sincos(fm, x)
    so, co = C.c_float(), C.c_float()
    for i in range(0, len(x), 7):
        O.lib().orc_fx_sincos(float(x[i]), C.byref(so), C.byref(co))
        assert np.float32(so.value) == s[i] and np.float32(co.value) == c[i], x[i]


def test_oracle_atan2_is_the_walker_method(fm):
    rng = np.random.default_rng(8)
    y = (rng.uniform(-1, 1, 20000) * 10.0 ** rng.uniform(-4, 4, 20000)).astype(np.float32)
    x = (rng.uniform(-1, 1, 20000) * 10.0 ** rng.uniform(-4, 4, 20000)).astype(np.float32)
    t = atan2(fm, y, x)
    for i in range(0, len(x), 5):
        assert np.float32(O.lib().orc_fm_atan2(float(y[i]), float(x[i]))) == t[i], (y[i], x[i])
