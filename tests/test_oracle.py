"""CPU checks that pin the oracle to the reference (no GPU).

Each test checks the oracle against something other than its own formula:
the reference's known-answer tests (values asserted in
src/renderer/compute_kernel.rs), the reference's rendered baseline image,
independent float64/float16 implementations (numpy) and properties the
mathematics fixes.
"""
import ctypes as C
import json
import math
import os
import struct

import numpy as np
import pytest

from oracle import oracle as O
import pyfae
from pyfae import _abi, configs

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def bits(f):
    return struct.unpack("<I", struct.pack("<f", f))[0]


def f32(u):
    return struct.unpack("<f", struct.pack("<I", int(u) & 0xFFFFFFFF))[0]


# ---------------------------------------------------------------- RNG

def np_pcg_hash(x):
    """PCG-RXS-M-XS-32 (O'Neill 2014) of one LCG step from x, in numpy."""
    x = np.asarray(x, dtype=np.uint64)
    m = np.uint64(0xFFFFFFFF)
    s = (x * np.uint64(747796405) + np.uint64(2891336453)) & m
    w = (((s >> ((s >> np.uint64(28)) + np.uint64(4))) ^ s) * np.uint64(277803737)) & m
    return (w >> np.uint64(22)) ^ w


def test_pcg_hash_matches_published_rxs_m_xs():
    xs = np.arange(0, 1 << 16, dtype=np.uint64) * np.uint64(65521)
    L = O.lib()
    got = np.array([L.orc_pcg_hash(int(x) & 0xFFFFFFFF) for x in xs[:4096]], dtype=np.uint64)
    assert np.array_equal(got, np_pcg_hash(xs[:4096] & np.uint64(0xFFFFFFFF)))


def test_pcg_hash_is_a_bijection_on_a_block():
    # LCG step (odd multiplier) and RXS-M-XS are both permutations of u32, so
    # the hash has no collisions; a dropped term would create some.
    xs = np.arange(0, 1 << 20, dtype=np.uint64)
    h = np_pcg_hash(xs)
    assert len(np.unique(h)) == len(xs)
    L = O.lib()
    for x in (0, 1, 2, 12345, 0xFFFFFFFF, 0x80000000):
        assert L.orc_pcg_hash(x) == int(np_pcg_hash(np.uint64(x)))


def test_lcg_multiplier_gives_full_period():
    # Hull-Dobell for m = 2^32: odd increment and a - 1 divisible by 4.
    assert (747796405 - 1) % 4 == 0
    # Walk a 2^16-state analogue of the same recurrence to its full period.
    a, c, m = 747796405 % (1 << 16), 2891336453 % (1 << 16) | 1, 1 << 16
    s, seen = 0, 0
    while True:
        s = (s * a + c) % m
        seen += 1
        if s == 0:
            break
    assert seen == m


def test_stream_matches_numpy_restatement_and_init():
    L = O.lib()
    for tid, seed in [(0, 1), (5, 12345), (65535, 2654435762), (1 << 20, 0xDEADBEEF)]:
        out = np.zeros(64, dtype=np.uint32)
        L.orc_rng_stream(tid, seed, 64, O.u32p(out))
        st = int(np_pcg_hash(np.uint64(tid ^ seed)))
        inc = ((int(np_pcg_hash(np.uint64((int(np_pcg_hash(np.uint64(tid))) + seed) & 0xFFFFFFFF))) << 1)
               & 0xFFFFFFFF) | 1
        ref = []
        for _ in range(64):
            old = st
            st = (old * 747796405 + inc) & 0xFFFFFFFF
            w = (((old >> ((old >> 28) + 4)) ^ old) * 277803737) & 0xFFFFFFFF
            ref.append((w >> 22) ^ w)
        assert out.tolist() == ref


def test_nextf_is_top_24_bits():
    r = O.Rng()
    r.state, r.inc = 0, 1
    L = O.lib()
    vals = [L.orc_rng_nextf(C.byref(r)) for _ in range(20000)]
    assert min(vals) >= 0.0 and max(vals) < 1.0
    assert all((v * (1 << 24)).is_integer() for v in vals)
    # the largest draw is 1 - 2^-24, never 1.0 (rng.wgsl:47-52)
    assert np.float32(0xFFFFFF) * np.float32(1.0 / 16777216.0) == np.float32(1.0 - 2.0 ** -24)
    # roughly uniform
    hist, _ = np.histogram(vals, bins=20, range=(0, 1))
    assert hist.min() > 800 and hist.max() < 1200


# ---------------------------------------------------------- selection

def test_select_brute_force_frequencies():
    fd = pyfae.flame_desc([pyfae.xform(weight=w) for w in (0.5, 0.0, 1.5, 2.0)])
    L = O.lib()
    counts = np.zeros(4)
    n = 1 << 14
    for i in range(n):
        u = np.float32(i) / np.float32(n)
        counts[L.orc_select(C.byref(fd), float(u))] += 1
    np.testing.assert_allclose(counts / n, [0.125, 0.0, 0.375, 0.5], atol=2.0 / n)
    # target == cumulative lands on the lower transform (<=)
    assert L.orc_select(C.byref(fd), 0.125) == 0
    # the zero-weight transform is chosen only by u = 0 exactly... never here
    assert L.orc_select(C.byref(fd), 0.0) == 0


# ----------------------------------------------------------------- view

def test_world_to_pixel_kat():
    # src/renderer/compute_kernel.rs:6498-6508: 64x64, zoom 1: x = -1.96875
    # lights column 0 of the middle row; x = -2.03125 is off the frame.
    v = pyfae.View(64, 64, 1.0, 0.0, 0.0, 0.0)
    px, py = C.c_int32(), C.c_int32()
    O.lib().orc_world_to_pixel(C.byref(v), -1.96875, 0.0, C.byref(px), C.byref(py))
    assert (px.value, py.value) == (0, 32)
    O.lib().orc_world_to_pixel(C.byref(v), -2.03125, 0.0, C.byref(px), C.byref(py))
    assert px.value == -1
    O.lib().orc_world_to_pixel(C.byref(v), 0.0, 0.0, C.byref(px), C.byref(py))
    assert (px.value, py.value) == (32, 32)


def test_world_to_pixel_rotation_closed_form():
    v = pyfae.View(200, 100, 2.0, 0.25, -0.5, math.pi / 2)
    px, py = C.c_int32(), C.c_int32()
    for x, y in [(0.3, 0.1), (-0.7, 0.45), (1.1, -0.2)]:
        O.lib().orc_world_to_pixel(C.byref(v), x, y, C.byref(px), C.byref(py))
        tx, ty = x - 0.25, y + 0.5
        rx, ry = -ty, tx  # rotation by +90 degrees
        ex = math.floor(100 + rx * 2 * 25)
        ey = math.floor(50 + ry * 2 * 25)
        assert abs(px.value - ex) <= 1 and abs(py.value - ey) <= 1


# ------------------------------------------------------------ OrbitFuses

def test_orbit_fuses_kats():
    """src/renderer/compute_kernel.rs:6056-6104, value for value."""
    L = O.lib()

    def new():
        return O.OrbitFuses()

    c = new()
    d = lambda *a: L.orc_orbit_fuses_dispatch(C.byref(c), *a)  # noqa: E731
    assert d(8192, 256, 20, 0, 0) == 8192 * 236
    assert d(8192, 256, 20, 0, 0) == 8192 * 256
    L.orc_orbit_fuses_free(C.byref(c))
    assert d(8192, 256, 20, 0, 0) == 8192 * 236
    L.orc_orbit_fuses_free(C.byref(c))
    assert d(64, 64, 150, 0, 0) == 0
    assert d(64, 64, 150, 0, 0) == 0
    assert d(64, 64, 150, 0, 0) == 64 * (192 - 150)
    assert d(64, 64, 150, 0, 0) == 64 * 64
    L.orc_orbit_fuses_free(C.byref(c))
    assert d(128, 256, 20, 0, 0) == 128 * 236
    assert d(64, 256, 20, 0, 0) == 64 * 256
    assert d(256, 256, 20, 0, 0) == 128 * 256 + 128 * 236
    assert d(256, 256, 20, 0, 0) == 256 * 256
    L.orc_orbit_fuses_free(C.byref(c))
    assert d(128, 64, 100, 0, 0) == 0
    assert d(64, 64, 100, 0, 0) == 64 * 28
    assert d(128, 64, 100, 0, 0) == 64 * 64 + 64 * 28
    L.orc_orbit_fuses_free(C.byref(c))
    assert d(1024, 256, 20, 16, 7) == 1024 * 236
    assert d(1024, 256, 20, 16, 3) == 960 * 256 + 64 * 236
    assert d(16, 256, 20, 16, 0xFFFFFFFF) == 15 * 256 + 236
    L.orc_orbit_fuses_free(C.byref(c))


# ------------------------------------------------------------ accumulate

def _acc_passes(prev, hist, n, count=None, mode=0):
    a = prev
    for _ in range(n):
        a = O.accumulate(hist, a, count=count, mode=mode)
    return a


def test_accumulate_long_count_keeps_every_hit():
    # compute_kernel.rs:6997-7026: 2^28 hits, three more a pass, 1000 passes.
    start = np.float32(1 << 28)
    prev = np.array([[[0.5, 0.5, 0.5, start]]], dtype=np.float32)
    hist = np.array([[[150, 150, 150, 300]]], dtype=np.uint32)
    f32sum = _acc_passes(prev, hist, 1000)
    assert f32sum[0, 0, 3] == start  # the f32 sum is stuck at 2^28
    count = np.zeros((1, 1, 2), dtype=np.uint32)
    ex = _acc_passes(prev, hist, 1000, count=count)
    truth = (1 << 28) + 3 * 1000
    assert (int(count[0, 0, 1]) << 32) | int(count[0, 0, 0]) == truth * 100
    assert abs(float(ex[0, 0, 3]) / truth - 1.0) < 2.0 * float(np.finfo(np.float32).eps)


def test_accumulate_count_carries():
    # compute_kernel.rs:7028-7052
    prev = np.array([[[0.5, 0.5, 0.5, 1.0]]], dtype=np.float32)
    count = np.array([[[0xFFFFFF00, 7]]], dtype=np.uint32)
    a = O.accumulate(np.array([[[0x100, 0x100, 0x100, 0x200]]], dtype=np.uint32), prev, count=count)
    assert count[0, 0].tolist() == [0x100, 8]
    want = ((8 << 32) | 0x100) / 100.0
    assert abs(float(a[0, 0, 3]) / want - 1.0) < 1e-7


def test_accumulate_count_seeded_at_threshold_and_reset():
    # compute_kernel.rs:7054-7100
    prev = np.array([[[0.5, 0.5, 0.5, 0.0]]], dtype=np.float32)
    hist = np.array([[[100, 100, 100, 200]]], dtype=np.uint32)
    a = _acc_passes(prev, hist, 4)
    assert a[0, 0, 3] == 8.0
    count = np.zeros((1, 1, 2), dtype=np.uint32)
    _acc_passes(a, hist, 1, count=count)
    assert count[0, 0].tolist() == [1000, 0]
    prev = np.array([[[0.5, 0.5, 0.5, 1000.0]]], dtype=np.float32)
    count = np.zeros((1, 1, 2), dtype=np.uint32)
    _acc_passes(prev, hist, 3, count=count)
    assert count[0, 0].tolist() == [100_600, 0]
    _acc_passes(np.zeros((1, 1, 4), np.float32), hist, 1, count=count)
    assert count[0, 0].tolist() == [200, 0]
    _acc_passes(prev, hist, 1, count=count, mode=2)
    assert count[0, 0].tolist() == [200, 0]


def test_accumulate_steady_colour_does_not_wander():
    # compute_kernel.rs:7102-7120
    hist = np.array([[[30, 30, 30, 100]]], dtype=np.uint32)
    first = O.accumulate(hist, np.zeros((1, 1, 4), np.float32))
    after = _acc_passes(first, hist, 2000)
    assert after[0, 0, :3].tobytes() == first[0, 0, :3].tobytes()
    assert first[0, 0, 0] == np.float32(30.0) / np.float32(100.0)


def test_accumulate_is_the_density_weighted_mean():
    rng = np.random.default_rng(3)
    hs = [rng.integers(0, 100, (4, 4, 4)).astype(np.uint32) for _ in range(5)]
    for h in hs:
        h[..., 3] = rng.integers(100, 1000, (4, 4)).astype(np.uint32)
        h[..., :3] = np.minimum(h[..., :3], h[..., 3:4])
    a = np.zeros((4, 4, 4), np.float32)
    for h in hs:
        a = O.accumulate(h, a)
    tot = sum(h.astype(np.float64) for h in hs)
    np.testing.assert_allclose(a[..., 3], tot[..., 3] / 100.0, rtol=1e-6)
    np.testing.assert_allclose(a[..., :3], tot[..., :3] / tot[..., 3:4], rtol=1e-5)


# --------------------------------------------------------------- palette

def test_f16_matches_numpy_float16():
    L = O.lib()
    vals = np.concatenate([np.linspace(0, 1, 20001, dtype=np.float32),
                           np.float32(2.0) ** -np.arange(10, 30, dtype=np.float32),
                           np.array([65504, 65519.99, 65520, 1e-8, 5.96e-8], np.float32)])
    for v in vals:
        h = L.orc_f32_to_f16(float(v))
        assert h == int(np.float16(v).view(np.uint16)), v
        assert L.orc_f16_to_f32(h) == float(np.float16(v))


def test_palette_table_of_xml_stops_is_the_stops():
    rng = configs.SplitMix64(5)
    stops = configs.random_palette(rng)
    lin = O.palette_linear(stops, 256).reshape(256, 3)
    want = np.power(np.maximum(stops[:, 1:].astype(np.float16).astype(np.float32), 0), np.float32(2.2))
    np.testing.assert_allclose(lin, want, rtol=2e-7)
    # f16 storage holds k/255 within 1/16 of a level (src/gpu/buffers.rs:3205-3214)
    k = np.arange(256) / 255.0
    assert np.max(np.abs(k.astype(np.float16).astype(np.float64) - k)) < 1.0 / 255.0 / 16.0


def test_palette_sample_edges_and_interpolation():
    stops = np.array([[0.25, 1, 0, 0], [0.75, 0, 0, 1]], np.float32)
    out = (C.c_float * 3)()
    L = O.lib()
    for t, want in [(0.0, (1, 0, 0)), (0.25, (1, 0, 0)), (0.5, (0.5, 0, 0.5)), (1.0, (0, 0, 1))]:
        L.orc_palette_sample(O.fp(stops), 2, t, out)
        np.testing.assert_allclose(list(out), want, atol=1e-7)
    for ci, size, want in [(0.0, 256, 0), (0.999, 256, 255), (1.0, 256, 255), (-0.5, 256, 0),
                           (0.5, 256, 128), (float("nan"), 256, 0)]:
        assert L.orc_palette_texel(ci, size) == want


# ------------------------------------------------------------ variations

PROBE = [(0.0, 0.0), (-0.0, 0.0), (1.0, 0.0), (0.0, 1.0), (-1.0, 0.0), (0.0, -1.0), (0.5, 0.5),
         (-0.5, 0.25), (1e-20, 1e-20), (3.0, -4.0), (1e5, 2e5), (-2.5, -0.125), (0.1, 0.9),
         # the rest of the reference's probe grid where WGSL pins the arithmetic
         # (src/probe/inputs.rs:44-99): signed zeros, diagonals, the unit circle
         (0.0, -0.0), (-0.0, -0.0), (1.0, -0.0), (-0.0, 1.0), (1.0, 1.0), (-1.0, 1.0),
         (-1.0, -1.0), (1.0, -1.0), (0.5, 0.3), (0.12345679, -0.9876543), (2.5, -1.5),
         (math.pi, math.e), (0.70710677, 0.70710677)]


def _var(name, x, y, params=(0, 0, 0, 0), xf=None, rng_state=(0, 1)):
    L = O.lib()
    p = (C.c_float * 4)(*params)
    xfm = xf if xf is not None else pyfae.Xform()
    r = O.Rng()
    r.state, r.inc = rng_state
    ox, oy = C.c_float(), C.c_float()
    L.orc_variation(pyfae.VARIATIONS[name], p, C.byref(xfm), x, y, C.byref(r), C.byref(ox),
                    C.byref(oy))
    return ox.value, oy.value, r


def _near(a, b, rel=2e-6, abs_=1e-6):
    return abs(a - b) <= abs_ + rel * abs(b)


def test_variations_against_float64_closed_forms():
    for x, y in PROBE:
        X, Y = float(np.float32(x)), float(np.float32(y))
        r2 = X * X + Y * Y
        r = math.sqrt(r2)
        cases = {
            "linear": (X, Y),
            "sinusoidal": (math.sin(X), math.sin(Y)),
            "spherical": (X / (r2 + 1e-6), Y / (r2 + 1e-6)),
            "swirl": (X * math.sin(r2) - Y * math.cos(r2), X * math.cos(r2) + Y * math.sin(r2)),
            "horseshoe": ((X - Y) * (X + Y) / (r + 1e-6), 2 * X * Y / (r + 1e-6)),
            "polar": (math.atan2(X, Y) / math.pi, r - 1.0),
            "disc": (math.atan2(X, Y) / math.pi * math.sin(math.pi * r),
                     math.atan2(X, Y) / math.pi * math.cos(math.pi * r)),
            "hyperbolic": (X / (r2 + 1e-6), Y),
            "bubble": (X / (r2 / 4 + 1), Y / (r2 / 4 + 1)),
        }
        for name, (ex, ey) in cases.items():
            # sin/cos of a large argument: the float32 rounding of the argument
            # itself dominates (r2 for swirl, pi*r for disc, x for sinusoidal)
            if name in ("swirl", "disc", "sinusoidal") and r2 > 1e4:
                continue
            ox, oy, _ = _var(name, x, y)
            assert _near(ox, ex, 4e-6, 1e-6) and _near(oy, ey, 4e-6, 1e-6), (name, x, y, ox, oy, ex, ey)


def test_variation_invariants():
    for x, y in PROBE[2:]:
        r = math.hypot(np.float32(x), np.float32(y))
        ox, oy, _ = _var("swirl", x, y)
        if r < 100:
            assert _near(math.hypot(ox, oy), r, 1e-5)  # swirl is a rotation
        ox, oy, st = _var("julia", x, y, rng_state=(12345, 7))
        assert _near(math.hypot(ox, oy), math.sqrt(r), 1e-5)  # |julia(p)| = sqrt|p|
        ox, oy, st = _var("julian", x, y, params=(3.0, 1.0, 1.0 / 3.0 / 2.0, 0))
        assert _near(math.hypot(ox, oy), (r * r) ** (1.0 / 6.0), 1e-5)
        # waves with b = d = 0 is the identity
        ox, oy, _ = _var("waves", x, y)
        assert (ox, oy) == (np.float32(x), np.float32(y))


def test_julia_and_julian_consume_one_draw():
    for name, params in (("julia", (0, 0, 0, 0)), ("julian", (2.0, 1.0, 0.25, 0))):
        _, _, r = _var(name, 0.3, 0.4, params=params, rng_state=(99, 5))
        assert (r.state, r.inc) == ((99 * 747796405 + 5) & 0xFFFFFFFF, 5)
        _, _, r = _var("spherical", 0.3, 0.4, rng_state=(99, 5))
        assert (r.state, r.inc) == (99, 5)


def _draws(state, inc, n):
    """The next n rng_nextf values of a stream (the pinned PCG step)."""
    L = O.lib()
    r = O.Rng()
    r.state, r.inc = state, inc
    return [L.orc_rng_nextf(C.byref(r)) for _ in range(n)], r


# ---- independent FP64 pins of the bodies that draw or read the transform
# (advanced.rs:329-340 julia, :414-422 waves, :467-506 julian with
# init_julian) and of the colour-index update (main_template.wgsl:496-503)

SMALL = [(x, y) for x, y in PROBE[2:] if 1e-6 < math.hypot(x, y) < 10.0]


def test_julia_closed_form_with_its_branch():
    # r = |p|, theta = atan2(y, x), omega = pi when the draw < 0.5,
    # out = sqrt(r) (cos, sin)(theta / 2 + omega); the draw is the pinned
    # stream's next value. f32 rounds theta/2 + omega at magnitude up to
    # 4.7, so the bar is that half-ulp times sqrt(r) plus a few ulp.
    branches = set()
    for x, y in SMALL:
        X, Y = float(np.float32(x)), float(np.float32(y))
        r = math.hypot(X, Y)
        for state in range(1, 40, 3):
            u, _ = _draws(state, 7, 1)
            om = math.pi if u[0] < 0.5 else 0.0
            branches.add(om)
            t = math.atan2(Y, X) / 2 + om
            ex, ey = math.sqrt(r) * math.cos(t), math.sqrt(r) * math.sin(t)
            ox, oy, st = _var("julia", x, y, rng_state=(state, 7))
            tol = 4e-7 * math.sqrt(r) + 4e-7 * math.sqrt(r)
            assert abs(ox - ex) <= tol and abs(oy - ey) <= tol, (x, y, state, ox, oy, ex, ey)
    assert branches == {0.0, math.pi}
    # the swapped operand order atan2(x, y) is another map: it fails the form
    X, Y = 0.3, 0.8
    ox, oy, _ = _var("julia", X, Y, rng_state=(1, 7))
    u, _ = _draws(1, 7, 1)
    t = math.atan2(X, Y) / 2 + (math.pi if u[0] < 0.5 else 0.0)
    assert abs(ox - math.sqrt(math.hypot(X, Y)) * math.cos(t)) > 1e-2


def _julian_flame(power, dist):
    return pyfae.flame_desc([pyfae.xform(variations={"julian": 1.0},
                                         params={"julian.power": power, "julian.dist": dist})])


@pytest.mark.parametrize("power,dist", [(2, 1.0), (3, 0.5), (-3, 1.0), (5, -1.5), (-7, 2.0)])
def test_julian_init_and_closed_form(power, dist):
    # init_julian: cpower = dist / max(|power|, 1e-30) / 2 (advanced.rs:467-474)
    f2 = O.prepared(_julian_flame(power, dist))
    cp = f2.xforms[0].var_param[0][2]
    assert cp == float(np.float32(np.float32(dist) / np.float32(abs(power)) / np.float32(2)))
    assert abs(cp - dist / abs(power) / 2) <= 1e-7 * abs(dist)
    # body: r = (x^2 + y^2)^cpower, theta = atan2(y, x), k = floor(|power| u),
    # t = (theta + 2 pi k) / power, out = r (cos t, sin t)
    ks = set()
    for x, y in SMALL:
        X, Y = float(np.float32(x)), float(np.float32(y))
        r = (X * X + Y * Y) ** (dist / abs(power) / 2)
        for state in range(1, 160, 4):
            u, _ = _draws(state, 11, 1)
            k = math.floor(abs(power) * u[0])
            ks.add(k)
            t = (math.atan2(Y, X) + 2 * math.pi * k) / power
            ox, oy, st = _var("julian", x, y, params=(power, dist, cp, 0), rng_state=(state, 11))
            tol = 2e-6 * r + 1e-6 * r * abs(power)
            assert abs(ox - r * math.cos(t)) <= tol and abs(oy - r * math.sin(t)) <= tol, \
                (power, dist, x, y, state, ox, oy, r * math.cos(t), r * math.sin(t))
            assert (st.state, st.inc) == (_draws(state, 11, 1)[1].state, 11)
    assert ks == set(range(abs(power)))  # every branch angle is taken


def test_waves_reads_its_own_b_e_d_f():
    # x + b sin(y / (e^2 + 1e-6)), y + d sin(x / (f^2 + 1e-6)) with the
    # transform's own b, d, e, f (advanced.rs:414-422); distinct values, so
    # an e/f or b/d swap is caught
    xf = pyfae.Xform()
    xf.b, xf.d, xf.e, xf.f = np.float32(0.37), np.float32(-0.81), np.float32(0.7), np.float32(1.3)
    b, d, e, f = (float(v) for v in (xf.b, xf.d, xf.e, xf.f))
    for x, y in SMALL:
        X, Y = float(np.float32(x)), float(np.float32(y))
        ex = X + b * math.sin(Y / (e * e + 1e-6))
        ey = Y + d * math.sin(X / (f * f + 1e-6))
        ox, oy, st = _var("waves", x, y, xf=xf, rng_state=(5, 3))
        assert _near(ox, ex, 4e-6, 2e-6) and _near(oy, ey, 4e-6, 2e-6), (x, y, ox, oy, ex, ey)
        assert (st.state, st.inc) == (5, 3)
        sx = X + b * math.sin(Y / (f * f + 1e-6))  # e and f swapped
        if abs(sx - ex) > 1e-3:
            assert abs(ox - sx) > 1e-4


@pytest.mark.parametrize("color,speed", [(1.0, 0.5), (0.6, -0.3), (0.25, 0.9)])
def test_colour_index_update_is_the_symmetry_form(color, speed):
    # c <- c (1 + s)/2 + colour (1 - s)/2 (main_template.wgsl:496-503):
    # from c0 = 0 with one transform, c_n = colour (1 - ((1 + s)/2)^n).
    # flam3's c <- c (1 - s) + colour s gives colour (1 - (1 - s)^n)
    # instead, which differs for s != 1/2... and at s = 1/2 (0.75^n vs 0.5^n)
    fd = pyfae.flame_desc([pyfae.xform(0.5, 0.0, 0.0, 0.5, color=color, color_speed=speed)])
    view = _abi.View(64, 64, 1.0, 0.0, 0.0, 0.0)
    tr = O.trace(fd, np.array([[0, 0, 0, 0], [1, 1, 1, 1]], np.float32), view, 3, 1, 0, 30)
    ci = tr[:, 4].view(np.float32).astype(np.float64)
    n = np.arange(1, 31)
    want = color * (1 - ((1 + speed) / 2) ** n)
    flam3 = color * (1 - (1 - speed) ** n)
    assert np.abs(ci - want).max() <= 1e-6
    assert np.abs(ci - flam3).max() > 1e-2


def test_blob_against_float64_closed_form():
    high, low, waves = 1.3, 0.4, 5.0
    for x, y in PROBE[2:]:
        X, Y = float(np.float32(x)), float(np.float32(y))
        if X * X + Y * Y > 1e4:
            continue  # sin(waves * theta) is fine, but scale * cos loses the float32 digits
        r = math.hypot(X, Y)
        th = math.atan2(X, Y)  # Apophysis's atan2(x, y)
        sc = r * (low + (high - low) / 2 * (math.sin(waves * th) + 1))
        ox, oy, st = _var("blob", x, y, params=(high, low, waves, 0), rng_state=(5, 3))
        assert _near(ox, sc * math.cos(th), 1e-5, 1e-6) and _near(oy, sc * math.sin(th), 1e-5, 1e-6)
        assert (st.state, st.inc) == (5, 3)  # no draw


def test_blur_noise_gaussian_blur_draws_and_closed_forms():
    tp = float(np.float32(6.28318530718))
    for x, y in PROBE[2:8]:
        X, Y = float(np.float32(x)), float(np.float32(y))
        u, end2 = _draws(77, 9, 2)
        th = float(np.float32(u[0]) * np.float32(tp))
        ox, oy, st = _var("blur", x, y, rng_state=(77, 9))
        assert _near(ox, u[1] * math.cos(th), 1e-5, 1e-7) and _near(oy, u[1] * math.sin(th), 1e-5, 1e-7)
        assert (st.state, st.inc) == (end2.state, end2.inc)
        ox, oy, st = _var("noise", x, y, rng_state=(77, 9))
        assert _near(ox, X * u[1] * math.cos(th), 1e-5, 1e-7)
        assert _near(oy, Y * u[1] * math.sin(th), 1e-5, 1e-7)
        assert (st.state, st.inc) == (end2.state, end2.inc)
        u, end5 = _draws(77, 9, 5)
        g = float(np.float32(np.float32(np.float32(np.float32(u[1]) + np.float32(u[2])) + np.float32(u[3]))
                             + np.float32(u[4])) - np.float32(2.0))
        ox, oy, st = _var("gaussian_blur", x, y, rng_state=(77, 9))
        assert _near(ox, g * math.cos(th), 1e-5, 1e-7) and _near(oy, g * math.sin(th), 1e-5, 1e-7)
        assert (st.state, st.inc) == (end5.state, end5.inc)


def test_pre_blur_runs_before_the_sum_whatever_its_slot():
    """pre_blur is pre-phase: it moves the affine point by w * (six draws
    - 3) at angle 2 pi * a seventh draw, and the normal variations read the
    moved point, even when pre_blur's local index comes after theirs."""
    L = O.lib()
    xs = [pyfae.xform(0.8, 0.1, -0.2, 0.9, 0.05, -0.1, variations={"spherical": 0.7,
                                                                   "pre_blur": 0.35})]
    fd = pyfae.flame_desc(xs)
    assert [fd.var_id[k] for k in range(fd.n_vars)] == [2, 54]
    for x, y in ((0.3, 0.4), (-0.7, 0.2), (1.5, -1.1)):
        X, Y = np.float32(x), np.float32(y)
        ax = float(np.float32(np.float32(0.8) * X + np.float32(0.1) * Y) + np.float32(0.05))
        ay = float(np.float32(np.float32(-0.2) * X + np.float32(0.9) * Y) + np.float32(-0.1))
        u, end = _draws(1234, 11, 7)
        g = np.float32(0)
        for v in u[:6]:
            g = np.float32(g + np.float32(v))
        rg = float(np.float32(0.35) * np.float32(g - np.float32(3.0)))
        ang = float(np.float32(u[6]) * np.float32(6.28318530717959))
        tx, ty = ax + rg * math.cos(ang), ay + rg * math.sin(ang)
        r2 = tx * tx + ty * ty + 1e-6
        px, py = C.c_float(x), C.c_float(y)
        r = O.Rng()
        r.state, r.inc = 1234, 11
        L.orc_apply_xform(C.byref(fd), C.byref(fd.xforms[0]), C.byref(px), C.byref(py), C.byref(r))
        assert _near(px.value, 0.7 * tx / r2, 1e-5, 1e-6) and _near(py.value, 0.7 * ty / r2, 1e-5, 1e-6)
        assert (r.state, r.inc) == (end.state, end.inc)  # seven draws


def test_curl_is_the_moebius_like_closed_form():
    c1, c2 = 0.7, 0.2
    for x, y in PROBE[2:8]:
        X, Y = float(np.float32(x)), float(np.float32(y))
        z = complex(X, Y)
        w = z / (1 + c1 * z + c2 * z * z)  # curl: p / (1 + c1 p + c2 p^2)
        ox, oy, _ = _var("curl", x, y, params=(c1, c2, 0, 0))
        assert _near(ox, w.real, 1e-5, 1e-6) and _near(oy, w.imag, 1e-5, 1e-6)


# ------------------------------------------------------ whole chaos game

def test_plain_trace_is_self_consistent():
    fd, pal, job = configs.plain(walkers=64, ipt=50)
    tr = O.trace(fd, pal, job.view, 3, 1, 20, 50)
    # burn-in: 20 steps flagged burning, then plotted or off-frame
    assert all(t[7] == 4 for t in tr[:20])
    assert all(t[7] in (0, 1) for t in tr[20:])
    # a linear flame: each step is exactly x' = a x + b y + e of the chosen xform
    p = O.prepared(fd)
    x0, y0 = f32(tr[19][2]), f32(tr[19][3])
    for t in tr[20:30]:
        xf = p.xforms[int(t[1])]
        ax = np.float32(np.float32(xf.a) * np.float32(x0)) + np.float32(np.float32(xf.b) * np.float32(y0))
        ax = np.float32(ax + np.float32(xf.e))
        ay = np.float32(np.float32(xf.c) * np.float32(x0)) + np.float32(np.float32(xf.d) * np.float32(y0))
        ay = np.float32(ay + np.float32(xf.f))
        ax, ay = np.float32(np.float32(1.0) * ax), np.float32(np.float32(1.0) * ay)
        assert (bits(float(ax)), bits(float(ay))) == (t[2], t[3])
        x0, y0 = f32(t[2]), f32(t[3])


def test_histogram_counts_and_colour_bounds():
    fd, pal, job = configs.plain(walkers=256, ipt=200)
    h, plotted, inframe = O.iterate(fd, pal, job.view, 256, 200, 20, 1)
    assert plotted == 256 * 180
    assert int(h[..., 3].sum()) == 100 * inframe
    assert (h[..., :3] <= h[..., 3:4]).all()


def test_mt_iterate_equals_single_thread():
    fd, pal, job = configs.single4k(walkers=1 << 10, ipt=64, width=320, height=180)
    h1, p1, _ = O.iterate(fd, pal, job.view, 1 << 10, 64, 20, 1)
    h2, p2 = O.iterate_mt(fd, pal, job.view, 1 << 10, 64, 20, 1, 4)
    assert p1 == p2 and np.array_equal(h1, h2)


def _simple_linear_job():
    d = json.load(open(os.path.join(GOLDEN, "ref_simple_linear.fflame")))
    t = d["flame"]["transforms"][0]
    fd = pyfae.flame_desc([pyfae.xform(t["a"], t["b"], t["c"], t["d"], t["e"], t["f"],
                                       weight=t["weight"], color=t["color"],
                                       color_speed=t["color_speed"], opacity=t["opacity"],
                                       variations=t["variations"])])
    stops = np.array([[s["position"]] + s["color"] for s in d["palette"]["stops"]], np.float32)
    tp = pyfae.tonemap_defaults()
    tp.gamma, tp.gamma_threshold = d["gamma"], d["gamma_threshold"]
    tp.brightness, tp.vibrancy, tp.exposure = d["brightness"], d["vibrancy"], d["exposure"]
    tp.use_curve = int(d.get("use_curve", True))
    # the CLI export the baseline came from: 800x600, ipt 256, burn-in 20,
    # 128 x 64 threads, deterministic seeds (tests/visual/run_tests.py:273-284)
    job = pyfae.make_job(800, 600, walkers=8192, ipt=256, burn_in=20,
                         target_samples=d["max_iterations"], zoom=d["zoom"], tonemap=tp)
    return fd, stops, job


def visual_rule(rgba, baseline_png):
    """The reference's visual comparison (tests/visual/run_tests.py:50-75)."""
    from PIL import Image
    a = np.asarray(Image.fromarray(np.ascontiguousarray(rgba[..., :3])).resize(
        (160, 120), Image.LANCZOS)).astype(np.float64)
    b = np.asarray(Image.open(baseline_png).convert("RGB")).astype(np.float64)
    d = np.abs(a - b)
    return d.mean(), (d.max(-1) > 40).mean()


def test_oracle_render_matches_reference_baseline_image():
    fd, stops, job = _simple_linear_job()
    rgba, acc, total = O.render(fd, stops, job, threads=8)
    assert total >= job.target_samples
    mean, frac = visual_rule(rgba, os.path.join(GOLDEN, "ref_2d_simple_linear.png"))
    assert mean <= 2.0 and frac <= 0.01, (mean, frac)


# -------------------------------------------------------------- stencils

def test_blur_constant_field_and_gaussian_taps():
    h = np.full((9, 11, 4), 500, np.uint32)
    out = O.blur(h, 1.3, 1e9)
    assert (np.abs(out.astype(np.int64) - 500) <= 1).all()
    # a single spike with a huge density sigma spreads as a separable
    # Gaussian of sigma 1 px (histogram_blur.wgsl:53-56), truncated at 3 sigma
    h = np.zeros((15, 15, 4), np.uint32)
    h[7, 7] = [1_000_000] * 4
    out = O.blur(h, 1.0, 1e12).astype(np.float64)
    k = np.exp(-np.arange(-3, 4) ** 2 / 2.0)
    k /= k.sum()
    want = 1e6 * np.outer(k, k)
    np.testing.assert_allclose(out[4:11, 4:11, 3], want, rtol=1e-4, atol=2)


def test_density_estimate_conserves_mass_and_keeps_flat_fields():
    rng = np.random.default_rng(0)
    acc = np.zeros((40, 40, 4), np.float32)
    acc[12:28, 12:28, 3] = rng.integers(1, 50, (16, 16))
    acc[12:28, 12:28, :3] = rng.uniform(0, 1, (16, 16, 3))
    out = O.density_estimate(acc, 9.0, 0.0, 0.4)
    assert abs(out[..., 3].sum() / acc[..., 3].sum() - 1) < 1e-5
    colour_mass = (acc[..., :3] * acc[..., 3:]).sum((0, 1))
    np.testing.assert_allclose((out[..., :3] * out[..., 3:]).sum((0, 1)), colour_mass, rtol=1e-5)
    flat = np.zeros((60, 60, 4), np.float32)
    flat[..., 3] = 7.0
    flat[..., :3] = 0.25
    o2 = O.density_estimate(flat, 9.0, 0.0, 0.4)
    np.testing.assert_allclose(o2[15:45, 15:45, 3], 7.0, rtol=1e-5)
    np.testing.assert_allclose(o2[15:45, 15:45, :3], 0.25, rtol=1e-5)


def test_density_estimate_dense_pixels_stay_put():
    # a count past the last filter uses the narrowest disc (radius min_r + 1
    # = 1): its mass spreads over the 3x3 cross at most
    acc = np.zeros((21, 21, 4), np.float32)
    acc[10, 10] = [0.5, 0.5, 0.5, 1e6]
    out = O.density_estimate(acc, 9.0, 0.0, 0.4)
    assert out[8:13, 8:13, 3].sum() / 1e6 > 0.999999
    # a single-hit pixel uses the widest disc (radius 10)
    acc[10, 10, 3] = 1.0
    out = O.density_estimate(acc, 9.0, 0.0, 0.4)
    assert (out[..., 3] > 0).sum() > 250


# ------------------------------------------------------ final attachments

def _attach_flame(attach):
    """3 linear transforms and 2 linear finals; `attach` per transform, or
    None for the legacy chain (every final, in order)."""
    rng = configs.SplitMix64(77)
    xs = []
    for i in range(3):
        a = configs._affine(rng, 0.7)
        xs.append(pyfae.xform(*a, weight=1.0, color=i / 2.0, variations={"linear": 1.0},
                              final_attachments=None if attach is None else attach[i]))
    fins = [pyfae.xform(0.9, 0.1, -0.1, 0.9, 0.05, 0.0, variations={"linear": 1.0}),
            pyfae.xform(0.8, 0.0, 0.0, -0.8, 0.0, 0.1, variations={"linear": 1.0})]
    return pyfae.flame_desc(xs, fins)


def test_final_attachments_all_equal_the_legacy_chain():
    """Every transform attaching [0, 1] is the single shared chain
    (migrate_legacy_final attaches the legacy final to every transform,
    src/scene/transforms.rs:2257-2266)."""
    fd0 = _attach_flame(None)
    fd1 = _attach_flame([[0, 1]] * 3)
    assert fd0.per_transform_finals == 0 and fd1.per_transform_finals == 1
    pal = configs.random_palette(configs.SplitMix64(5))
    view = pyfae.View(256, 256, 1.0, 0.0, 0.0, 0.0)
    for w in (0, 9, 1234):
        assert np.array_equal(O.trace(fd0, pal, view, w, 1, 5, 300), O.trace(fd1, pal, view, w, 1, 5, 300))


def test_final_attachments_per_transform():
    """A transform attaching no final plots its own output; one attaching
    [1, 0] applies final 1 then final 0 (AttachmentList order,
    shaders/core/main_template.wgsl:663-717)."""
    fd = _attach_flame([[], [1, 0], [0]])
    pal = configs.random_palette(configs.SplitMix64(5))
    view = pyfae.View(256, 256, 1.0, 0.0, 0.0, 0.0)
    tr = O.trace(fd, pal, view, 3, 1, 5, 400)
    f = O.prepared(fd)
    seen = set()
    for row in tr[5:]:  # past the burn-in
        xi = int(row[1])
        x, y = np.frombuffer(row[2:4].tobytes(), np.float32)
        fx, fy = C.c_float(x), C.c_float(y)
        for k in {0: [], 1: [1, 0], 2: [0]}[xi]:
            O.lib().orc_apply_xform(C.byref(f), C.byref(f.finals[k]), C.byref(fx), C.byref(fy), None)
        px, py = C.c_int32(), C.c_int32()
        O.lib().orc_world_to_pixel(C.byref(view), fx, fy, C.byref(px), C.byref(py))
        assert (int(np.int32(row[5])), int(np.int32(row[6]))) == (px.value, py.value), xi
        seen.add(xi)
    assert seen == {0, 1, 2}


# ------------------------------------------------ exact density count switch

def test_render_exact_count_switch():
    """From accumulate pass `exact_after` on, densities are counted in 64
    bits (src/renderer/compute_kernel.rs:360-372,1562-1575): alpha is then
    f32(count) / 100 of an integer count, seeded from the running sum when
    the switch comes mid-render; before it, the f32 running sum."""
    fd, pal, job = configs.plain(walkers=2048, ipt=100)
    job.view.width, job.view.height = 96, 64
    job.target_samples = 2048 * 80 * 4 * 6  # six accumulate passes
    _, a_sum, t0 = O.render(fd, pal, job)
    for after in (1, 3):
        _, a_ex, t1 = O.render(fd, pal, job, exact_after=after)
        assert t1 == t0
        alpha = a_ex[..., 3]
        cnt = np.round(alpha.astype(np.float64) * 100.0)
        np.testing.assert_array_equal(np.float32(cnt) / np.float32(100), alpha)
        np.testing.assert_allclose(alpha, a_sum[..., 3], rtol=2e-6, atol=0)
        np.testing.assert_array_equal(a_ex[..., :3] == 0, a_sum[..., :3] == 0)


def test_tonemap_transparent_modes():
    """Transparent export (shaders/tonemap.wgsl:531-576). Premultiplied: rgb
    is the over-black picture, so it equals the opaque tonemap over a black
    background bit for bit; straight alpha: the coverage alpha is the largest
    channel of that picture, and rgb x a over black gives it back."""
    rng = np.random.default_rng(7)
    H, W = 40, 50
    acc = np.zeros((H, W, 4), np.float32)
    acc[..., :3] = rng.uniform(0, 1, (H, W, 3))
    acc[..., 3] = np.exp(rng.uniform(-4, 9, (H, W))).astype(np.float32)
    acc[::5, ::4] = 0
    out = {}
    for mode in (0, 1, 2):
        tp = configs.log_tonemap(2.2)
        tp.transparent_mode = mode
        out[mode] = O.tonemap(acc, 1.0, 2e5, tp).astype(int)
    opaque, straight, premul = out[0], out[1], out[2]
    assert (opaque[..., 3] == 255).all()
    np.testing.assert_array_equal(premul[..., :3], opaque[..., :3])
    np.testing.assert_array_equal(straight[..., 3], premul[..., :3].max(-1))
    flat = np.round(straight[..., :3] * straight[..., 3:4] / 255.0)
    assert np.abs(flat - opaque[..., :3]).max() <= 1
    empty = acc[..., 3] == 0
    assert (straight[empty] == 0).all() and (premul[empty] == 0).all()
