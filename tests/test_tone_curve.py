"""The tone curve (ToneCurve, src/scene/tonemap.rs:142-303; applied at
shaders/tonemap.wgsl:479-492) in the oracle and the scene loader, on the CPU.

The oracle's curve is pinned by the reference's own tests of it
(tonemap.rs:400-423), by what a cubic Hermite with Fritsch-Carlson limits
must do (pass through every point, reproduce a line, stay monotone and
inside the data where the data is monotone, stay flat on a flat segment),
and its f16 rounding by numpy's float16."""
import json

import numpy as np
import pytest

from oracle import oracle as O
import pyfae
from pyfae import _abi, formats

LINEAR = [(0.0, 0.0), (1.0, 1.0)]
S_CURVE = [(0.0, 0.0), (0.25, 0.15), (0.5, 0.5), (0.75, 0.85), (1.0, 1.0)]
SHADOWS = [(0.0, 0.0), (0.25, 0.4), (0.5, 0.6), (0.75, 0.8), (1.0, 1.0)]


def f32(v):
    return float(np.float32(v))


def curve_tp(points):
    tp = _abi.tonemap_defaults()
    tp.use_curve = 1
    tp.curve_n = len(points)
    for i, (x, y) in enumerate(points):
        tp.curve_x[i], tp.curve_y[i] = x, y
    return tp


def test_round_f16_matches_numpy():
    rng = np.random.default_rng(5)
    v = np.concatenate([
        rng.uniform(0, 1, 4000), rng.uniform(-1, 1, 1000),
        np.exp(rng.uniform(np.log(1e-9), np.log(7e4), 4000)) * rng.choice([-1, 1], 4000),
        # ties between f16 neighbours, both parities, normal and subnormal
        (np.arange(1000, 1200) + 0.5) / 1024.0, (np.arange(1, 200) + 0.5) * 2.0 ** -24,
        [0.0, -0.0, 2.0 ** -25, 2.0 ** -24, 65504.0, 65519.0, 65520.0, 1e30, np.inf, -np.inf],
    ]).astype(np.float32)
    with np.errstate(over="ignore"):
        want = v.astype(np.float16).astype(np.float32)
    got = np.array([O.round_f16(x) for x in v], dtype=np.float32)
    assert got.tobytes() == want.tobytes()


def test_reference_curve_tests():
    # tonemap.rs:400-423: linear, s-curve and add_point's (0.5, 0.7)
    assert O.curve_eval(LINEAR, 0.0) == 0.0
    assert O.curve_eval(LINEAR, 0.5) == 0.5
    assert O.curve_eval(LINEAR, 1.0) == 1.0
    assert O.curve_eval(S_CURVE, 0.25) < 0.25 or O.curve_eval(S_CURVE, 0.2) < 0.2
    assert O.curve_eval(S_CURVE, 0.5) == 0.5
    assert O.curve_eval(S_CURVE, 0.8) > 0.8
    assert O.curve_eval([(0.0, 0.0), (0.5, 0.7), (1.0, 1.0)], 0.5) == f32(0.7)


@pytest.mark.parametrize("pts", [S_CURVE, SHADOWS, [(0.0, 0.1), (0.1, 0.9), (1.0, 1.0)]])
def test_curve_passes_through_its_points(pts):
    for x, y in pts:
        assert O.curve_eval(pts, x) == f32(y)


def test_collinear_points_give_the_line():
    pts = [(0.0, 0.2), (0.3, 0.35), (0.6, 0.5), (1.0, 0.7)]
    for x in np.linspace(0, 1, 101):
        assert abs(O.curve_eval(pts, x) - (0.2 + 0.5 * x)) < 1e-6


def test_clamps_outside_unit_interval():
    assert O.curve_eval(SHADOWS, -0.5) == 0.0 and O.curve_eval(SHADOWS, 7.0) == 1.0


def test_flat_segment_is_flat():
    pts = [(0.0, 0.0), (0.3, 0.5), (0.6, 0.5), (1.0, 1.0)]
    for x in np.linspace(0.3, 0.6, 31):
        assert O.curve_eval(pts, x) == 0.5


def test_tangent_limit_prevents_overshoot():
    # segment 1 rises 0.1 over 0.9 but its left tangent, the mean of the
    # neighbouring secants, is 4.6: unlimited, the cubic overshoots past 1
    pts = [(0.0, 0.0), (0.1, 0.9), (1.0, 1.0)]
    ys = [O.curve_eval(pts, x) for x in np.linspace(0.1, 1.0, 901)]
    assert max(ys) <= 1.0 and all(b >= a for a, b in zip(ys, ys[1:]))
    # the limited tangent: tau * alpha * delta with alpha^2 + beta^2 = 9
    delta = 0.1 / 0.9
    alpha, beta = (9.0 + delta) / 2 / delta, 1.0
    tau = 3 / np.hypot(alpha, beta)
    h = 1e-3
    slope = (O.curve_eval(pts, 0.1 + h) - O.curve_eval(pts, 0.1)) / h
    assert abs(slope - tau * alpha * delta) < 0.02


def test_opposing_tangent_falls_back_to_the_secant():
    # a peak: at (0.5, 1) the mean slope is 0 -> alpha 0 (kept); at the
    # next segment's left end m0 = 0, beta from the far secant: a segment
    # whose neighbour tangent has the opposite sign is a straight line
    pts = [(0.0, 0.0), (0.4, 0.8), (0.5, 0.2), (1.0, 1.0)]
    # segment [0.4, 0.5]: delta -6, m0 = (2 - 6)/2 < 0 -> alpha > 0; m1 =
    # (-6 + 1.6)/2 < 0 -> beta > 0, so it is Hermite; segment [0.5, 1]:
    # m0 = (-6 + 1.6)/2 < 0 against delta 1.6 > 0 -> the secant
    for x in np.linspace(0.5, 1.0, 51):
        assert abs(O.curve_eval(pts, x) - (0.2 + 1.6 * (x - 0.5))) < 1e-6


def test_monotone_data_gives_monotone_curve():
    rng = np.random.default_rng(11)
    for _ in range(20):
        n = rng.integers(3, 17)
        xs = np.sort(np.concatenate([[0.0, 1.0], rng.uniform(0.01, 0.99, n - 2)]))
        ys = np.sort(rng.uniform(0, 1, n))
        pts = list(zip(xs.tolist(), ys.tolist()))
        vals = [O.curve_eval(pts, x) for x in np.linspace(0, 1, 1021)]
        assert all(b >= a - 1e-6 for a, b in zip(vals, vals[1:]))
        assert min(vals) >= f32(ys[0]) - 1e-6 and max(vals) <= f32(ys[-1]) + 1e-6


def test_lut_of_the_linear_curve():
    tp = _abi.tonemap_defaults()
    assert tp.use_curve == 1 and tp.curve_n == 0
    lut = O.curve_lut(tp)
    want = (np.arange(256, dtype=np.float32) / np.float32(255)).astype(np.float16).astype(np.float32)
    assert lut.tobytes() == want.tobytes()
    assert O.curve_lut(curve_tp(LINEAR)).tobytes() == lut.tobytes()
    s = O.curve_lut(curve_tp(S_CURVE))
    assert s[0] == 0 and s[255] == 1 and np.all(np.diff(s) >= 0)


def accum(seed=2, H=48, W=64):
    rng = np.random.default_rng(seed)
    acc = np.zeros((H, W, 4), np.float32)
    acc[..., :3] = rng.uniform(0, 1, (H, W, 3))
    acc[..., 3] = np.exp(rng.uniform(-3, 12, (H, W))).astype(np.float32)
    acc[::7, ::3] = 0
    return acc


def test_curve_table_is_read_with_linear_filtering():
    # the curve sampler filters linearly (src/gpu/buffers.rs:1877-1886) and
    # the shader reads it at u = (255c + 0.5)/256 (tonemap.wgsl:479-488):
    # texel-space position 255c, so between texel centres the read is the
    # straight line through the two texels, and it clamps at both ends. A
    # nearest read would miss the midpoints by half a texel step.
    rng = np.random.default_rng(11)
    lut = np.sort(rng.uniform(0, 1, 256)).astype(np.float32)
    cs = np.concatenate([rng.uniform(0, 1, 2000), (np.arange(255) + 0.5) / 255.0,
                         [0.0, 1.0, -0.5, 1.5, np.nan]]).astype(np.float32)
    for c in cs:
        s = np.float64(np.float32(c * np.float32(255)))
        want = np.interp(s, np.arange(256.0), lut.astype(np.float64)) if s == s else lut[0]
        assert abs(O.curve_sample(lut, c) - want) <= 2e-7, (c, O.curve_sample(lut, c), want)


def ramp_accum(H=96, W=128):
    # a black-to-white ramp of mean colours at a count well past the
    # curve's 0.001 gate
    acc = np.zeros((H, W, 4), np.float32)
    g = np.linspace(0.0, 1.0, H * W, dtype=np.float32).reshape(H, W)
    acc[..., 0] = g ** np.float32(2.2)  # grey after the 1/2.2 encode
    acc[..., 1] = g ** 4
    acc[..., 2] = g
    acc[..., 3] = 50.0
    return acc


def test_the_identity_tone_curve_changes_nothing():
    # src/escape/app_repro_test.rs:10455-10507: through a black-to-white
    # ramp, linear tone map at gamma 1, no levels, black background, the
    # identity curve on is the curve off to a level, and the ramp's dark end
    # is in the frame
    acc = ramp_accum()
    on = _abi.tonemap_defaults()
    on.tonemap_mode = 0
    on.gamma = 1.0
    on.exposure = 1.0
    on.levels_enabled = 0
    on.background[:] = [0.0, 0.0, 0.0]
    on.use_curve = 1
    on.curve_n = 0
    off = _abi.tonemap_defaults()
    ctypes_copy(off, on)
    off.use_curve = 0
    a = O.tonemap(acc, 4.0 / 3.0, 5e6, on).astype(int)
    b = O.tonemap(acc, 4.0 / 3.0, 5e6, off).astype(int)
    assert (b[..., 0] < 40).sum() > 500
    assert np.abs(a - b)[..., :3].max() <= 1
    # the log tone map with the default (identity) curve, too: a level at most
    acc2 = accum()
    on2 = _abi.tonemap_defaults()
    off2 = _abi.tonemap_defaults()
    off2.use_curve = 0
    d = np.abs(O.tonemap(acc2, 4.0 / 3.0, 5e6, on2).astype(int)
               - O.tonemap(acc2, 4.0 / 3.0, 5e6, off2).astype(int))
    assert d.max() <= 1 and d[..., 3].max() == 0


def ctypes_copy(dst, src):
    import ctypes
    ctypes.memmove(ctypes.addressof(dst), ctypes.addressof(src), ctypes.sizeof(src))


def test_identity_curve_nearest_read_would_fail():
    # the rule above distinguishes the filters: reading the identity table
    # at the nearest texel moves dark channels by many levels
    acc = ramp_accum()
    lut = O.curve_lut(_abi.tonemap_defaults())
    c = acc[..., 0].ravel()
    nearest = lut[np.clip(np.floor(c * np.float32(255) + np.float32(0.5)), 0, 255).astype(int)]
    enc = lambda v: np.floor(np.clip(v, 0, 1) ** (1 / 2.2) * 255 + 0.5)
    assert np.abs(enc(nearest) - enc(c)).max() > 5
    linear = np.array([O.curve_sample(lut, v) for v in c[::37]])
    assert np.abs(enc(linear) - enc(c[::37])).max() <= 1


def test_s_curve_darkens_and_brightens():
    acc = accum()
    off = _abi.tonemap_defaults()
    off.use_curve = 0
    off.levels_enabled = 0
    on = curve_tp(S_CURVE)
    on.levels_enabled = 0
    a = O.tonemap(acc, 4.0 / 3.0, 5e6, on)[..., :3].astype(int)
    b = O.tonemap(acc, 4.0 / 3.0, 5e6, off)[..., :3].astype(int)
    # the curve's fixed point 0.5 is 186 after the 1/2.2 encode: darker
    # below, brighter above (the darkest channels, within the table's first
    # few 1/255 steps, can gain a level or lose many)
    mid = (b >= 40) & (b < 180)
    assert a[mid].mean() < b[mid].mean() - 5 and (a[b < 40] <= b[b < 40] + 1).all()
    assert a[b > 215].mean() > b[b > 215].mean() + 1 and (a[b > 215] >= b[b > 215]).all()
    assert (a != b).mean() > 0.3


def minimal(**extra):
    cfg = {"flame": {"name": "t", "transforms": [
        {"a": 0.5, "b": 0.0, "c": 0.0, "d": 0.5, "e": 0.0, "f": 0.0, "weight": 1.0,
         "variations": {"linear": 1.0}, "color": 0.25}]}}
    cfg.update(extra)
    return cfg


def test_scene_loads_the_curve_and_round_trips():
    pts = [{"x": 0.0, "y": 0.0}, {"x": 0.25, "y": 0.4}, {"x": 0.5, "y": 0.6}, {"x": 1.0, "y": 1.0}]
    sc = pyfae.load(json.dumps(minimal(tonemap_curve={"points": pts})))
    tp = sc.tonemap
    assert tp.use_curve == 1 and tp.curve_n == 4
    assert [tp.curve_x[i] for i in range(4)] == [f32(p["x"]) for p in pts]
    assert [tp.curve_y[i] for i in range(4)] == [f32(p["y"]) for p in pts]
    back = pyfae.load(json.dumps(formats.scene_to_config(sc)))
    assert bytes(back.tonemap) == bytes(tp)
    off = pyfae.load(json.dumps(minimal(use_curve=False)))
    assert off.tonemap.use_curve == 0 and off.tonemap.curve_n == 0


def test_scene_curve_limits():
    many = [{"x": i / 16, "y": i / 16} for i in range(17)]
    with pytest.raises(pyfae.EngineError) as e:
        pyfae.load(json.dumps(minimal(tonemap_curve={"points": many})))
    assert e.value.status == _abi.ERR_TOO_LARGE
    with pytest.raises(pyfae.EngineError) as e:
        pyfae.load(json.dumps(minimal(tonemap_curve={"points": [{"x": 0, "y": 0}]})))
    assert e.value.status == _abi.ERR_PARSE


def test_apophysis_curves_sample_the_bezier():
    # parse_apophysis_curves (src/flame_xml.rs:1084-1117): (0,0), the first
    # rational Bezier at t = 1/4, 1/2, 3/4, (1,1), sorted by x
    cv = "0 0 1 0.25 0.5 1 0.75 0.9 1 1 1 1 " + "0 0 1 0.25 0.25 1 0.75 0.75 1 1 1 1 " * 3
    xml = ('<flame curves="%s"><xform weight="1" color="0" coefs="1 0 0 1 0 0" linear="1"/>'
           '</flame>') % cv
    tp = pyfae.load(xml).tonemap
    assert tp.use_curve == 1 and tp.curve_n == 5
    ctrl = np.array([[0, 0], [0.25, 0.5], [0.75, 0.9], [1, 1]], np.float64)
    for k, t in enumerate((0.25, 0.5, 0.75)):
        b = np.array([(1 - t) ** 3, 3 * (1 - t) ** 2 * t, 3 * (1 - t) * t ** 2, t ** 3])
        x, y = b @ ctrl  # unit weights: the denominator is 1
        assert abs(tp.curve_x[k + 1] - x) < 1e-6 and abs(tp.curve_y[k + 1] - y) < 1e-6


def test_apophysis_curve_points_are_clamped_to_the_unit_square():
    # CurvePoint::new clamps x and y to [0, 1] (src/scene/tonemap.rs:134-139),
    # and parse_apophysis_curves builds every sampled point with it
    # (src/flame_xml.rs:1100-1110): a control point above the square gives
    # samples at y = 1, not past it
    cv = "0 0 1 0.25 2.5 1 0.75 2.5 1 1 1 1 " + "0 0 1 0.25 0.25 1 0.75 0.75 1 1 1 1 " * 3
    xml = ('<flame curves="%s"><xform weight="1" color="0" coefs="1 0 0 1 0 0" linear="1"/>'
           '</flame>') % cv
    tp = pyfae.load(xml).tonemap
    assert tp.curve_n == 5
    ys = [tp.curve_y[i] for i in range(5)]
    ctrl = np.array([[0, 0], [0.25, 2.5], [0.75, 2.5], [1, 1]], np.float64)
    for k, t in enumerate((0.25, 0.5, 0.75)):
        b = np.array([(1 - t) ** 3, 3 * (1 - t) ** 2 * t, 3 * (1 - t) * t ** 2, t ** 3])
        x, y = b @ ctrl
        assert y > 1.0
        assert ys[k + 1] == 1.0 and abs(tp.curve_x[k + 1] - x) < 1e-6
    assert max(ys) == 1.0 and min(ys) == 0.0
