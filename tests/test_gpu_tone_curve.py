"""The tone curve on the GPU (shaders/tonemap.wgsl:479-492 with the table
of ToneCurve::generate_lut, src/scene/tonemap.rs:289-303) against the
oracle: the library's f16 table equals the oracle's bit for bit and both
read it with the same linear filter, so a pixel differs only by what the
device's log/pow move a channel: one level after the encode. The
reference's own rule for the identity curve (on = off to a level,
src/escape/app_repro_test.rs:10455-10507) is checked on the GPU too."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402
import pyfae  # noqa: E402
from pyfae import _abi, configs  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = "cuda:0"

CURVES = {
    "linear": [],
    "s_curve": [(0.0, 0.0), (0.25, 0.15), (0.5, 0.5), (0.75, 0.85), (1.0, 1.0)],
    "brighten_shadows": [(0.0, 0.0), (0.25, 0.4), (0.5, 0.6), (0.75, 0.8), (1.0, 1.0)],
    "darken_highlights": [(0.0, 0.0), (0.25, 0.2), (0.5, 0.4), (0.75, 0.6), (1.0, 1.0)],
    "steep_limited": [(0.0, 0.0), (0.1, 0.9), (1.0, 1.0)],
    "sixteen": list(zip(np.linspace(0, 1, 16).tolist(),
                        np.sort(np.random.default_rng(4).uniform(0, 1, 16)).tolist())),
}


def curve_tp(points, base):
    tp = base
    tp.use_curve = 1
    tp.curve_n = len(points)
    for i, (x, y) in enumerate(points):
        tp.curve_x[i], tp.curve_y[i] = x, y
    return tp


@pytest.mark.parametrize("name", sorted(CURVES))
def test_tonemap_with_curve_matches_oracle(engine, name):
    print("tone curve", name, flush=True)
    rng = np.random.default_rng(7)
    H, W = 64, 96
    acc = np.zeros((H, W, 4), np.float32)
    acc[..., :3] = rng.uniform(0, 1, (H, W, 3))
    acc[..., 3] = np.exp(rng.uniform(-8, 12, (H, W))).astype(np.float32)
    acc[::7, ::3] = 0
    for base in (_abi.tonemap_defaults(), configs.log_tonemap(2.2)):
        tp = curve_tp(CURVES[name], base)
        tp.background[:] = [0.1, 0.2, 0.05]
        o = O.tonemap(acc, 4.0 / 3.0, 5e6, tp)
        ta = torch.from_numpy(acc).to(DEV)
        tr = torch.zeros((H, W, 4), dtype=torch.uint8, device=DEV)
        engine.tonemap(ta, tr, W, H, 4.0 / 3.0, 5e6, tp)
        torch.cuda.synchronize()
        d = np.abs(tr.cpu().numpy().astype(int) - o.astype(int))
        assert d.max() <= 1, (name, d.max())
        assert (d > 0).mean() < 0.02, (name, (d > 0).mean())


def test_identity_curve_changes_nothing_on_gpu(engine):
    # a black-to-white ramp, linear tone map at gamma 1, no levels, black
    # background: the identity curve on is the curve off to a level
    H, W = 96, 128
    g = np.linspace(0.0, 1.0, H * W, dtype=np.float32).reshape(H, W)
    acc = np.zeros((H, W, 4), np.float32)
    acc[..., 0] = g ** np.float32(2.2)
    acc[..., 1] = g ** 4
    acc[..., 2] = g
    acc[..., 3] = 50.0
    ta = torch.from_numpy(acc).to(DEV)
    out = []
    for use in (1, 0):
        tp = _abi.tonemap_defaults()
        tp.tonemap_mode = 0
        tp.gamma = 1.0
        tp.levels_enabled = 0
        tp.background[:] = [0.0, 0.0, 0.0]
        tp.use_curve = use
        tr = torch.zeros((H, W, 4), dtype=torch.uint8, device=DEV)
        engine.tonemap(ta, tr, W, H, 4.0 / 3.0, 5e6, tp)
        torch.cuda.synchronize()
        out.append(tr.cpu().numpy().astype(int))
    assert (out[1][..., 0] < 40).sum() > 500
    assert np.abs(out[0] - out[1])[..., :3].max() <= 1


def test_bad_curves_are_refused(engine):
    H, W = 8, 8
    ta = torch.zeros((H, W, 4), dtype=torch.float32, device=DEV)
    tr = torch.zeros((H, W, 4), dtype=torch.uint8, device=DEV)
    for pts in ([(0.0, 0.0)], [(0.0, 0.0), (0.6, 0.5), (0.5, 1.0)],
                [(0.0, 0.0), (0.5, float("nan")), (1.0, 1.0)]):
        tp = curve_tp(pts, _abi.tonemap_defaults())
        with pytest.raises(pyfae.EngineError) as e:
            engine.tonemap(ta, tr, W, H, 1.0, 1e3, tp)
        assert e.value.status == _abi.ERR_INVALID_ARGUMENT
    tp = curve_tp([(0.0, 0.0), (1.0, 1.0)], _abi.tonemap_defaults())
    tp.curve_n = _abi.MAX_CURVE_POINTS + 1
    with pytest.raises(pyfae.EngineError) as e:
        engine.tonemap(ta, tr, W, H, 1.0, 1e3, tp)
    assert e.value.status == _abi.ERR_INVALID_ARGUMENT


def test_render_refuses_a_bad_curve_before_any_work(engine):
    # the tone map's parameters are checked at the top of the render, also
    # on the accumulator-only path, so a curve_n past the point arrays is
    # refused rather than read
    fd, pal, job = configs.plain(walkers=1024, ipt=32)
    engine.set_flame(fd)
    engine.set_palette(pal, 256)
    job.tonemap.use_curve = 1
    job.tonemap.curve_n = 1000
    for call in (engine.render_accum, engine._render_job):
        with pytest.raises(pyfae.EngineError) as e:
            call(job)
        assert e.value.status == _abi.ERR_INVALID_ARGUMENT
