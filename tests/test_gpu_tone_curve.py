"""The tone curve on the GPU (shaders/tonemap.wgsl:479-492 with the table
of ToneCurve::generate_lut, src/scene/tonemap.rs:289-303) against the
oracle: the library's f16 table must equal the oracle's bit for bit, so a
pixel can differ only where the device's log/pow put a channel on the
other side of a texel boundary: by at most one table step (in levels,
after the encode) plus the tone map's own one level."""
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
        lut = O.curve_lut(tp).astype(np.float64)
        step = np.abs(np.diff(lut ** (1 / 2.2))).max() * 255
        assert d.max() <= int(np.ceil(step)) + 1, (name, d.max(), step)
        assert (d > 0).mean() < 0.02, (name, (d > 0).mean())


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
