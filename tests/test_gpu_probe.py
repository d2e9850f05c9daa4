"""Every variation on the reference's probe grid (src/probe/inputs.rs:44-99:
signed zeros, axes, diagonals, the unit circle, tiny, smallest-normal and
subnormal values, 1e20, either side of the 1e32 respawn threshold,
lopsided magnitudes) on the GPU against the oracle. A transform whose linear
part is zero maps every point to (e, f), so each chaos step evaluates the
variation at that probe point.

Where WGSL pins the arithmetic -- every point whose squared radius stays a
normal f32, |x|, |y| zero or in [1e-18, 1e18] -- the trace's records agree
in status, transform, colour, pixel and texel words, and in position to 4
ulp of the point's magnitude (bit-exact for the variations without
transcendentals). WGSL leaves quotients by divisors outside [2^-126, 2^126]
and subnormal values to the platform (the reference's probe exists to
classify exactly that), so at the overflow and underflow points the test
only prints where the two sides part: the device flushes subnormal
intermediates and gives NaN for a quotient by infinity (its fast division),
so such a point respawns where the oracle's IEEE arithmetic carries on."""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402
import pyfae  # noqa: E402
from pyfae import _abi  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
SN = 1.1754944e-38
PROBE = [(0.0, 0.0), (0.0, -0.0), (-0.0, 0.0), (-0.0, -0.0), (1.0, 0.0), (-1.0, 0.0),
         (0.0, 1.0), (0.0, -1.0), (1.0, 1.0), (-1.0, 1.0), (-1.0, -1.0), (1.0, -1.0),
         (0.5, 0.3), (0.12345679, -0.9876543), (2.5, -1.5), (math.pi, math.e),
         (0.70710677, 0.70710677), (1e-20, -1e-20), (SN, SN), (1e-40, -1e-40),
         (1e20, -1e20), (9e31, -9e31), (1e33, -1e33), (1e30, 1e-30), (1.0, -0.0), (-0.0, 1.0)]
EXACT = {"linear", "spherical", "hyperbolic", "bubble", "curl"}
STOPS = np.array([[0, 1, 0, 0], [1, 0, 0, 1]], np.float32)
VIEW = pyfae.View(64, 64, 1.0, 0.0, 0.0, 0.0)


def ulp_gap(got, want):
    """Largest position difference in ulps of the point's magnitude (a
    component near zero next to one near 1 is measured on the point's
    scale); non-finite entries excluded."""
    g = got[:, 2:4].copy().view(np.float32).astype(np.float64)
    w = want[:, 2:4].copy().view(np.float32)
    fin = np.isfinite(w) & np.isfinite(g)
    scale = np.maximum(np.abs(np.where(np.isfinite(w), w, 0)).max(1, keepdims=True),
                       np.float32(1e-30))
    d = np.where(fin, np.abs(g - w.astype(np.float64)) / np.spacing(scale), 0.0)
    return float(d.max())


def disagreement(name, got, want):
    """None when the traces agree to the bar, else what differs."""
    for k in (0, 1, 4, 5, 6, 7):
        if not np.array_equal(got[:, k], want[:, k]):
            return f"word {k}"
    g = got[:, 2:4].copy().view(np.float32)
    w = want[:, 2:4].copy().view(np.float32)
    if not (np.array_equal(np.isnan(g), np.isnan(w)) and np.array_equal(np.isinf(g), np.isinf(w))):
        return "class"
    if name in EXACT:
        return None if np.array_equal(got[:, 2:4], want[:, 2:4]) else "bits"
    gap = ulp_gap(got, want)
    return None if gap <= 4 else f"{gap:.0f} ulp"


def pinned(px, py):
    return all(v == 0.0 or 1e-18 <= abs(v) <= 1e18 for v in (px, py))


@pytest.mark.parametrize("name", sorted(_abi.VARIATIONS))
def test_variation_on_the_probe_grid(engine, name):
    print(f"probe grid {name}", flush=True)
    worst = 0
    parted = []
    for px, py in PROBE:
        fd = pyfae.flame_desc([pyfae.xform(0.0, 0.0, 0.0, 0.0, px, py, weight=1.0, color=0.3,
                                           color_speed=0.5, variations={name: 1.0})])
        engine.set_flame(fd)
        engine.set_palette(STOPS, 256)
        out = torch.zeros((6, 8), dtype=torch.int32, device=DEV)
        engine.trace(VIEW, 3, 17, 0, 6, out)
        torch.cuda.synchronize()
        got = out.cpu().numpy().view(np.uint32)
        want = O.trace(fd, STOPS, VIEW, 3, 17, 0, 6)
        why = disagreement(name, got, want)
        if pinned(px, py):
            assert why is None, (name, px, py, why)
            worst = max(worst, ulp_gap(got, want))
        elif why is not None:
            parted.append((px, py, why))
    print(f"  worst {worst:.2f} ulp; outside WGSL's pinned range the sides part at: {parted}",
          flush=True)
