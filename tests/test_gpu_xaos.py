"""Xaos on the GPU against the oracle: the selection chain step by step, the
integer histogram of a dispatch (atomic and LDS-tile splats), persistent
orbits with the stratified first pick and the refresh period across
dispatches, and a whole render. Affine flames, so every bar is bit-exact
(images within one level: the tone map's transcendentals)."""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402
import pyfae  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = "cuda:0"

# closed groups {1, 2} and {3}; transform 0 reaches both (the walk splits orbits)
SPLIT = [[1, 1, 0, 1], [0, 1, 1, 0], [0, 1, 1, 0], [0, 0, 0, 1]]
STOPS = np.array([[0, 0.9, 0.1, 0.1], [0.5, 0.1, 0.9, 0.2], [1, 0.2, 0.3, 1.0]], np.float32)


def say(*a):
    print(*a, flush=True)


def flame(xaos=SPLIT, weights=(0.5, 1.0, 0.8, 0.6)):
    xs = []
    for i, w in enumerate(weights):
        ang = 2 * np.pi * i / len(weights) + 0.3
        xs.append(pyfae.xform(0.48, 0.05 * i, -0.05 * i, 0.48, 0.55 * np.cos(ang),
                              0.55 * np.sin(ang), weight=w, color=i / 3.0, color_speed=0.4,
                              variations={"linear": 1.0}))
    return pyfae.flame_desc(xs, xaos=xaos)


VIEW = pyfae.View(256, 192, 1.1, 0.0, 0.0, 0.0)


def set_flame(engine, fd):
    engine.set_flame(fd)
    engine.set_palette(STOPS, 256)


def test_xaos_trace_bit_exact(engine):
    fd = flame()
    set_flame(engine, fd)
    for walker in (0, 4097, 123456):
        say(f"xaos trace walker {walker}")
        out = torch.zeros((300, 8), dtype=torch.int32, device=DEV)
        engine.trace(VIEW, walker, 3, 20, 300, out)
        torch.cuda.synchronize()
        got = out.cpu().numpy().view(np.uint32)
        want = O.trace(fd, STOPS, VIEW, walker, 3, 20, 300)
        np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("mode", [0, 2])
def test_xaos_histogram_bit_exact(engine, mode):
    say(f"xaos histogram mode {mode}")
    fd = flame()
    set_flame(engine, fd)
    h = torch.zeros((VIEW.height, VIEW.width, 4), dtype=torch.int32, device=DEV)
    cnt = torch.zeros(2, dtype=torch.int64, device=DEV)
    engine.iterate(VIEW, 1 << 15, 200, 20, 11, h, splat_mode=mode,
                   plotted=cnt[0:1].data_ptr(), inframe=cnt[1:2].data_ptr())
    torch.cuda.synchronize()
    want, plotted = O.iterate_mt(fd, STOPS, VIEW, 1 << 15, 200, 20, 11, 8)
    assert int(cnt[0]) == plotted
    assert np.array_equal(h.cpu().numpy().view(np.uint32), want)


@pytest.mark.parametrize("mode", [0, 2])
def test_xaos_persistent_orbits_and_refresh(engine, mode):
    """Four dispatches resuming orbits: the first starts every walker with
    the stratified pick, later ones resume (x, y, colour, fuse, previous
    transform) except the walkers the refresh period restarts."""
    fd = flame()
    assert O.xaos_refresh_period(fd) == 4
    set_flame(engine, fd)
    walkers, ipt, burn = 4096, 48, 20
    orb_d = torch.zeros((walkers, 4), dtype=torch.float32, device=DEV)
    orb_o = (O.Walker * walkers)()
    for k in range(4):
        seed = (12345 + k * 2654435761) & 0xFFFFFFFF
        say(f"xaos orbits mode {mode} dispatch {k}")
        h = torch.zeros((VIEW.height, VIEW.width, 4), dtype=torch.int32, device=DEV)
        engine.iterate(VIEW, walkers, ipt, burn, seed, h, orbits=orb_d, fresh=k == 0,
                       splat_mode=mode)
        torch.cuda.synchronize()
        want, _, _ = O.iterate(fd, STOPS, VIEW, walkers, ipt, burn, seed, orbits=orb_o,
                               fresh=k == 0, refresh_period=4)
        assert np.array_equal(h.cpu().numpy().view(np.uint32), want), k
        # the stored state: point and colour as the oracle's, fuse | prev << 24
        o = orb_d.cpu().numpy()
        ow = np.ctypeslib.as_array(C.cast(orb_o, C.POINTER(C.c_uint32)), (walkers * 5,))
        ow = ow.reshape(walkers, 5)
        np.testing.assert_array_equal(o[:, :3].view(np.uint32), ow[:, :3])
        np.testing.assert_array_equal(o[:, 3].view(np.uint32), ow[:, 3] | (ow[:, 4] << 24))


def test_xaos_render_equals_oracle(engine):
    say("xaos render")
    fd = flame()
    set_flame(engine, fd)
    job = pyfae.make_job(VIEW.width, VIEW.height, walkers=8192, ipt=64, target_samples=3_000_000,
                         splat_mode=2, zoom=VIEW.zoom)
    rgba, stats = engine.render(job)
    orgba, _, total = O.render(fd, STOPS, job, threads=8)
    assert stats.total_samples == total
    d = np.abs(rgba.astype(int) - orgba.astype(int))
    say(f"  samples {total} max diff {d.max()}")
    assert d.max() <= 1
