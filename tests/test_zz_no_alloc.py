"""ffm_iterate_noalloc through the C ABI against ffm_iterate: the same
dispatch with the LDS tile splat's tables and records in the library's
per-GPU device scratch instead of context buffers. Integer histograms,
sample counters and orbits equal ffm_iterate's bit for bit; float
histograms to rounding (f32 atomics in another order). Calls run back to
back, the later ones smaller, so each starts on what the last one left in
the shared scratch."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import pyfae  # noqa: E402
from pyfae import _abi, configs  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def say(*a):
    print(*a, flush=True)


def run(engine, fn, view, walkers, ipt, seed, mode, orbits=False):
    dt = torch.float32 if mode in (3, 4) else torch.int32
    h = torch.zeros((view.height, view.width, 4), dtype=dt, device=DEV)
    cnt = torch.zeros(2, dtype=torch.int64, device=DEV)
    orb = torch.zeros((walkers, 4), dtype=torch.float32, device=DEV) if orbits else None
    fn(view, walkers, ipt, 20, seed, h, splat_mode=mode, orbits=orb, fresh=True,
       plotted=cnt[0:1].data_ptr(), inframe=cnt[1:2].data_ptr())
    torch.cuda.synchronize()
    return h, tuple(int(v) for v in cnt.cpu()), orb


def compare(engine, view, walkers, ipt, seed, mode, orbits=False):
    h0, c0, o0 = run(engine, engine.iterate, view, walkers, ipt, seed, mode, orbits)
    h1, c1, o1 = run(engine, engine.iterate_noalloc, view, walkers, ipt, seed, mode, orbits)
    assert c0 == c1 and c0[1] > 0
    if mode in (3, 4):
        assert torch.equal(h0[..., 3], h1[..., 3])  # sums of 100.0: exact
        assert torch.allclose(h1[..., :3], h0[..., :3], rtol=1e-4, atol=1e-3)
    else:
        assert torch.equal(h0, h1)
    if orbits:
        assert torch.equal(o0, o1)


def test_noalloc_iterate_equals_iterate(engine):
    # 1: 4K, 2^20 walkers x 64: 68M records, more than the scratch's record
    # space takes at once, so the walkers run in two chunks
    fd, pal, job = configs.single4k(walkers=1 << 20, ipt=64)
    engine.set_flame(fd)
    engine.set_palette(pal, 256)
    say("noalloc: single4k 3840x2160, 2^20 x 64, u32 LDS splat")
    compare(engine, job.view, 1 << 20, 64, 3, 2)
    # 2: a smaller input right after, on what call 1 left in the scratch;
    # persistent orbits too
    fd, pal, job = configs.plain(walkers=1 << 14, ipt=200)
    engine.set_flame(fd)
    engine.set_palette(pal, 256)
    say("noalloc: plain 512x512, 2^14 x 200, u32 and f32 LDS splat, global atomics")
    compare(engine, job.view, 1 << 14, 200, 1, 2, orbits=True)
    compare(engine, job.view, 1 << 14, 200, 1, 4)
    compare(engine, job.view, 1 << 12, 100, 9, 0)


def test_noalloc_iterate_with_tile_bins(engine):
    # 6144x3072 has 2304 tiles: bins of two tiles, records sorted by tile in
    # the second half of the scratch
    fd, pal, job = configs.stress(walkers=1 << 17, ipt=32, side=6144)
    job.view.height = 3072
    engine.set_flame(fd)
    engine.set_palette(pal, 256)
    say("noalloc: 6144x3072 tile bins, 2^17 x 32")
    compare(engine, job.view, 1 << 17, 32, 5, 2)
    say("noalloc: 960x540 after it")
    fd, pal, job = configs.single4k(walkers=1 << 15, ipt=64, width=960, height=540)
    engine.set_flame(fd)
    engine.set_palette(pal, 256)
    compare(engine, job.view, 1 << 15, 64, 2, 2)


def test_noalloc_refusals(engine):
    fd, pal, job = configs.plain(walkers=1 << 10, ipt=10)
    engine.set_flame(fd)
    engine.set_palette(pal, 256)
    h = torch.zeros((512, 512, 4), dtype=torch.int32, device=DEV)
    # one walker's records past three quarters of the record space
    with pytest.raises(pyfae.EngineError) as e:
        engine.iterate_noalloc(job.view, 1 << 10, 60_000_000, 20, 1, h, splat_mode=2)
    assert e.value.status == _abi.ERR_TOO_LARGE
    # the per-XCD histograms need memory the scratch does not have
    with pytest.raises(pyfae.EngineError) as e:
        engine.iterate_noalloc(job.view, 1 << 10, 10, 20, 1, h, splat_mode=1)
    assert e.value.status == _abi.ERR_INVALID_ARGUMENT
