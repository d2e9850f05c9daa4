"""GPU parity: the HIP path (through the C ABI) against the oracle.

Bars: bit-exact for the RNG stream, transform choices, bins and integer
histograms; per-step FP32 positions within 4 ulp of the step's scale for
transcendental flames (see test_step_parity_transcendental); whole images
within the seed-to-seed Monte-Carlo noise floor the oracle measures.
"""
import os
import struct

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402
import pyfae  # noqa: E402
from pyfae import configs  # noqa: E402

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
DEV = "cuda:0"


def say(*a):
    print(*a, flush=True)


def dev_u32(shape):
    return torch.zeros(shape, dtype=torch.int32, device=DEV)


def as_u32(t):
    return t.cpu().numpy().view(np.uint32)


def set_flame(engine, fd, pal, size=256):
    engine.set_flame(fd)
    engine.set_palette(pal, size)


def gpu_trace(engine, fd, pal, view, walker, seed, burn_in, steps):
    set_flame(engine, fd, pal)
    out = dev_u32((steps, 8))
    engine.trace(view, walker, seed, burn_in, steps, out)
    torch.cuda.synchronize()
    return as_u32(out)


def gpu_hist(engine, fd, pal, view, walkers, ipt, burn_in, seed, mode=0):
    set_flame(engine, fd, pal)
    dt = torch.float32 if mode in (3, 4) else torch.int32
    h = torch.zeros((view.height, view.width, 4), dtype=dt, device=DEV)
    cnt = torch.zeros(2, dtype=torch.int64, device=DEV)
    engine.iterate(view, walkers, ipt, burn_in, seed, h, splat_mode=mode, plotted=cnt[0:1].data_ptr(),
                   inframe=cnt[1:2].data_ptr())
    torch.cuda.synchronize()
    c = cnt.cpu().numpy()
    return (h.cpu().numpy() if mode in (3, 4) else as_u32(h)), int(c[0]), int(c[1])


# ------------------------------------------------------------ bit-exact

@pytest.mark.parametrize("walker", [0, 1, 777, 65535, 1 << 20])
def test_trace_linear_is_bit_exact(engine, walker):
    say("trace plain walker", walker)
    fd, pal, job = configs.plain()
    g = gpu_trace(engine, fd, pal, job.view, walker, 1, 20, 1000)
    o = O.trace(fd, pal, job.view, walker, 1, 20, 1000)
    assert np.array_equal(g, o)


def test_plain_histogram_is_bit_exact(engine):
    """The full plain workload: 2^16 walkers x 1000 iterations, 512x512."""
    say("plain histogram")
    fd, pal, job = configs.plain()
    g, plotted, inframe = gpu_hist(engine, fd, pal, job.view, 1 << 16, 1000, 20, 1)
    o, oplotted = O.iterate_mt(fd, pal, job.view, 1 << 16, 1000, 20, 1, 8)
    assert plotted == oplotted == (1 << 16) * 980
    assert inframe == int(o[..., 3].sum()) // 100
    assert np.array_equal(g, o)


def test_xcd_private_histograms_equal_global(engine):
    say("xcd vs global")
    fd, pal, job = configs.single4k(walkers=1 << 16, ipt=64, width=960, height=540)
    g0, p0, i0 = gpu_hist(engine, fd, pal, job.view, 1 << 16, 64, 20, 1, mode=0)
    g1, p1, i1 = gpu_hist(engine, fd, pal, job.view, 1 << 16, 64, 20, 1, mode=1)
    assert (p0, i0) == (p1, i1)
    assert np.array_equal(g0, g1)


def test_lds_tile_splat_equals_oracle_and_global(engine):
    """The binned LDS-tile splat gives the same integer histogram."""
    say("lds tile splat: plain vs oracle")
    fd, pal, job = configs.plain()
    g, plotted, inframe = gpu_hist(engine, fd, pal, job.view, 1 << 16, 1000, 20, 1, mode=2)
    o, oplotted = O.iterate_mt(fd, pal, job.view, 1 << 16, 1000, 20, 1, 8)
    assert plotted == oplotted
    assert np.array_equal(g, o)
    for (w, h, walkers, ipt) in ((960, 540, 1 << 16, 64), (3840, 2160, 1 << 18, 64), (300, 200, 5000, 37)):
        say(f"lds tile splat: single4k {w}x{h} {walkers}x{ipt}")
        fd, pal, job = configs.single4k(walkers=walkers, ipt=ipt, width=w, height=h)
        g0, p0, i0 = gpu_hist(engine, fd, pal, job.view, walkers, ipt, 20, 7, mode=0)
        g2, p2, i2 = gpu_hist(engine, fd, pal, job.view, walkers, ipt, 20, 7, mode=2)
        assert (p0, i0) == (p2, i2)
        assert np.array_equal(g0, g2)


def gpu_hist_dev(engine, fd, pal, view, walkers, ipt, burn_in, seed, mode):
    """As gpu_hist, but the histogram stays on the device."""
    set_flame(engine, fd, pal)
    dt = torch.float32 if mode in (3, 4) else torch.int32
    h = torch.zeros((view.height, view.width, 4), dtype=dt, device=DEV)
    cnt = torch.zeros(2, dtype=torch.int64, device=DEV)
    engine.iterate(view, walkers, ipt, burn_in, seed, h, splat_mode=mode, plotted=cnt[0:1].data_ptr(),
                   inframe=cnt[1:2].data_ptr())
    torch.cuda.synchronize()
    return h, tuple(int(v) for v in cnt.cpu())


def test_single4k_full_walk_lds_equals_atomics(engine):
    """The headline dispatch itself -- 2^24 walkers x 256 at 3840x2160 in
    one walker launch, 4.3e9 records, so bins' regions pass 2^32 record
    slots -- gives the u32 histogram of the global-atomic splat bit for bit."""
    say("single4k full walk: LDS splat (one launch) vs global atomics")
    fd, pal, job = configs.single4k()
    h2, c2 = gpu_hist_dev(engine, fd, pal, job.view, 1 << 24, 256, 20, 1, 2)
    h0, c0 = gpu_hist_dev(engine, fd, pal, job.view, 1 << 24, 256, 20, 1, 0)
    say(f"  plotted {c2[0]}, in frame {c2[1]} (regions sized 1.25x: past 2^32 slots)")
    assert c2 == c0
    assert torch.equal(h2, h0)
    del h2, h0


@pytest.mark.parametrize("side", [(6144, 3072), (12288, 12288), (32768, 32768)])
def test_lds_splat_with_tile_bins(engine, side):
    """Images past 2048 tiles sort records into bins of 2^k tiles (6144x3072:
    2 tiles per bin; 12288^2, the stress size: 16; 32768^2, the reference's
    largest side (src/config/defaults.rs:33), 64 tiles per bin in one 17 GB
    histogram where the reference would tile it). The integer histogram
    still equals the atomic splat's bit for bit; the float one to rounding."""
    w, h = side
    say(f"lds splat with tile bins {w}x{h}")
    fd, pal, job = configs.stress(walkers=1 << 18, ipt=32, side=w)
    job.view.height = h
    g0, c0 = gpu_hist_dev(engine, fd, pal, job.view, 1 << 18, 32, 20, 5, mode=0)
    g2, c2 = gpu_hist_dev(engine, fd, pal, job.view, 1 << 18, 32, 20, 5, mode=2)
    assert c0 == c2
    assert c0[1] > 1 << 20  # most samples land in frame
    assert torch.equal(g0, g2)
    del g0, g2
    g3, c3 = gpu_hist_dev(engine, fd, pal, job.view, 1 << 18, 32, 20, 5, mode=3)
    g4, c4 = gpu_hist_dev(engine, fd, pal, job.view, 1 << 18, 32, 20, 5, mode=4)
    assert c3 == c4
    assert torch.equal(g3[..., 3], g4[..., 3])  # sums of 100.0: exact
    assert torch.allclose(g4[..., :3], g3[..., :3], rtol=1e-4, atol=1e-3)


def test_region_overflow_fallback_is_exact(engine, monkeypatch):
    """Region capacities scaled to 1/1000 of the estimate
    (FFM_TILE_CAPACITY_SCALE) send nearly every record past its region to
    the direct-atomic fallback: the one-flame LDS splat still equals the
    global-atomic splat bit for bit (one tile a bin, and bins of two tiles),
    and a fresh context's batched gallery still equals its one-by-one
    renders."""
    say("region overflow fallback (capacities x 0.001)")
    monkeypatch.setenv("FFM_TILE_CAPACITY_SCALE", "0.001")
    for (w, h) in ((960, 540), (6144, 3072)):
        fd, pal, job = configs.single4k(walkers=1 << 16, ipt=64, width=w, height=h)
        g0, c0 = gpu_hist_dev(engine, fd, pal, job.view, 1 << 16, 64, 20, 3, 0)
        g2, c2 = gpu_hist_dev(engine, fd, pal, job.view, 1 << 16, 64, 20, 3, 2)
        assert c0 == c2
        assert torch.equal(g0, g2)
        del g0, g2
    other = pyfae.Engine(0)  # its batch plan captures the scaled capacities
    try:
        items = configs.gallery(n=4, spp=10.0, ipt=64)
        ob, sb = other.batch_render(items)
        os_, ss = other.batch_render(items, serial=True)
    finally:
        other.close()
    assert sb.total_samples == ss.total_samples
    for i in range(len(items)):
        assert np.array_equal(ob[i], os_[i]), i


@pytest.mark.parametrize("side", [(6144, 3072), (12288, 12288)])
def test_binned_splat_equals_oracle(engine, side):
    """The binned LDS splat (bins of 2 and 16 tiles) against the oracle's
    direct splat (shaders/core/accumulate_samples.wgsl:280-365 deposits the
    same u32 sums): an affine flame with a post-affine and a final, whose
    steps are +,* chains the device rounds as the oracle does, so the u32
    histogram is bit-identical; the f32 one has the same density sums
    exactly and colour sums between the u32 truncations and one per hit."""
    w, h = side
    say(f"binned splat vs oracle {w}x{h}")
    xs = [pyfae.xform(0.5, 0.0, 0.0, 0.5, -0.5, -0.5, weight=1.0, color=0.0),
          pyfae.xform(0.5, 0.1, -0.1, 0.5, 0.5, -0.5, weight=1.0, color=0.4,
                      post=(0.9, 0.05, -0.05, 0.9, 0.0, 0.02)),
          pyfae.xform(0.45, -0.2, 0.2, 0.45, 0.0, 0.5, weight=0.8, color=1.0)]
    fin = [pyfae.xform(1.1, 0.0, 0.0, 1.1, 0.0, 0.0)]
    fd = pyfae.flame_desc(xs, finals=fin)
    _, pal, _ = configs.plain()
    view = pyfae.View(w, h, 4.0 / 3.0, 0.0, 0.0, 0.0)
    walkers, ipt = 1 << 16, 64
    o, oplotted, oin = O.iterate(fd, pal, view, walkers, ipt, 20, 9)
    g, c = gpu_hist_dev(engine, fd, pal, view, walkers, ipt, 20, 9, mode=2)
    assert c == (oplotted, oin) and oin > walkers * 30
    assert np.array_equal(as_u32(g), o)
    del g
    g4, c4 = gpu_hist_dev(engine, fd, pal, view, walkers, ipt, 20, 9, mode=4)
    assert c4 == (oplotted, oin)
    g4 = g4.cpu().numpy()
    assert np.array_equal(g4[..., 3], o[..., 3].astype(np.float32))
    hits = o[..., 3:4].astype(np.float64) / 100.0
    d = g4[..., :3].astype(np.float64) - o[..., :3]
    assert (d >= -1e-3 * g4[..., :3] - 1e-3).all() and (d <= hits + 1e-3 * g4[..., :3] + 1e-3).all()


def test_single4k_full_size_with_density_estimation(engine):
    """The headline configuration at its own size, 3840x2160 with density
    estimation 9 -> 0 (curve 0.4) and the log tone map, at 2^18 walkers:
    the GPU's accumulator block densities and tone-mapped block colours
    differ from the oracle's render no more than a second oracle seed's."""
    say("single4k 3840x2160 with DE vs oracle")
    fd, pal, job = configs.single4k(walkers=1 << 18, ipt=256)
    assert job.de_max_radius == 9.0 and job.view.width == 3840
    set_flame(engine, fd, pal)
    gacc, _ = engine.render_accum(job)
    grgba, _ = engine.render(job)
    org1, o1, _ = O.render(fd, pal, job, threads=16)
    job2 = configs.single4k(walkers=1 << 18, ipt=256)[2]
    job2.seed_base = 2
    org2, o2, _ = O.render(fd, pal, job2, threads=16)
    b = 32
    bg, b1, b2 = block_sums(gacc, b), block_sums(o1, b), block_sums(o2, b)
    floor = np.abs(b1 - b2).sum() / b1.sum()
    diff = np.abs(bg - b1).sum() / b1.sum()

    def blocks(img):
        f = img[..., :3].astype(np.float64)
        hh, ww = f.shape[0] // b * b, f.shape[1] // b * b
        return f[:hh, :ww].reshape(hh // b, b, ww // b, b, 3).mean((1, 3))

    cf = np.abs(blocks(org1) - blocks(org2)).mean()
    cd = np.abs(blocks(grgba) - blocks(org1)).mean()
    say(f"  density floor {floor:.4f} gpu {diff:.4f}; colour floor {cf:.3f} gpu {cd:.3f}")
    assert diff <= 1.25 * floor + 1e-3
    assert cd <= 1.25 * cf + 0.05


def test_f32_lds_tile_splat_matches_f32_atomics(engine):
    say("f32 lds tile splat")
    fd, pal, job = configs.single4k(walkers=1 << 16, ipt=64, width=960, height=540)
    g3, p3, i3 = gpu_hist(engine, fd, pal, job.view, 1 << 16, 64, 20, 3, mode=3)
    g4, p4, i4 = gpu_hist(engine, fd, pal, job.view, 1 << 16, 64, 20, 3, mode=4)
    assert (p3, i3) == (p4, i4)
    assert np.array_equal(g3[..., 3], g4[..., 3])  # sums of 100.0: exact
    # float sums in another order: within a few ulp of the sums
    np.testing.assert_allclose(g4[..., :3], g3[..., :3], rtol=1e-4, atol=1e-3)


def test_f32_histogram_matches_u32(engine):
    say("f32 splat")
    fd, pal, job = configs.plain(walkers=1 << 14, ipt=200)
    g0, _, _ = gpu_hist(engine, fd, pal, job.view, 1 << 14, 200, 20, 1, mode=0)
    g3, _, _ = gpu_hist(engine, fd, pal, job.view, 1 << 14, 200, 20, 1, mode=3)
    # density channel: sums of 100.0, exact below 2^24
    assert np.array_equal(g3[..., 3], g0[..., 3].astype(np.float32))
    # colour: f32 keeps the fraction the u32 path truncates (< 1 per hit)
    hits = g0[..., 3] / 100.0
    d = g3[..., :3] - g0[..., :3]
    assert (d >= -1e-3 * g3[..., :3] - 1e-3).all() and (d <= hits[..., None] + 1e-3 * g3[..., :3] + 1e-3).all()


# ---------------------------------------------------- per-step tolerance

def ulp(x):
    x = np.abs(np.asarray(x, dtype=np.float32))
    return np.spacing(np.maximum(x, np.float32(2.0 ** -20))).astype(np.float64)


@pytest.mark.parametrize("cfg", ["single4k", "stress", "blurs"])
def test_step_parity_transcendental(engine, cfg):
    """One iterator step on the GPU vs the oracle's step from the same state,
    24000 steps over six walkers. RNG words, transform choices, respawn
    flags and the colour index are exact. x and y are within 4 ulp of the
    step's term scale at EVERY step: the largest magnitude among the terms
    the step sums (the point, the affine point, each weighted variation
    output, the post-affine products), which is the scale its FP32 sums
    round at -- an output that cancels terms of magnitude 1 down to 0.1
    carries their rounding, on any implementation (the oracle's own step
    is then ~7 ulp of the output from the exact value).

    The oracle's sin/cos/atan2 are the device's fixed method (pinned
    against double-precision libm by tests/test_fastmath.py). Against the
    host libm's instead, a one-ulp atan2 difference moves julia's
    theta/2 + pi by a whole ulp at magnitude up to 4.7 (and julian's angle
    alike), up to 8 ulp of a unit-scale term: that comparison is held to
    p99.9 <= 4 and max <= 8 ulp of the term scale."""
    say("step parity", cfg)
    fd, pal, job = getattr(configs, cfg)()
    errs, errs_libm = [], []
    for walker in (3, 7, 1000, 99999, 123456, 5555555):
        g = gpu_trace(engine, fd, pal, job.view, walker, 1, 20, 4000)
        o, term = O.replay_steps(fd, pal, job.view, walker, 1, 20, g, with_scale=True)
        assert np.array_equal(g[:, 0], o[:, 0]), "rng stream"
        assert np.array_equal(g[:, 1], o[:, 1]), "transform choice"
        assert np.array_equal(g[:, 7] & 2, o[:, 7] & 2), "respawns"
        # colour index: a pure +,* chain
        assert np.array_equal(g[:, 4], o[:, 4])
        gx = g[:, 2:4].view(np.float32).astype(np.float64)
        ox = o[:, 2:4].view(np.float32).astype(np.float64)
        ok = np.isfinite(ox).all(1) & np.isfinite(gx).all(1) & ((g[:, 7] & 2) == 0)
        errs.append((np.abs(gx - ox).max(1) / ulp(term))[ok])
        O.lib().orc_use_libm(1)
        try:
            ol, lterm = O.replay_steps(fd, pal, job.view, walker, 1, 20, g, with_scale=True)
        finally:
            O.lib().orc_use_libm(0)
        lx = ol[:, 2:4].view(np.float32).astype(np.float64)
        okl = np.isfinite(lx).all(1) & np.isfinite(gx).all(1) & ((g[:, 7] & 2) == 0)
        errs_libm.append((np.abs(gx - lx).max(1) / ulp(lterm))[okl])
    e, el = np.concatenate(errs), np.concatenate(errs_libm)
    say(f"  ulp of the term scale: median {np.median(e):.2f}  p99.9 {np.quantile(e, 0.999):.2f}"
        f"  max {e.max():.2f};  vs libm: p99.9 {np.quantile(el, 0.999):.2f}  max {el.max():.2f}")
    assert e.max() <= 4.0
    assert np.quantile(el, 0.999) <= 4.0 and el.max() <= 8.0


# -------------------------------------------------------------- stencils

def test_accumulate_bit_exact_and_count(engine):
    say("accumulate")
    rng = np.random.default_rng(1)
    H, W = 37, 53
    hist = rng.integers(0, 5000, (H, W, 4)).astype(np.uint32)
    hist[..., :3] = np.minimum(hist[..., :3], hist[..., 3:4])
    hist[::3, ::4] = 0
    prev = np.zeros((H, W, 4), np.float32)
    prev[..., 3] = rng.uniform(0, 1e6, (H, W)).astype(np.float32)
    prev[..., :3] = rng.uniform(0, 1, (H, W, 3))
    prev[::5, ::2] = 0
    for mode, blend in ((0, 0.0), (1, 0.25), (2, 0.0)):
        count = np.zeros((H, W, 2), np.uint32)
        o = O.accumulate(hist, prev, count=count, mode=mode, blend=blend)
        th = torch.from_numpy(hist.view(np.int32)).to(DEV)
        tp = torch.from_numpy(prev).to(DEV)
        to = torch.zeros_like(tp)
        tc = torch.zeros((H, W, 2), dtype=torch.int32, device=DEV)
        engine.accumulate(th, tp, to, W, H, count=tc, mode=mode, blend=blend)
        torch.cuda.synchronize()
        assert to.cpu().numpy().tobytes() == o.tobytes(), mode
        assert np.array_equal(as_u32(tc), count), mode


def test_tonemap_within_one_level(engine):
    say("tonemap")
    rng = np.random.default_rng(2)
    H, W = 64, 96
    acc = np.zeros((H, W, 4), np.float32)
    acc[..., :3] = rng.uniform(0, 1, (H, W, 3))
    acc[..., 3] = np.exp(rng.uniform(-3, 12, (H, W))).astype(np.float32)
    acc[::7, ::3] = 0
    tps = [pyfae.tonemap_defaults(), configs.log_tonemap(2.2)]
    for mode in (1, 2):  # transparent export: straight, premultiplied
        t = configs.log_tonemap(2.2)
        t.transparent_mode = mode
        tps.append(t)
    for tp in tps:
        tp.saturation = 1.3
        tp.background[:] = [0.1, 0.2, 0.05]
        o = O.tonemap(acc, 4.0 / 3.0, 5e6, tp)
        ta = torch.from_numpy(acc).to(DEV)
        tr = torch.zeros((H, W, 4), dtype=torch.uint8, device=DEV)
        engine.tonemap(ta, tr, W, H, 4.0 / 3.0, 5e6, tp)
        torch.cuda.synchronize()
        d = np.abs(tr.cpu().numpy().astype(int) - o.astype(int))
        assert d.max() <= 1 and (d > 0).mean() < 0.01


def test_blur_within_one_count(engine):
    say("blur")
    rng = np.random.default_rng(3)
    H, W = 40, 70
    hist = rng.integers(0, 20000, (H, W, 4)).astype(np.uint32)
    o = O.blur(hist, 1.7, 3000.0)
    th = torch.from_numpy(hist.view(np.int32)).to(DEV)
    sc = torch.zeros((2, H, W, 4), dtype=torch.int32, device=DEV)
    engine.blur(th, sc, W, H, 1.7, 3000.0)
    torch.cuda.synchronize()
    d = np.abs(as_u32(th).astype(np.int64) - o.astype(np.int64))
    assert d.max() <= 1


def test_density_estimate_matches_oracle(engine):
    say("density estimate")
    rng = np.random.default_rng(4)
    H, W = 50, 70
    acc = np.zeros((H, W, 4), np.float32)
    acc[..., 3] = np.where(rng.uniform(size=(H, W)) < 0.5, rng.integers(1, 400, (H, W)), 0)
    acc[..., :3] = rng.uniform(0, 1, (H, W, 3))
    o = O.density_estimate(acc, 9.0, 0.0, 0.4)
    ta = torch.from_numpy(acc).to(DEV)
    to = torch.zeros_like(ta)
    engine.density_estimate(ta, to, W, H, 9.0, 0.0, 0.4)
    torch.cuda.synchronize()
    g = to.cpu().numpy()
    np.testing.assert_allclose(g[..., 3], o[..., 3], rtol=2e-4, atol=1e-4)
    m = o[..., 3] > 1e-3
    np.testing.assert_allclose(g[..., :3][m], o[..., :3][m], rtol=2e-4, atol=1e-5)


# ---------------------------------------------------------- whole images

def test_render_matches_reference_baseline_image(engine):
    """The reference's own rendered baseline (tests/visual/baseline) through
    the GPU path, by the reference's visual rule."""
    say("render simple-linear vs reference baseline")
    from test_oracle import _simple_linear_job, visual_rule
    fd, stops, job = _simple_linear_job()
    set_flame(engine, fd, stops)
    rgba, stats = engine.render(job)
    assert stats.total_samples >= job.target_samples
    mean, frac = visual_rule(rgba, os.path.join(GOLDEN, "ref_2d_simple_linear.png"))
    say(f"  mean {mean:.3f} frac>40 {frac:.4f}")
    assert mean <= 2.0 and frac <= 0.01


def test_plain_render_equals_oracle(engine):
    say("plain render")
    fd, pal, job = configs.plain()
    set_flame(engine, fd, pal)
    rgba, stats = engine.render(job)
    orgba, oacc, tot = O.render(fd, pal, job, threads=8)
    assert stats.total_samples == tot
    d = np.abs(rgba.astype(int) - orgba.astype(int))
    assert d.max() <= 1


@pytest.mark.parametrize("after", [1, 3])
def test_exact_count_switch_equals_oracle(engine, after, monkeypatch):
    """The exact 64-bit density count from accumulate pass `after` on
    (FFM_EXACT_COUNT_AFTER, as the reference's FAE_EXACT_COUNT_AFTER):
    the accumulator equals the oracle's bit for bit (affine flame)."""
    say(f"exact count switch after {after}")
    monkeypatch.setenv("FFM_EXACT_COUNT_AFTER", str(after))
    fd, pal, job = configs.plain(walkers=1 << 14, ipt=100)
    job.view.width, job.view.height = 192, 128
    job.target_samples = (1 << 14) * 80 * 4 * 6
    set_flame(engine, fd, pal)
    gacc, stats = engine.render_accum(job)
    _, oacc, tot = O.render(fd, pal, job, threads=8, exact_after=after)
    assert stats.total_samples == tot
    np.testing.assert_array_equal(gacc, oacc)


def block_sums(acc, b=16):
    h, w = acc.shape[:2]
    d = acc[: h // b * b, : w // b * b, 3].reshape(h // b, b, w // b, b).sum((1, 3))
    return d


def test_single4k_image_within_noise_floor(engine):
    """Transcendental flame, reduced walkers: GPU vs oracle block densities
    differ no more than two oracle seeds differ from each other."""
    say("single4k noise floor")
    fd, pal, job = configs.single4k(walkers=1 << 15, ipt=256, width=960, height=540)
    job.de_max_radius = 0.0
    set_flame(engine, fd, pal)
    gacc, _ = engine.render_accum(job)
    _, o1, _ = O.render(fd, pal, job, threads=8)
    job2 = configs.single4k(walkers=1 << 15, ipt=256, width=960, height=540)[2]
    job2.de_max_radius = 0.0
    job2.seed_base = 2
    _, o2, _ = O.render(fd, pal, job2, threads=8)
    bg, b1, b2 = block_sums(gacc), block_sums(o1), block_sums(o2)
    floor = np.abs(b1 - b2).sum() / b1.sum()
    diff = np.abs(bg - b1).sum() / b1.sum()
    say(f"  floor {floor:.4f} gpu-vs-oracle {diff:.4f}")
    assert diff <= 1.25 * floor + 1e-3


# -------------------------------------------------------- batched gallery

@pytest.mark.parametrize("tile", [None, "128x64", "64x64"])
def test_batch_equals_serial_and_oracle(engine, tile, monkeypatch):
    """A small gallery (mixed sizes and transform counts) in one batched
    launch gives exactly the images of the same flames rendered one by one,
    and those match the oracle's within one level -- with the tuned tile
    shape and with each one fixed."""
    say(f"batch vs serial, tile {tile or 'tuned'}")
    if tile:
        monkeypatch.setenv("FFM_BATCH_TILE", tile)
    items = configs.gallery(n=6, spp=20.0, ipt=64)
    outs_b, st_b = engine.batch_render(items)
    outs_b2, _ = engine.batch_render(items)  # graph replay
    outs_s, st_s = engine.batch_render(items, serial=True)
    assert st_b.total_samples == st_s.total_samples
    for i, (a, a2, b) in enumerate(zip(outs_b, outs_b2, outs_s)):
        assert np.array_equal(a, b), i
        assert np.array_equal(a, a2), i
    for i in (0, 3):
        fd, pal, job = items[i]
        job.target_samples = 0
        job.persistent_orbits = 0
        orgba, _, _ = O.render(fd, pal, job, threads=8)
        d = np.abs(outs_b[i].astype(int) - orgba.astype(int))
        say(f"  flame {i}: {job.view.width}x{job.view.height} max diff {d.max()} frac>1 {(d > 1).mean():.5f}")
        # transcendental flame: trajectories diverge after a few ulp, so the
        # images agree statistically (block means), not pixel for pixel
        ga, oa = outs_b[i][..., :3].astype(float), orgba[..., :3].astype(float)
        b = 16
        h, w = ga.shape[0] // b * b, ga.shape[1] // b * b
        gm = ga[:h, :w].reshape(h // b, b, w // b, b, 3).mean((1, 3))
        om = oa[:h, :w].reshape(h // b, b, w // b, b, 3).mean((1, 3))
        assert np.abs(gm - om).mean() < 2.0


def test_batch_repeat_with_new_cameras_seeds_and_tonemaps(engine):
    """A second batched call with the same shapes (so the captured graph is
    reused) but every camera moved, every seed changed and every gamma
    changed renders the new pictures: each equals the same flame rendered
    one by one with the new job, and differs from the first call's."""
    say("batch: repeat call with new cameras, seeds and tone maps")
    items = configs.gallery(n=4, spp=10.0, ipt=64)
    first, _ = engine.batch_render(items)
    for fd, pal, job in items:
        job.view.pan_x += 0.05
        job.view.pan_y -= 0.03
        job.view.rotation += 0.1
        job.seed_base += 7
        job.tonemap.gamma = 1.8
    again, sb = engine.batch_render(items)
    serial, ss = engine.batch_render(items, serial=True)
    assert sb.total_samples == ss.total_samples
    for i in range(len(items)):
        assert np.array_equal(again[i], serial[i]), i
        assert not np.array_equal(again[i], first[i]), i


def test_batch_device_outputs_equal_host_outputs(engine):
    """The device-resident forms write the same pictures into caller device
    buffers: batched (then its graph replayed into other buffers) and one by
    one, against the host-buffer batched render."""
    say("batch: device outputs")
    items = configs.gallery(n=5, spp=10.0, ipt=64)
    outs_h, st_h = engine.batch_render(items)
    outs_d, st_d = engine.batch_render(items, device=True)
    outs_d2, _ = engine.batch_render(items, device=True)
    outs_sd, st_sd = engine.batch_render(items, serial=True, device=True)
    torch.cuda.synchronize()
    assert st_h.total_samples == st_d.total_samples == st_sd.total_samples
    for i, h in enumerate(outs_h):
        for d in (outs_d[i], outs_d2[i], outs_sd[i]):
            assert d.device.type == "cuda" and tuple(d.shape) == h.shape
            assert np.array_equal(d.cpu().numpy(), h), i


def test_batch_full_gallery_equals_serial(engine):
    """The whole 512-flame gallery (28.1e9 samples, so its waves stage
    billions of records) gives every flame exactly its one-by-one render:
    wave record buffers past 2^32 records once overflowed the bins' u32
    record offsets and mixed flames' records."""
    say("batch: the full 512-flame gallery against the one-by-one renders")
    items = configs.gallery(n=512)
    ob, sb = engine.batch_render(items, device=True)
    os_, ss = engine.batch_render(items, serial=True, device=True)
    torch.cuda.synchronize()
    assert sb.total_samples == ss.total_samples
    bad = [i for i in range(len(items)) if not torch.equal(ob[i], os_[i])]
    assert not bad, bad[:20]


def test_batch_tuning_cache_skips_the_tuning(engine, monkeypatch, tmp_path):
    say("batch: FFM_TUNE_CACHE reused by a new context")
    path = tmp_path / "tune.txt"
    monkeypatch.setenv("FFM_TUNE_CACHE", str(path))
    items = configs.gallery(n=3, spp=8.0, ipt=64)
    outs, _ = engine.batch_render(items)
    lines = path.read_text().splitlines()
    assert len(lines) == 1
    h, iw, fe, mode = lines[0].split()
    assert int(iw) in (2048, 8192, 32768) and int(fe) in (2, 4, 8) and int(mode) in (0, 1)
    other = pyfae.Engine(0)  # a fresh context: nothing tuned in memory
    try:
        outs2, _ = other.batch_render(items)
    finally:
        other.close()
    assert path.read_text().splitlines() == lines  # found, not tuned again
    for a, b in zip(outs, outs2):
        assert np.array_equal(a, b)


def test_batch_tile_override_is_checked(engine, monkeypatch):
    say("batch: FFM_BATCH_TILE refused unless 128x64 or 64x64")
    items = configs.gallery(n=2, spp=4.0, ipt=64)
    monkeypatch.setenv("FFM_BATCH_TILE", "32x32")
    with pytest.raises(pyfae.EngineError) as e:
        engine.batch_render(items)
    assert "FFM_BATCH_TILE" in str(e.value)


# ------------------------------------------------------------ error paths

def test_errors_are_status_codes(engine):
    say("errors")
    e2 = pyfae.Engine(0)
    v = pyfae.View(64, 64, 1.0, 0, 0, 0)
    h = dev_u32((64, 64, 4))
    with pytest.raises(pyfae.EngineError) as ei:
        e2.iterate(v, 64, 10, 0, 1, h)
    assert ei.value.status == pyfae._abi.ERR_NO_FLAME
    fd = pyfae.flame_desc([pyfae.xform()])
    e2.set_flame(fd)
    with pytest.raises(pyfae.EngineError) as ei:
        e2.iterate(v, 64, 10, 0, 1, h)
    assert ei.value.status == pyfae._abi.ERR_NO_PALETTE
    e2.set_palette([[0, 1, 1, 1]])
    with pytest.raises(pyfae.EngineError) as ei:
        e2.iterate(pyfae.View(40000, 10, 1, 0, 0, 0), 64, 10, 0, 1, h)
    assert ei.value.status == pyfae._abi.ERR_TOO_LARGE
    with pytest.raises(pyfae.EngineError) as ei:
        e2.render(pyfae.make_job(64, 64, walkers=0))
    assert ei.value.status == pyfae._abi.ERR_INVALID_ARGUMENT
    e2.close()


# ------------------------------------------------------ final attachments

@pytest.mark.parametrize("attach", [None, [[], [1, 0], [0]], [[0, 1], [0, 1], [0, 1]]])
def test_final_attachments_bit_exact(engine, attach):
    """Per-transform final chains (the reference's AttachmentList) on a
    linear flame: the device trace and the LDS-splat histogram equal the
    oracle's bit for bit, for the shared chain (uniform across a wave) and
    for chains that differ by transform."""
    say("final attachments", attach)
    rng = configs.SplitMix64(77)
    xs = []
    for i in range(3):
        a = configs._affine(rng, 0.7)
        xs.append(pyfae.xform(*a, weight=1.0, color=i / 2.0, variations={"linear": 1.0},
                              final_attachments=None if attach is None else attach[i]))
    fins = [pyfae.xform(0.9, 0.1, -0.1, 0.9, 0.05, 0.0, variations={"linear": 1.0}),
            pyfae.xform(0.8, 0.0, 0.0, -0.8, 0.0, 0.1, variations={"linear": 1.0})]
    fd = pyfae.flame_desc(xs, fins)
    pal = configs.random_palette(configs.SplitMix64(5))
    view = pyfae.View(320, 240, 1.0, 0.0, 0.0, 0.0)
    for w in (0, 4097):
        assert np.array_equal(gpu_trace(engine, fd, pal, view, w, 1, 20, 300),
                              O.trace(fd, pal, view, w, 1, 20, 300))
    g, plotted, _ = gpu_hist(engine, fd, pal, view, 1 << 14, 200, 20, 3, mode=2)
    o, oplotted = O.iterate_mt(fd, pal, view, 1 << 14, 200, 20, 3, 8)
    assert plotted == oplotted
    assert np.array_equal(g, o)


def test_flame_without_variation_slots_bit_exact(engine):
    """A flame with no variation slots (every transform's sum is empty, so
    the point is its post-affine of the origin) takes the EXT kernels; its
    trace and histograms equal the oracle's bit for bit."""
    say("no variation slots")
    rng = configs.SplitMix64(99)
    xs = []
    for i in range(3):
        a = configs._affine(rng, 0.7)
        xs.append(pyfae.xform(*a, weight=1.0, color=i / 2.0, variations={"linear": 0.0},
                              post=(0.9, 0.1, -0.1, 0.9, 0.3 * i, -0.2 * i)))
    fd = pyfae.flame_desc(xs)
    fd.n_vars = 0
    pal = configs.random_palette(configs.SplitMix64(6))
    view = pyfae.View(320, 240, 1.0, 0.0, 0.0, 0.0)
    for w in (0, 777):
        assert np.array_equal(gpu_trace(engine, fd, pal, view, w, 1, 20, 200),
                              O.trace(fd, pal, view, w, 1, 20, 200))
    for mode in (0, 2):
        g, plotted, _ = gpu_hist(engine, fd, pal, view, 1 << 12, 100, 20, 3, mode=mode)
        o, oplotted = O.iterate_mt(fd, pal, view, 1 << 12, 100, 20, 3, 8)
        assert plotted == oplotted
        assert np.array_equal(g, o)


@pytest.mark.parametrize("kind", ["dense", "negative", "max"])
def test_selection_buckets_bit_exact(engine, kind):
    """The pick through the device's selection buckets (1024 buckets of the
    24-bit draw) equals select_transform_const's scan, on linear flames
    whose traces and histograms are bit-exact: many boundaries inside one
    bucket (weights from 1e-7 to 1), a negative weight (the cumulative sums
    decrease, so the buckets are off and the scan runs), and the 128-transform
    maximum."""
    say("selection buckets", kind)
    rng = configs.SplitMix64(4242)
    n = {"dense": 40, "negative": 7, "max": 128}[kind]
    xs = []
    for i in range(n):
        a = configs._affine(rng, 0.7)
        w = 10.0 ** (-7.0 * rng.uniform()) if kind == "dense" else rng.uniform(0.1, 1.0)
        if kind == "negative" and i == 3:
            w = -0.4
        xs.append(pyfae.xform(*a, weight=w, color=i / (n - 1), variations={"linear": 1.0}))
    fd = pyfae.flame_desc(xs)
    pal = configs.random_palette(configs.SplitMix64(6))
    view = pyfae.View(320, 240, 1.0, 0.0, 0.0, 0.0)
    for w in (0, 12345):
        assert np.array_equal(gpu_trace(engine, fd, pal, view, w, 1, 20, 400),
                              O.trace(fd, pal, view, w, 1, 20, 400))
    for mode in (0, 2):
        g, plotted, _ = gpu_hist(engine, fd, pal, view, 1 << 14, 200, 20, 3, mode=mode)
        o, oplotted = O.iterate_mt(fd, pal, view, 1 << 14, 200, 20, 3, 8)
        assert plotted == oplotted
        assert np.array_equal(g, o)
