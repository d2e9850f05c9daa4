"""2x antialiasing (src/export/supersample.rs) on the CPU: the oracle's
downsample against the reference's own known-answer tests (:210-267)
restated, closed forms of the firefly clamp, and the job the config path
builds for a supersampled render (scale_config_for_supersample, :54-58)."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
import pyfae

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def quad_image(quads, w):
    """`quads` left to right on a (2 x 2w) image, as the reference's helper
    lays them out (samples (0,0), (1,0), (0,1), (1,1))."""
    img = np.zeros((2, 2 * w, 4), np.uint8)
    for q, quad in enumerate(quads):
        for i, px in enumerate(quad):
            dx, dy = [(0, 0), (1, 0), (0, 1), (1, 1)][i]
            img[dy, 2 * q + dx] = px
    return img


def test_uniform_quad_averages_exactly():
    out = O.downsample2x(quad_image([[[100, 150, 200, 255]] * 4], 1))
    assert out.ravel().tolist() == [100, 150, 200, 255]


def test_plain_average_when_no_outlier():
    out = O.downsample2x(quad_image([[[100] * 3 + [255], [120] * 3 + [255], [140] * 3 + [255],
                                      [160] * 3 + [255]]], 1))
    assert out.ravel()[:3].tolist() == [130, 130, 130]


def test_firefly_is_clamped():
    out = O.downsample2x(quad_image([[[2, 2, 2, 255]] * 3 + [[255, 255, 255, 255]]], 1))
    assert out[0, 0, 0] < 25
    assert out[0, 0, 3] == 255  # alpha is coverage: never clamped


def test_legitimate_bright_detail_survives():
    out = O.downsample2x(quad_image([[[10, 10, 10, 255]] * 2 + [[250, 250, 250, 255]] * 2], 1))
    assert out.ravel()[:3].tolist() == [130, 130, 130]


def test_firefly_closed_form():
    # three samples at lum L, one at 4L: median L, cap 4L + 64: unclamped;
    # a sample past the cap is scaled to exactly the cap's luminance
    rng = np.random.default_rng(3)
    for _ in range(200):
        q = rng.integers(0, 256, (4, 4)).astype(np.float64)
        lum = q[:, 0] + 2 * q[:, 1] + q[:, 2]
        med = np.sort(lum)[1:3].mean()
        cap = med * 4 + 64
        scale = np.where(lum > cap, cap / np.maximum(lum, 1e-30), 1.0)
        want = np.concatenate([(q[:, :3] * scale[:, None]).sum(0), [q[:, 3].sum()]]) / 4
        want = np.clip(np.floor(want + 0.5), 0, 255)
        out = O.downsample2x(quad_image([q.astype(np.uint8)], 1)).ravel().astype(float)
        assert np.abs(out - want).max() <= 1 and np.abs(out - want).mean() < 0.5


def test_downsample_is_per_quad_and_row_major():
    rng = np.random.default_rng(4)
    img = rng.integers(0, 256, (6, 10, 4)).astype(np.uint8)
    out = O.downsample2x(img)
    assert out.shape == (3, 5, 4)
    for y in range(3):
        for x in range(5):
            one = O.downsample2x(np.ascontiguousarray(img[2 * y:2 * y + 2, 2 * x:2 * x + 2]))
            assert np.array_equal(out[y, x], one[0, 0])


def test_supersampled_job_scales_iterations_filter_and_size():
    # scale_config_for_supersample: filter_radius x 2, max_iterations x 4;
    # the render is twice the size
    cfg = json.load(open(os.path.join(GOLDEN, "ref_simple_linear.fflame")))
    cfg["filter_radius"] = 1.5
    sc = pyfae.load(json.dumps(cfg))
    j1 = pyfae.scene_job(sc, 400, 300, iterations=1_000_000, iterations_per_thread=256)
    j2 = pyfae.scene_job(sc, 400, 300, iterations=1_000_000, iterations_per_thread=256,
                         supersample=True)
    assert (j2.view.width, j2.view.height) == (800, 600)
    assert j2.target_samples == 4 * j1.target_samples == 4_000_000
    assert j2.filter_radius == 2 * j1.filter_radius == 3.0
    assert j2.view.zoom == j1.view.zoom


def plotted(j, n):
    """Samples n dispatches of job j plot (every walker's first dispatch
    loses the burn-in; later ones plot ipt with persistent orbits)."""
    first = j.ipt - j.burn_in
    return j.walkers * (first + (n - 1) * (j.ipt if j.persistent_orbits else first))


def test_config_jobs_take_the_reference_dispatch_unless_asked_to_fill_the_gpu():
    sc = pyfae.load(os.path.join(GOLDEN, "ref_simple_linear.fflame"))
    ref = pyfae.scene_job(sc, 800, 600, iterations_per_thread=256)
    assert ref.walkers == 8192
    jobs = [ref]
    for kw, n in ((dict(iterations_per_thread=256), 2),  # 501M samples
                  (dict(iterations=100_000, iterations_per_thread=64), 2),
                  (dict(iterations=10_000_000, iterations_per_thread=256), 2),
                  (dict(iterations=20_000_000_000, iterations_per_thread=1024), 19)):
        j = pyfae.scene_job(sc, 64, 48, fill_gpu=True, **kw)
        jobs.append(j)
        # the fewest walkers, a multiple of 8192 up to 2^20, that plot the
        # budget in n dispatches (two at least)
        assert j.walkers % 8192 == 0 and 8192 <= j.walkers <= 1 << 20
        assert plotted(j, n) >= j.target_samples or j.walkers == 8192
        assert j.walkers == 8192 or plotted(j, n) - n * 8192 * j.ipt < j.target_samples
        if n > 2:  # 2^20 walkers would not plot the budget in one dispatch fewer
            assert plotted(j, n - 1) * (1 << 20) // j.walkers < j.target_samples
    assert jobs[1].walkers > 100 * 8192
    for j in jobs:
        assert (j.seed_base, j.seed_stride, j.burn_in) == (12345, 2654435761, 20)


def test_the_reference_visual_config_draws_its_walkers_start_points():
    # simple-linear is one transform, affine 0.5 I under linear weight 2: the
    # identity map. Its walkers never move, so its picture is the set of
    # their start points, and depends on how many walkers the render has:
    # why the config path keeps the reference's 8192 by default
    sc = pyfae.load(os.path.join(GOLDEN, "ref_simple_linear.fflame"))
    xf = sc.flame.xforms[0]
    assert sc.flame.n_xforms == 1 and sc.flame.n_finals == 0
    assert (xf.a, xf.b, xf.c, xf.d, xf.e, xf.f) == (0.5, 0.0, 0.0, 0.5, 0.0, 0.0)
    assert xf.var_weight[0] == 2.0 and sc.flame.var_id[0] == pyfae.VARIATIONS["linear"]
    fd, stops = sc.flame, pyfae.scene_stops(sc)
    for walkers, lit in ((8192, 8192), (1 << 14, 1 << 14)):
        job = pyfae.scene_job(sc, 800, 600, iterations=walkers * 50, iterations_per_thread=64)
        job.walkers = walkers
        _, acc, _ = O.render(fd, stops, job, pal_size=sc.palette_size, threads=8)
        # every walker sits on one pixel: at most `walkers` lit pixels, and
        # almost that many (a few share a pixel)
        n = int((acc[..., 3] > 0).sum())
        assert 0.9 * lit <= n <= lit


def test_the_gpu_filling_dispatch_is_not_for_a_slowly_mixing_flame():
    # Plastic spends 99% of its steps in one deterministic contraction and
    # mixes only through the 1% of its rare noisy transform: two dispatches of
    # a GPU-filling walker count sample its transient, not the picture the
    # reference's 8192 long persistent orbits draw -- why FFM_RENDER_FILL_GPU
    # is a choice the caller makes per flame, not the default
    text = open(os.path.join(GOLDEN, "ref_presets_2d.fae")).read()
    sc = pyfae.load(text, 2)
    assert sc.name == b"Plastic"
    stops = pyfae.scene_stops(sc)

    def density(fill, seed):
        j = pyfae.scene_job(sc, 160, 120, iterations=4_000_000, iterations_per_thread=64,
                            fill_gpu=fill)
        j.seed_base, j.de_max_radius = seed, 0.0
        _, acc, _ = O.render(sc.flame, stops, j, pal_size=sc.palette_size,
                             palette_xf=sc.palette_xf, threads=8)
        h = acc[:, :, 3].reshape(15, 8, 20, 8).sum((1, 3))
        return h / h.sum(), j.walkers

    ref, w0 = density(False, 12345)
    ref2, _ = density(False, 777)
    fill, w1 = density(True, 12345)
    assert w0 == 8192 and w1 > 4 * w0
    floor = np.abs(ref - ref2).sum()
    assert np.abs(fill - ref).sum() > 10 * floor
