"""2x antialiasing and the config path's dispatch on the GPU: ffm_downsample2x
equals the oracle's downsample_2x_firefly (src/export/supersample.rs:117-187)
bit for bit; a supersampled config render equals the oracle's render at
twice the size boxed down; the config path renders the reference's visual
config at its full budget and passes the reference's baseline; and its
GPU-filling dispatch renders a mixing flame within the Monte-Carlo noise of
the reference's own 8192-walker dispatch, in far fewer dispatches."""
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402
import pyfae  # noqa: E402
from test_gpu_formats import affine_config  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("ow,oh", [(1, 1), (37, 21), (640, 360)])
def test_downsample_equals_oracle(engine, ow, oh):
    say(f"downsample2x {ow}x{oh}")
    rng = np.random.default_rng(ow * 1000 + oh)
    img = rng.integers(0, 256, (2 * oh, 2 * ow, 4)).astype(np.uint8)
    # dark quads with one blown-out sample: the firefly clamp engages
    img[: 2 * (oh // 2), : 2 * (ow // 2)] //= 16
    img[0:2 * oh:8, 1:2 * ow:6, :3] = 255
    src = torch.from_numpy(img).to(DEV)
    dst = torch.zeros((oh, ow, 4), dtype=torch.uint8, device=DEV)
    engine.downsample2x(src, dst, ow, oh)
    torch.cuda.synchronize()
    want = O.downsample2x(img)
    assert np.array_equal(dst.cpu().numpy(), want)


def test_supersampled_config_render_equals_oracle(engine):
    say("render_config supersample=True vs oracle")
    cfg = affine_config()
    img = engine.render(cfg, 160, 120, iterations=1_000_000, iterations_per_thread=64,
                        supersample=True).to_numpy()
    assert img.shape == (120, 160, 4)
    sc = pyfae.load(cfg)
    job = pyfae.scene_job(sc, 160, 120, iterations=1_000_000, iterations_per_thread=64,
                          supersample=True)
    assert (job.view.width, job.view.height, job.target_samples) == (320, 240, 4_000_000)
    orgba, _, _ = O.render(sc.flame, pyfae.scene_stops(sc), job, pal_size=sc.palette_size,
                           threads=8, palette_xf=sc.palette_xf)
    want = O.downsample2x(orgba)
    d = np.abs(img.astype(int) - want.astype(int))
    say(f"  max diff {d.max()} frac>0 {(d > 0).mean():.5f}")
    # the tone map's own one level can move a quad's mean across a rounding
    assert d.max() <= 1 and (d > 0).mean() < 0.01


def block_sums(acc, b=16):
    h, w = acc.shape[:2]
    return acc[: h // b * b, : w // b * b, 3].reshape(h // b, b, w // b, b).sum((1, 3))


def test_reference_config_through_the_config_path_timed(engine):
    """simple-linear.fflame at 800x600 and its 501M-sample budget through
    ffm_render_config, the reference's 8192-walker dispatch: the dispatches
    go back to back (no host wait between them) and the picture passes the
    reference's baseline by its own rule."""
    say("config path: simple-linear 800x600, 501M samples")
    from test_oracle import visual_rule
    path = os.path.join(GOLDEN, "ref_simple_linear.fflame")
    engine.render(path, 800, 600, iterations=10_000_000, iterations_per_thread=256)  # warm
    img = engine.render(path, 800, 600, iterations_per_thread=256)
    mean, frac = visual_rule(img.to_numpy(), os.path.join(GOLDEN, "ref_2d_simple_linear.png"))
    say(f"  {img.info['dispatches']} dispatches, {img.info['ms']:.1f} ms; "
        f"baseline mean {mean:.3f} frac>40 {frac:.4f}")
    assert img.info["iterations"] >= 501187233 and mean <= 2.0 and frac <= 0.01


@pytest.mark.parametrize("i", [0, 3, 5])
def test_gpu_filling_dispatch_within_the_noise_floor(engine, i):
    """Presets whose walks mix within the burn-in (Grand Julian, Spherical3,
    Flower): the GPU-filling dispatch against the 8192-walker dispatch at its
    seed and at another."""
    text = open(os.path.join(GOLDEN, "ref_presets_2d.fae")).read()
    sc = pyfae.load(text, i)
    say(f"config path: GPU-filling vs 8192-walker dispatch, preset {i}")
    engine.set_scene(sc)
    kw = dict(iterations=80_000_000, iterations_per_thread=64)
    fill = pyfae.scene_job(sc, 320, 240, fill_gpu=True, **kw)
    ref = pyfae.scene_job(sc, 320, 240, **kw)
    ref2 = pyfae.scene_job(sc, 320, 240, **kw)
    ref.seed_base, ref2.seed_base = 12345, 777
    fill.seed_base = 12345
    for j in (fill, ref, ref2):
        j.de_max_radius = 0.0
    assert fill.walkers > 16 * ref.walkers
    a_fill, s_fill = engine.render_accum(fill)
    a_ref, s_ref = engine.render_accum(ref)
    a_ref2, _ = engine.render_accum(ref2)
    say(f"  GPU-filling: {s_fill.dispatches} dispatches, {s_fill.render_time_ms:.1f} ms; "
        f"8192 walkers: {s_ref.dispatches} dispatches, {s_ref.render_time_ms:.1f} ms")
    assert s_fill.total_samples >= fill.target_samples
    # densities per sample: the two dispatches plot slightly different totals
    b0, b1, b2 = (b / b.sum() for b in map(block_sums, (a_ref, a_ref2, a_fill)))
    floor = np.abs(b0 - b1).sum()
    diff = np.abs(b2 - b0).sum()
    say(f"  seed floor {floor:.5f} GPU-filling vs 8192 walkers {diff:.5f}")
    assert diff <= 1.5 * floor + 1e-3
