"""The tail's effect kernels on the GPU against the oracle: every built-in
density and colour effect at its defaults and away from them (the blend
library's 13 modes through the vignette), on a non-square image whose
sides are not multiples of the 16x16 launch tile; a render with both chains
against the oracle's render; a config's effect lists through
ffm_render_config. Density effects agree to float rounding (exp and sqrt
are each side's own), colour effects within one level of RGBA8."""
import json

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402
import pyfae  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
H, W = 61, 97
RNG = np.random.default_rng(11)


def say(*a):
    print(*a, flush=True)


def density_image():
    img = RNG.random((H, W, 4)).astype(np.float32)
    # hit counts from sparse to dense around a mean of 1
    img[..., 3] = (RNG.random((H, W)) ** 3 * 4.0).astype(np.float32)
    return img


def colour_image():
    y, x = np.mgrid[0:H, 0:W]
    base = np.stack([np.sin(x * 0.21) * 0.5 + 0.5, np.cos(y * 0.17) * 0.5 + 0.5,
                     ((x + y) % 23) / 22.0], -1)
    img = np.clip(base + RNG.normal(0, 0.05, base.shape), 0, 1)
    rgba = np.concatenate([img * 255, RNG.integers(0, 256, (H, W, 1))], -1)
    return np.round(rgba).astype(np.uint8)


DENSITY_CASES = [
    ("density_blur", {}), ("density_blur", {"radius": 9.0, "threshold": 0.6, "falloff": 0.2}),
    ("density_blur", {"radius": 1.5}),
    ("bilateral_blur", {}), ("bilateral_blur", {"radius": 10.0, "sigma_spatial": 6.0,
                                                "sigma_range": 0.4}),
    ("sharpen", {}), ("sharpen", {"amount": 2.0, "radius": 4.5}),
]


@pytest.mark.parametrize("name,params", DENSITY_CASES)
def test_density_effect_equals_oracle(engine, name, params):
    say(f"density effect {name} {params}")
    img = density_image()
    fx = pyfae.effect(name, **params)
    src = torch.from_numpy(img).to(DEV)
    dst = torch.empty_like(src)
    engine.density_effect(fx, src, dst, density_mean=1.0, pixel_scale=1.0)
    torch.cuda.synchronize()
    got = dst.cpu().numpy()
    want = O.density_effect(fx, img, 1.0, 1.0)
    np.testing.assert_allclose(got, want, rtol=2e-5, atol=2e-6)


COLOUR_CASES = [
    ("vignette", {}), ("vignette", {"intensity": 1.0, "radius": 0.2, "softness": 0.9}),
    ("film_grain", {}), ("film_grain", {"intensity": 0.6, "seed": 17.0, "blend_mode": 5.0}),
    ("chromatic_aberration", {}), ("chromatic_aberration", {"amount": 9.0, "radial": 0.0}),
    ("chromatic_aberration", {"amount": 14.0, "intensity": 0.6, "blend_mode": 3.0}),
    ("hue_shift", {"offset": 75.0}), ("hue_shift", {"offset": 200.0, "intensity": 0.5,
                                                    "blend_mode": 11.0}),
    ("sobel_edges", {}), ("sobel_edges", {"threshold": 0.3, "glow": 1.0, "preserve_color": 0.0}),
    ("kaleidoscope", {}), ("kaleidoscope", {"segments": 9.0, "rotation": 40.0, "zoom": 0.7,
                                            "square_mode": 0.0, "edge_offset": 30.0}),
    ("plasma", {}), ("plasma", {"scale": 12.0, "time": 7.5, "direction": 10.0,
                                "blend_mode": 0.0, "intensity": 1.0}),
    ("tunnel", {}), ("tunnel", {"time": 4.0, "distortion": 0.7, "speed": 2.0}),
    ("julia", {}), ("julia", {"mode": 0.0, "color_mode": 0.0, "max_iter": 250.0, "zoom": 2.5,
                              "center_x": -0.6}),
    ("julia", {"color_mode": 2.0, "blend_mode": 2.0}),
    ("worley_noise", {}), ("worley_noise", {"mode": 1.0, "time": 12.0, "scale": 15.0}),
    ("worley_noise", {"mode": 3.0, "color_mode": 2.0, "blend_mode": 0.0}),
    ("simplex_noise", {}), ("simplex_noise", {"mode": 1.0, "intensity": 0.8, "octaves": 6.0}),
    ("simplex_noise", {"mode": 2.0, "scale": 11.0, "time": 3.0, "blend_mode": 0.0}),
    ("domain_warp", {}), ("domain_warp", {"intensity": 0.4, "scale": 7.0, "octaves": 2.0,
                                          "time": 20.0, "blend_mode": 5.0}),
    ("bloom", {}), ("bloom", {"threshold": 0.1, "softness": 0.8, "intensity": 3.0, "spread": 1.0}),
    ("distance_field", {}), ("distance_field", {"threshold": 0.5, "mode": 1.0, "band_count": 7.0}),
    ("distance_field", {"threshold": 0.6, "mode": 2.0, "spread": 0.6, "intensity": 0.8}),
    ("trails", {}),
] + [("vignette", {"intensity": 0.8, "radius": 0.3, "softness": 0.4, "blend_mode": float(m)})
     for m in range(13)]

# effects whose value wraps (fract) or jumps (atan2's branch, a segment
# boundary) where one more ulp on one side moves a pixel across: the
# platform's sin / atan2 bits decide those few pixels
WRAPS = {"kaleidoscope", "tunnel", "julia"}


@pytest.mark.parametrize("name,params", COLOUR_CASES)
def test_colour_effect_equals_oracle(engine, name, params):
    say(f"colour effect {name} {params}")
    img = colour_image()
    fx = pyfae.effect(name, **params)
    src = torch.from_numpy(img).to(DEV)
    dst = torch.empty_like(src)
    engine.color_effect(fx, src, dst, pixel_scale=1.0)
    torch.cuda.synchronize()
    got = dst.cpu().numpy().astype(int)
    want = O.color_effect(fx, img, 1.0).astype(int)
    d = np.abs(got - want)
    if name in WRAPS:
        # film grain's hash reads fract(sin(n) * 43758.5453) with n past
        # 32768 here, where sin is the library's on each side: one ulp
        # moves the hash by ~1e-3 and, next to a whole number, wraps it
        # from ~1 to ~0; the atan2 branch (kaleidoscope, tunnel) and the
        # fract wraps of tunnel and julia's smooth colouring move a few
        # pixels across a jump. The platforms' sin / atan2 / log bits
        # decide these. (Below 32768 both sides' sin is the same method,
        # so worley's hash agrees to the bit.)
        m = d.max(-1)
        assert (m > 1).mean() < 0.03 and (m > 8).mean() < 0.005, ((m > 1).mean(), (m > 8).mean())
    else:
        assert d.max() <= 1, (d.max(), np.argwhere(d > 1)[:5])
    if not (name == "simplex_noise" and params.get("mode") == 1.0):
        assert np.array_equal(got[..., 3], img[..., 3])  # the input's alpha kept
    # (simplex's distortion mode writes its warped read unblended, alpha too)


def test_film_grain_values_agree_modulo_one(engine):
    say("film grain raw")
    img = colour_image()
    fx = pyfae.effect("film_grain", intensity=1.0, blend_mode=0.0, seed=5.0)
    src = torch.from_numpy(img).to(DEV)
    dst = torch.empty_like(src)
    engine.color_effect(fx, src, dst)
    torch.cuda.synchronize()
    got = dst.cpu().numpy()[..., 0].astype(int)
    want = O.color_effect(fx, img)[..., 0].astype(int)
    d = np.abs(got - want)
    assert np.minimum(d, 255 - d).max() <= 2


STOPS = np.array([[0, 0.9, 0.1, 0.1], [0.5, 0.1, 0.9, 0.2], [1, 0.2, 0.3, 1.0]], np.float32)


def lin_flame():
    xs = []
    for i in range(3):
        ang = 2 * np.pi * i / 3 + 0.3
        xs.append(pyfae.xform(0.5, 0.02 * i, -0.02 * i, 0.5, 0.6 * np.cos(ang), 0.6 * np.sin(ang),
                              weight=1.0, color=i / 2.0, color_speed=0.5,
                              variations={"linear": 1.0}))
    return pyfae.flame_desc(xs)


def test_render_with_both_chains_equals_oracle(engine):
    say("render with effects")
    fd = lin_flame()
    engine.set_flame(fd)
    engine.set_palette(STOPS, 256)
    fx = pyfae.effects(density=[("density_blur", {"radius": 4.0}), ("sharpen", {"amount": 0.5})],
                       color=[("hue_shift", {"offset": 40.0}), ("bloom", {"intensity": 2.0}),
                              ("vignette", {"intensity": 0.7})])
    job = pyfae.make_job(200, 150, walkers=8192, ipt=64, target_samples=2_000_000, splat_mode=2,
                         zoom=1.1)
    try:
        engine.set_effects(fx)
        rgba, stats = engine.render(job)
        acc_dev = engine.render_accum(job)[0]
    finally:
        engine.set_effects(None)
    orgba, oacc, total = O.render(fd, STOPS, job, threads=8, effects=fx)
    assert stats.total_samples == total
    # the density chain's output is what the tone map read
    np.testing.assert_allclose(acc_dev, oacc, rtol=1e-4, atol=1e-5)
    d = np.abs(rgba.astype(int) - orgba.astype(int))
    say(f"  max diff {d.max()} mean {d.mean():.4f}")
    assert d.max() <= 2 and d.mean() < 0.05
    plain, _ = engine.render(job)  # cleared: no effects
    assert not np.array_equal(plain, rgba)


def test_config_effects_render_through_render_config(engine):
    say("config effects")
    t = json.load(open("tests/golden/ref_simple_linear.fflame"))
    t["density_effects"] = [{"effect_type": "bilateral_blur", "enabled": True,
                             "params": {"radius": 2.0}}]
    t["color_effects"] = [{"effect_type": "sobel_edges", "enabled": True, "params": {}},
                          {"effect_type": "film_grain", "enabled": True,
                           "params": {"intensity": 0.2}}]
    text = json.dumps(t)
    img = engine.render(text, 160, 120, iterations_per_thread=64, iterations=3_000_000)
    sc = pyfae.load(text)
    engine.set_scene(sc)
    job = pyfae.scene_job(sc, 160, 120, iterations=3_000_000, iterations_per_thread=64)
    try:
        engine.set_effects(sc.effects)
        rgba, _ = engine.render(job)
    finally:
        engine.set_effects(None)
    assert np.array_equal(img.to_numpy(), rgba)
    plain = engine.render(json.dumps(json.load(open("tests/golden/ref_simple_linear.fflame"))),
                          160, 120, iterations_per_thread=64, iterations=3_000_000)
    assert not np.array_equal(plain.to_numpy(), rgba)


def test_effect_parameters_that_would_run_away_are_refused(engine):
    for fx in (pyfae.effects(color=[("julia", {"max_iter": 6000.0})]),
               pyfae.effects(density=[("sharpen", {"radius": 80.0})])):
        with pytest.raises(pyfae.EngineError):
            engine.set_effects(fx)
