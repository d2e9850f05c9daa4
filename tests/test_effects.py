"""The render tail's effects in the oracle (CPU): the reference's own effect
tests restated (src/renderer/effect_chain.rs:1378-1498), closed forms of
each effect and of the blend library, an independent float64 restatement of
the blurs on non-square images, and the config loader's effect lists."""
import json

import numpy as np
import pytest

import pyfae
from oracle import oracle as O

RNG = np.random.default_rng(7)


def stripes(W=32, H=8):
    """Columns alternating red 0 and 1, 1000 hits each."""
    img = np.zeros((H, W, 4), np.float32)
    img[..., 0] = (np.arange(W) % 2)[None, :]
    img[..., 3] = 1000.0
    return img


# ---------------------------------------------------------- reference KATs

def test_a_hit_count_past_half_float_survives_the_density_chain():
    img = np.tile(np.array([0.5, 0.25, 0.75, 1.0e6], np.float32), (4, 64, 1))
    out = O.density_effect(pyfae.effect("sharpen", amount=0.0), img, 1.0)
    assert np.all(np.abs(out[..., 3] / 1.0e6 - 1.0) < 1e-6)
    assert np.all(np.abs(out[..., 0] - 0.5) < 1e-6) and np.all(np.abs(out[..., 2] - 0.75) < 1e-6)


def test_density_blur_measures_density_in_multiples_of_the_mean():
    img = stripes()
    fx = pyfae.effect("density_blur")
    sparse = O.density_effect(fx, img, 10_000.0)[4, 16, 0]
    assert abs(sparse - 0.5) < 0.2
    dense = O.density_effect(fx, img, 100.0)[4, 16, 0]
    assert dense in (0.0, 1.0)


def test_a_size_in_pixels_follows_the_pixel_scale():
    img = stripes()
    b = lambda r, s: O.density_effect(pyfae.effect("density_blur", radius=r), img, 10_000.0, s)
    assert np.array_equal(b(8.0, 0.5), b(4.0, 1.0))
    assert not np.array_equal(b(8.0, 1.0), b(4.0, 1.0))
    W, H = 32, 8
    i = np.arange(W * H)
    col = np.stack([(i % W) * 8, np.zeros_like(i), (i // W) * 30, np.full_like(i, 255)], -1)
    col = col.astype(np.uint8).reshape(H, W, 4)
    ca = lambda a, s: O.color_effect(pyfae.effect("chromatic_aberration", amount=a, radial=0.0),
                                     col, s)
    assert np.array_equal(ca(6.0, 0.5), ca(3.0, 1.0))
    assert not np.array_equal(ca(6.0, 1.0), ca(3.0, 1.0))


# ------------------------------------------------------------ closed forms

@pytest.mark.parametrize("name", ["density_blur", "bilateral_blur", "sharpen"])
def test_density_effects_keep_a_flat_image(name):
    img = np.tile(np.array([0.3, 0.6, 0.1, 0.02], np.float32), (9, 13, 1))
    out = O.density_effect(pyfae.effect(name), img, 1.0)
    np.testing.assert_allclose(out, img, rtol=2e-6, atol=0)


def test_dense_pixels_and_sharpen_off_are_exact():
    img = RNG.random((11, 17, 4)).astype(np.float32)
    img[..., 3] = 50.0  # 50x the mean: past threshold + falloff, no blur at all
    assert np.array_equal(O.density_effect(pyfae.effect("density_blur"), img, 1.0), img)
    assert np.array_equal(O.density_effect(pyfae.effect("sharpen", amount=0.005), img, 1.0), img)


def bilinear64(img, u, v):
    H, W = img.shape[:2]
    p = np.array([u * W - 0.5, v * H - 0.5])
    p = np.clip(p, -1, [W, H])
    f = np.floor(p)
    t = p - f
    x0, y0 = int(np.clip(f[0], 0, W - 1)), int(np.clip(f[1], 0, H - 1))
    x1, y1 = int(np.clip(f[0] + 1, 0, W - 1)), int(np.clip(f[1] + 1, 0, H - 1))
    top = img[y0, x0] * (1 - t[0]) + img[y0, x1] * t[0]
    bot = img[y1, x0] * (1 - t[0]) + img[y1, x1] * t[0]
    return top * (1 - t[1]) + bot * t[1]


def test_blurs_against_a_float64_restatement_on_a_non_square_image():
    H, W = 5, 7
    img = RNG.random((H, W, 4)).astype(np.float32)
    img[..., 3] = RNG.random((H, W)) * 0.2  # sparse: density_blur blurs fully
    img64 = img.astype(np.float64)
    db = O.density_effect(pyfae.effect("density_blur", radius=3.0), img, 1.0)
    bl = O.density_effect(pyfae.effect("bilateral_blur", radius=2.0, sigma_spatial=1.5,
                                       sigma_range=0.3), img, 1.0)
    sh = O.density_effect(pyfae.effect("sharpen", amount=0.7, radius=1.5), img, 1.0)
    lum = lambda c: c[0] * 0.2126 + c[1] * 0.7152 + c[2] * 0.0722
    for y in range(H):
        for x in range(W):
            u, v = (x + 0.5) / W, (y + 0.5) / H
            c = img64[y, x]
            # density_blur at radius 3: 9x9 taps one pixel apart within the disc
            s, ws = np.zeros(4), 0.0
            for yi in range(-4, 5):
                for xi in range(-4, 5):
                    d = np.hypot(xi, yi)
                    if d <= 3.0:
                        w = np.exp(-d * d / (2 * 1.5 ** 2))
                        s += bilinear64(img64, u + xi / W, v + yi / H) * w
                        ws += w
            np.testing.assert_allclose(db[y, x], s / ws, rtol=1e-5, atol=1e-6)
            s, ws = np.zeros(4), 0.0
            for yi in range(-2, 3):
                for xi in range(-2, 3):
                    t = bilinear64(img64, u + xi / W, v + yi / H)
                    w = (np.exp(-(xi * xi + yi * yi) / (2 * 1.5 ** 2))
                         * np.exp(-(lum(c) - lum(t)) ** 2 / (2 * 0.3 ** 2)))
                    s += t * w
                    ws += w
            np.testing.assert_allclose(bl[y, x], s / ws, rtol=1e-5, atol=1e-6)
            s, ws = np.zeros(4), 0.0
            for yi in range(-2, 3):
                for xi in range(-2, 3):
                    d = np.hypot(xi, yi)
                    if d <= 1.5:
                        w = np.exp(-d * d / (2 * 0.75 ** 2))
                        s += bilinear64(img64, u + xi / W, v + yi / H) * w
                        ws += w
            want = np.clip(c[:3] + (c[:3] - s[:3] / ws) * 0.7, 0, 1)
            np.testing.assert_allclose(sh[y, x, :3], want, rtol=1e-5, atol=1e-6)
            assert sh[y, x, 3] == img[y, x, 3]


def test_blend_library_closed_forms():
    base = RNG.random((50, 3)).astype(np.float32)
    eff = RNG.random((50, 3)).astype(np.float32)
    one, zero = np.ones(3, np.float32), np.zeros(3, np.float32)
    for b, e in zip(base, eff):
        assert np.array_equal(O.apply_blend(b, e, 0, 1.0), e)          # normal
        for m in range(13):
            assert np.array_equal(O.apply_blend(b, e, m, 0.0), b)      # intensity 0
        np.testing.assert_allclose(O.apply_blend(b, e, 1, 1.0), np.minimum(b + e, 1), atol=0)
        assert np.array_equal(O.apply_blend(b, one, 2, 1.0), b)        # multiply by white
        assert np.array_equal(O.apply_blend(b, e, 2, 1.0), O.apply_blend(e, b, 2, 1.0))
        np.testing.assert_allclose(O.apply_blend(b, zero, 3, 1.0), b, atol=1e-7)  # screen black
        # screen is the complement of multiplying the complements
        np.testing.assert_allclose(O.apply_blend(b, e, 3, 1.0),
                                   1 - O.apply_blend(1 - b, 1 - e, 2, 1.0), atol=1e-6)
        # overlay and hard light swap roles
        np.testing.assert_allclose(O.apply_blend(b, e, 4, 1.0), O.apply_blend(e, b, 6, 1.0),
                                   atol=1e-6)
        np.testing.assert_allclose(O.apply_blend(b, zero, 7, 1.0), b / 1.00001, rtol=1e-6)
        np.testing.assert_allclose(O.apply_blend(b, one, 8, 1.0),
                                   1 - (1 - b) / 1.00001, atol=1e-6)
        for m in (9, 10, 11, 12):  # HSL modes of a colour with itself
            np.testing.assert_allclose(O.apply_blend(b, b, m, 1.0), b, atol=2e-6)
    # luminosity: the effect's HSL lightness on the base's hue and saturation
    r = O.apply_blend(np.array([0.8, 0.2, 0.2], np.float32), np.array([0.3, 0.3, 0.3], np.float32),
                      12, 1.0)
    np.testing.assert_allclose((r.max() + r.min()) / 2, 0.3, atol=1e-6)
    np.testing.assert_allclose(r, [0.48, 0.12, 0.12], atol=1e-6)
    # hue of pure green on red's saturation and lightness: pure green
    g = O.apply_blend(np.array([1, 0, 0], np.float32), np.array([0, 0.4, 0], np.float32), 9, 1.0)
    np.testing.assert_allclose(g, [0, 1, 0], atol=1e-6)


def test_colour_effects_identities():
    img = (RNG.random((13, 19, 4)) * 255).astype(np.uint8)
    flat = np.tile(np.array([40, 120, 200, 77], np.uint8), (13, 19, 1))
    assert np.array_equal(O.color_effect(pyfae.effect("film_grain", intensity=0.0), img), img)
    assert np.array_equal(O.color_effect(pyfae.effect("chromatic_aberration", amount=0.0), img),
                          img)
    assert np.array_equal(O.color_effect(pyfae.effect("sobel_edges"), flat), flat)
    for off in (0.0, 360.0):
        d = O.color_effect(pyfae.effect("hue_shift", offset=off), img).astype(int) - img
        assert np.abs(d).max() <= 1 and np.array_equal(d[..., 3], 0 * d[..., 3])
    red = np.tile(np.array([255, 0, 0, 255], np.uint8), (3, 3, 1))
    assert np.array_equal(O.color_effect(pyfae.effect("hue_shift", offset=120.0), red)[1, 1],
                          [0, 255, 0, 255])
    # the vignette leaves the centre of an odd-sized image as it is and
    # darkens a far corner fully at radius 0, softness 0.1
    v = O.color_effect(pyfae.effect("vignette", intensity=1.0, radius=0.0, softness=0.1), img)
    assert np.array_equal(v[6, 9], img[6, 9]) and np.array_equal(v[0, 0, :3], [0, 0, 0])
    assert np.array_equal(v[..., 3], img[..., 3])
    # film grain is a value in [0, 1), the same at every call, different per seed
    g = pyfae.effect("film_grain", intensity=1.0, blend_mode=0)
    a = O.color_effect(g, img)
    assert np.array_equal(a, O.color_effect(g, img)) and np.all(a[..., 0] == a[..., 1])
    g2 = pyfae.effect("film_grain", intensity=1.0, blend_mode=0, seed=3.0)
    assert not np.array_equal(a, O.color_effect(g2, img))
    assert 100 < a[..., 0].mean() < 155  # uniform-ish grain


def test_sobel_finds_a_vertical_edge():
    img = np.zeros((8, 16, 4), np.uint8)
    img[:, 8:, :3] = 255
    img[..., 3] = 255
    out = O.color_effect(pyfae.effect("sobel_edges", blend_mode=0, preserve_color=0.0, glow=0.0),
                         img)
    # edge columns glow white (gain 1 * smoothstep = 1), flat regions keep nothing
    assert np.all(out[:, 7:9, :3] == 255) and np.all(out[:, :6, :3] == 0)


# ------------------------------------------------------------ config lists

def cfg_with(density=(), color=()):
    t = json.load(open("tests/golden/ref_simple_linear.fflame"))
    t["density_effects"] = list(density)
    t["color_effects"] = list(color)
    return json.dumps(t)


def test_config_effect_lists_load():
    sc = pyfae.load(cfg_with(
        [{"effect_type": "density_blur", "enabled": True, "params": {"radius": 5.0}},
         {"effect_type": "sharpen", "enabled": False, "params": {}}],
        [{"effect_type": "vignette", "enabled": True, "params": {"blend_mode": 2.0, "x": 9}},
         {"effect_type": "hue_shift", "enabled": True}]))
    fx = sc.effects
    assert (fx.n_density, fx.n_color, fx.pixel_scale) == (1, 2, 1.0)
    assert fx.density[0].type == 1 and list(fx.density[0].params[:3]) == [5.0, 0.25, 0.5]
    assert fx.color[0].type == 16 and list(fx.color[0].params[:4]) == [0.5, 0.5, 0.5, 2.0]
    assert fx.color[1].type == 19 and list(fx.color[1].params[:3]) == [0.0, 1.0, 0.0]
    assert pyfae.load(cfg_with()).effects.n_density == 0


def test_config_effects_this_engine_lacks_are_refused():
    for d, c in (([{"effect_type": "bloom", "enabled": True}], []),
                 ([], [{"effect_type": "my_rhai_effect", "enabled": True}]),
                 ([{"effect_type": "vignette", "enabled": True}], [])):
        with pytest.raises(pyfae.EngineError) as e:
            pyfae.load(cfg_with(d, c))
        assert e.value.status == -10
    # a disabled one is dropped, whatever it is
    assert pyfae.load(cfg_with([{"effect_type": "bloom", "enabled": False}])).effects.n_density == 0


def test_effect_builders():
    fx = pyfae.effects(density=["sharpen", ("density_blur", {"radius": 2})],
                       color=[{"effect_type": "sobel_edges", "enabled": True,
                               "params": {"glow": 0.1}}], pixel_scale=2.0)
    assert (fx.n_density, fx.n_color, fx.pixel_scale) == (2, 1, 2.0)
    assert fx.density[1].params[0] == 2.0 and abs(fx.color[0].params[2] - 0.1) < 1e-7
    with pytest.raises(ValueError):
        pyfae.effects(density=["vignette"])
    with pytest.raises(ValueError):
        pyfae.effect("sharpen", sigma=1)


def test_effect_lists_round_trip_through_scene_to_config():
    text = cfg_with([{"effect_type": "sharpen", "enabled": True, "params": {"amount": 0.3}}],
                    [{"effect_type": "film_grain", "enabled": True, "params": {"seed": 9}},
                     {"effect_type": "sobel_edges", "enabled": True, "params": {}}])
    sc = pyfae.load(text)
    again = pyfae.load(pyfae.formats.scene_to_json(sc))
    assert pyfae.formats.scene_summary(again) == pyfae.formats.scene_summary(sc)
    assert again.effects.n_density == 1 and again.effects.n_color == 2
    assert abs(again.effects.density[0].params[0] - 0.3) < 1e-7


def test_generator_effects_closed_forms():
    H, W = 31, 31
    img = (RNG.random((H, W, 4)) * 255).astype(np.uint8)
    for name in ("kaleidoscope", "plasma", "tunnel", "julia", "worley_noise", "simplex_noise",
                 "domain_warp"):
        assert np.array_equal(O.color_effect(pyfae.effect(name, intensity=0.0), img), img), name
    # simplex's distortion mode writes the warped read unblended, alpha too:
    # at intensity 0 that is the image itself
    assert np.array_equal(O.color_effect(pyfae.effect("simplex_noise", mode=1.0, intensity=0.0),
                                         img), img)
    # its mask mode at intensity 1 (normal blend) scales the colour by a
    # smoothstep in [0, 1]: never brighter
    m = O.color_effect(pyfae.effect("simplex_noise", mode=2.0, intensity=1.0, blend_mode=0.0), img)
    assert np.all(m[..., :3] <= img[..., :3] + 1)
    # the kaleidoscope folds the angle into mirrored segments: a square
    # image mirrored top to bottom about its centre row reads the same
    k = O.color_effect(pyfae.effect("kaleidoscope", segments=5.0), img)
    same = np.all(k[:, :, :3] == k[::-1, :, :3], axis=-1)
    assert same.mean() > 0.97
    # the tunnel keeps the pixel at its centre (no wall there)
    t = O.color_effect(pyfae.effect("tunnel", time=3.0, distortion=0.5), img)
    assert np.array_equal(t[15, 15], img[15, 15])
    # plasma: a sum of sines, against float64 (intensity 1, normal blend)
    pl = O.color_effect(pyfae.effect("plasma", intensity=1.0, blend_mode=0.0, time=2.0), img)
    y, x = np.mgrid[0:H, 0:W]
    u, v = (x + 0.5) / W, (y + 0.5) / H
    X, Y = u * 3.0, v * 3.0
    d = np.deg2rad(225.0)
    t = 2.0 + X * np.cos(d) + Y * np.sin(d)
    w = np.sin(X + t) + np.sin((Y + t) * 0.5) + np.sin((X + Y + t) * 0.5)
    cx, cy = X + 0.5 * np.sin(t * 0.25), Y + 0.5 * np.cos(t * 0.5)
    w = (w + np.sin(np.sqrt(cx * cx + cy * cy + 1.0) + t)) * 0.25
    want = np.stack([np.sin(w * np.pi + ph) * 0.5 + 0.5 for ph in (0.0, 2.094, 4.189)], -1)
    assert np.abs(pl[..., :3].astype(float) - np.round(want * 255)).max() <= 1


def julia_counts(W, H, max_iter, cx=0.0, cy=0.0, zoom=1.0, c=None):
    """Escape counts, float32, in julia.rhai's operation order."""
    f = np.float32
    y, x = np.mgrid[0:H, 0:W].astype(np.float32)
    u, v = (x + f(0.5)) / f(W), (y + f(0.5)) / f(H)
    aspect = f(W) / f(H)
    px = (u - f(0.5)) * f(2) * aspect / f(zoom) + f(cx)
    py = (v - f(0.5)) * f(2) / f(zoom) + f(cy)
    if c is None:  # Mandelbrot
        zr, zi, kr, ki = np.zeros_like(px), np.zeros_like(py), px, py
    else:
        zr, zi = px.copy(), py.copy()
        kr, ki = np.full_like(px, f(c[0])), np.full_like(py, f(c[1]))
    it = np.zeros(px.shape, np.int32)
    x2, y2 = zr * zr, zi * zi
    for _ in range(max_iter):
        live = ~(x2 + y2 > f(4))
        nr, ni = x2 - y2 + kr, f(2) * zr * zi + ki
        zr, zi = np.where(live, nr, zr), np.where(live, ni, zi)
        x2, y2 = zr * zr, zi * zi
        it += live
    return it


def test_julia_escape_counts():
    H, W = 21, 29
    img = (RNG.random((H, W, 4)) * 255).astype(np.uint8)
    for mode, c in ((0.0, None), (1.0, (-0.7, 0.27))):
        fx = pyfae.effect("julia", mode=mode, color_mode=2.0, max_iter=40.0, zoom=1.3,
                          center_x=-0.3, blend_mode=0.0, intensity=1.0)
        out = O.color_effect(fx, img)
        it = julia_counts(W, H, 40, cx=-0.3, zoom=1.3, c=c)
        f = np.where(it >= 40, 0.0, it / np.float32(40)).astype(np.float32)
        want = np.round(np.clip(img[..., :3] / np.float32(255) * f[..., None], 0, 1) * 255)
        assert np.abs(out[..., :3].astype(int) - want).max() <= 1
        assert 0 < (it >= 40).mean() < 1  # inside and outside points both present
    # escape colouring: the set itself is black
    fx = pyfae.effect("julia", mode=0.0, color_mode=0.0, max_iter=30.0, blend_mode=0.0)
    out = O.color_effect(fx, img)
    inside = julia_counts(W, H, 30) >= 30
    assert np.all(out[inside][:, :3] == 0)


def test_worley_grayscale_is_a_distance():
    H, W = 24, 40
    img = (RNG.random((H, W, 4)) * 255).astype(np.uint8)
    fx = pyfae.effect("worley_noise", color_mode=0.0, mode=0.0, blend_mode=0.0, intensity=1.0,
                      time=3.0)
    out = O.color_effect(fx, img)
    assert np.all(out[..., 0] == out[..., 1]) and np.all(out[..., 1] == out[..., 2])
    # the nearest of the 3x3 cells' points: within a cell diagonal
    assert out[..., 0].max() <= 255 and out[..., 0].min() >= 0 and out[..., 0].std() > 10
    assert np.array_equal(out, O.color_effect(fx, img))


def test_effect_sincos_is_sin_and_cos():
    """The colour effects' sin/cos (Cody-Waite + Cephes; an f64 reduction
    past 32768) against the mathematical functions: a few ulp."""
    import math
    import ctypes as C
    xs = np.concatenate([np.linspace(-40.0, 40.0, 20001), np.linspace(-32768.0, 32768.0, 20001),
                         RNG.uniform(-1e4, 1e4, 5000), RNG.uniform(-2e6, 2e6, 5000),
                         RNG.uniform(-1e9, 1e9, 2000)]).astype(np.float32)
    s, c = C.c_float(), C.c_float()
    worst = 0.0
    for x in xs:
        O.lib().orc_fx_sincos(float(x), C.byref(s), C.byref(c))
        for got, want in ((s.value, math.sin(float(x))), (c.value, math.cos(float(x)))):
            ulp = np.spacing(np.float32(max(abs(want), 1e-30)))
            worst = max(worst, abs(got - want) / ulp if abs(want) > 1e-3 else abs(got - want) / 6e-8)
    assert worst <= 4.0, worst
    # quadrant bookkeeping: sin(pi/2 * k) signs
    for k, (ws, wc) in enumerate([(0, 1), (1, 0), (0, -1), (-1, 0)]):
        O.lib().orc_fx_sincos(float(np.float32(math.pi / 2 * k)), C.byref(s), C.byref(c))
        assert abs(s.value - ws) < 1e-6 and abs(c.value - wc) < 1e-6


def test_bloom_closed_forms():
    H, W = 31, 45
    black = np.zeros((H, W, 4), np.uint8)
    black[..., 3] = 200
    assert np.array_equal(O.color_effect(pyfae.effect("bloom"), black), black)
    img = (RNG.random((H, W, 4)) * 255).astype(np.uint8)
    assert np.array_equal(O.color_effect(pyfae.effect("bloom", intensity=0.0), img), img)
    # below threshold - knee nothing glows
    grey = np.full((H, W, 4), 60, np.uint8)
    assert np.array_equal(O.color_effect(pyfae.effect("bloom"), grey), grey)
    # a bright 3x3 patch glows around itself, falling off with distance,
    # symmetric left to right, never darkening anything
    dot = np.zeros((H, W, 4), np.uint8)
    dot[..., 3] = 255
    dot[14:17, 21:24, :3] = 255
    out = O.color_effect(pyfae.effect("bloom", intensity=4.0, spread=1.0), dot).astype(int)
    row = out[15, :, 0]
    assert np.all(out >= dot) and row[16] > row[12] > row[8] > 0
    assert np.abs(out[:, :22] - out[:, 44:22:-1]).max() <= 1
    assert out[15, 22, 0] == 255


def test_distance_field_is_the_distance_to_the_nearest_seed():
    """The jump flood against brute force: each unseeded pixel's nearest
    seed distance (ties aside) from the glow it shades by."""
    H, W = 23, 37
    img = np.zeros((H, W, 4), np.uint8)
    img[..., 3] = 255
    seeds = [(3, 4), (18, 30), (11, 20), (20, 2)]
    for y, x in seeds:
        img[y, x, :3] = 200
    inten = 1.0
    out = O.color_effect(pyfae.effect("distance_field", intensity=inten, spread=0.5), img)
    yy, xx = np.mgrid[0:H, 0:W]
    d = np.min([np.hypot(xx - x, yy - y) for y, x in seeds], axis=0)
    t = np.clip(d / (0.5 * max(W, H)), 0, 1)
    want = np.where(d > 0.5, 200 / 255 * np.exp(-4 * t) * inten, 200 / 255)
    got = out[..., 0] / 255.0
    assert np.abs(got - np.clip(want, 0, 1)).max() <= 1.5 / 255
    # no seed at all: the picture as it was
    dark = np.zeros((H, W, 4), np.uint8)
    assert np.array_equal(O.color_effect(pyfae.effect("distance_field"), dark), dark)


def test_trails_on_a_still_is_the_picture():
    """trails keeps max(c, history * fade); a still render starts from a
    zero history (src/renderer/effect_chain.rs:1746 'a still has no
    trail'), so c + (trail - c) * intensity is c."""
    img = (RNG.random((9, 14, 4)) * 255).astype(np.uint8)
    for inten in (0.0, 0.4, 1.0):
        assert np.array_equal(O.color_effect(pyfae.effect("trails", intensity=inten), img), img)


def test_effect_defaults_and_bad_types():
    import ctypes as C
    from pyfae import _abi
    lib = pyfae.library()
    fx = _abi.Effect()
    assert lib.ffm_effect_defaults(0, C.byref(fx)) != 0
    assert lib.ffm_effect_defaults(99, C.byref(fx)) != 0
    assert lib.ffm_effect_defaults(1, None) != 0
    # every registry effect has defaults, enabled, type echoed
    for name, (t, _, names) in _abi.EFFECTS.items():
        assert lib.ffm_effect_defaults(t, C.byref(fx)) == 0, name
        assert fx.type == t and fx.enabled == 1
        assert all(fx.params[i] == 0.0 for i in range(len(names), _abi.MAX_EFFECT_PARAMS))
