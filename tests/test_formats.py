"""CPU tests of the config loader (ffm_scene_parse / ffm_scene_job through the
C ABI; no kernel runs): FractalConfig JSON (.fflame) with its serde defaults
and version migrations (src/config/fractal_config.rs:178-700,1746-2050;
src/scene/transforms.rs:1307-1560,2567-2690), Apophysis / JWildfire XML
(src/flame_xml.rs:41-1067), and the palette lookup-table transform of the
oracle (src/scene/palette.rs:607-800).

Expected values come from the reference's documented mappings and from the
reference's own golden config (tests/golden/ref_simple_linear.fflame); the
XML documents here are written for these tests."""
import ctypes as C
import json
import math
import os

import numpy as np
import pytest

import pyfae
from pyfae import _abi
from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
F32 = np.float32


def f32(x):
    return float(F32(x))


def load_err(config, index=0):
    with pytest.raises(pyfae.EngineError) as e:
        pyfae.load(config, index)
    return e.value


def minimal(**top):
    cfg = {"flame": {"name": "t", "transforms": [
        {"a": 0.5, "b": 0.0, "c": 0.0, "d": 0.5, "e": 0.0, "f": 0.0, "weight": 1.0,
         "variations": {"linear": 1.0}, "color": 0.25}]}}
    cfg.update(top)
    return cfg


def vars_of(sc):
    fd = sc.flame
    return [_abi.VARIATION_NAMES[fd.var_id[k]] for k in range(fd.n_vars)]


# ------------------------------------------------------------------ JSON

def test_golden_config_loads_as_the_hand_built_one():
    from test_oracle import _simple_linear_job
    fd, stops, job = _simple_linear_job()
    sc = pyfae.load(os.path.join(GOLDEN, "ref_simple_linear.fflame"))
    assert sc.name == b"Simple Linear"
    assert bytes(sc.flame)[:C.sizeof(_abi.FlameDesc)] == bytes(fd)
    np.testing.assert_array_equal(pyfae.scene_stops(sc), stops)
    j = pyfae.scene_job(sc, 800, 600, iterations_per_thread=256)
    for k in ("walkers", "ipt", "burn_in", "seed_base", "seed_stride", "target_samples",
              "batch_size", "persistent_orbits"):
        assert getattr(j, k) == getattr(job, k), k
    for k, _ in _abi.TonemapParams._fields_:
        if k.startswith("curve_"):
            continue  # the file spells out the linear curve's two points
        a, b = getattr(j.tonemap, k), getattr(job.tonemap, k)
        assert (list(a) == list(b)) if hasattr(a, "__len__") else a == b, k
    from oracle import oracle as O
    assert O.curve_lut(j.tonemap).tobytes() == O.curve_lut(job.tonemap).tobytes()
    assert (j.view.width, j.view.height, j.view.zoom) == (800, 600, 1.0)
    # a v0 file: levels_gamma keeps the 0.5 default
    assert sc.tonemap.levels_gamma == 0.5 and sc.deterministic_rng == 1
    assert sc.render_3d == 0 and sc.palette_size == 256


def test_serde_defaults_of_a_minimal_config():
    sc = pyfae.load(minimal())
    tp = sc.tonemap
    # src/config/defaults.rs:19-102
    assert (tp.tonemap_mode, tp.highlight_mode) == (1, 0)
    assert (tp.exposure, tp.gamma, tp.gamma_threshold, tp.brightness) == (1.0, 4.0, 5.0, 4.0)
    assert (tp.vibrancy, tp.white_level, tp.saturation, tp.hue_shift) == (1.0, 200.0, 1.0, 0.0)
    assert (tp.alpha_blend_low, tp.alpha_blend_high) == (f32(0.3), f32(0.8))
    assert (tp.levels_enabled, tp.levels_low, tp.levels_high, tp.levels_gamma) == (1, 0, 10, 0.5)
    assert sc.max_iterations == 1_000_000_000 and sc.deterministic_rng == 0
    assert (sc.zoom, sc.pan_x, sc.pan_y, sc.rotation) == (1.0, 0.0, 0.0, 0.0)
    assert (sc.image_width, sc.image_height) == (1920, 1080)
    # default_palette: Palette::fire (src/scene/palette.rs:148-175)
    np.testing.assert_array_equal(pyfae.scene_stops(sc), F32(
        [[0, 0, 0, 0], [0.25, 0.5, 0, 0], [0.5, 1, 0.3, 0], [0.75, 1, 0.8, 0], [1, 1, 1, 0.8]]))
    x = sc.flame.xforms[0]
    # Transform serde fallbacks: color_speed 0, opacity 1, no post
    assert (x.color, x.color_speed, x.opacity, x.post_enabled) == (0.25, 0.0, 1.0, 0)
    j = pyfae.scene_job(sc, 64, 48)
    assert j.ipt == 1024 and j.target_samples == 1_000_000_000 and j.splat_mode == 2
    assert j.seed_stride == 2654435761


def test_version_migrations():
    # saved as v3: an absent levels_gamma was 1.0 (the v3 -> v4 arm)
    assert pyfae.load(minimal(version=3)).tonemap.levels_gamma == 1.0
    assert pyfae.load(minimal(version=4)).tonemap.levels_gamma == 0.5
    assert pyfae.load(minimal(version=7, levels_gamma=0.7)).tonemap.levels_gamma == f32(0.7)
    # no version, but a top-level render mode with none on the flame: a v3 file
    assert pyfae.load(minimal(render_mode="2d")).tonemap.levels_gamma == 1.0
    # v0-v2: the render mode is lifted from the flame
    cfg = minimal()
    cfg["flame"]["render_mode"] = "ThreeD"
    cfg["flame"]["preserve_z"] = False
    assert pyfae.load(cfg).render_3d == 1
    cfg["flame"]["render_mode"] = "TwoD"
    assert pyfae.load(cfg).render_3d == 0
    e = load_err(minimal(min_reader=8))
    assert e.status == _abi.ERR_PARSE and "newer" in str(e)


def test_3d_mode_needs_the_flag_and_a_flat_camera():
    sc = pyfae.load(minimal(version=7, render_mode="3d", preserve_z=False))
    assert sc.render_3d == 1
    with pytest.raises(pyfae.EngineError) as e:
        pyfae.scene_job(sc, 64, 64)
    assert e.value.status == _abi.ERR_UNSUPPORTED
    assert pyfae.scene_job(sc, 64, 64, as_2d=True).view.width == 64
    for key in ("camera_rotation_x", "perspective_strength", "dof_blur_strength"):
        e = load_err(minimal(version=7, render_mode="3d", preserve_z=False, **{key: 0.5}))
        assert e.status == _abi.ERR_UNSUPPORTED
    # preserve_z defaults to true on load: z would feed the 3D variations
    assert load_err(minimal(version=7, render_mode="3d")).status == _abi.ERR_UNSUPPORTED
    # the same camera fields are inert in 2D
    assert pyfae.load(minimal(version=7, camera_rotation_x=0.5)).render_3d == 0


def test_variation_local_index_order():
    t = lambda v, **kw: dict({"a": 1, "b": 0, "c": 0, "d": 1, "e": 0, "f": 0, "weight": 1,
                              "color": 0, "variations": v}, **kw)
    cfg = {"flame": {"transforms": [
        # registry order (linear 0 < spherical 2 < julian 24)
        t({"julian": 0.5, "spherical": 0.25, "linear": 0.25}),
        # variation_order first, then the rest in registry order
        t({"swirl": 1.0, "curl": 0.5, "polar": 0.0}, variation_order=["curl"]),
    ], "final_transforms": [t({"bubble": 1.0, "linear": 1.0})]}}
    sc = pyfae.load(cfg)
    assert vars_of(sc) == ["linear", "spherical", "julian", "curl", "swirl", "polar", "bubble"]
    x1 = sc.flame.xforms[1]
    assert [x1.var_weight[k] for k in range(7)] == [0, 0, 0, 0.5, 1.0, 0.0, 0]
    # params default from the registry: julian power 2 dist 1, curl c1 1 c2 0
    assert list(sc.flame.xforms[0].var_param[2][:2]) == [2.0, 1.0]
    assert list(x1.var_param[3][:2]) == [1.0, 0.0]
    cfg["flame"]["transforms"][0]["variation_params"] = {"julian.power": 5, "julian.dist": -1}
    sc = pyfae.load(cfg)
    assert list(sc.flame.xforms[0].var_param[2][:2]) == [5.0, -1.0]


def test_legacy_variation_array_and_rgb_colour():
    cfg = minimal()
    t = cfg["flame"]["transforms"][0]
    t["variations"] = [0.5, 0.0, 0.25] + [0.0] * 21 + [0.125]  # linear, spherical, julian
    t["color"] = [0.2, 0.4, 0.9]
    sc = pyfae.load(cfg)
    assert vars_of(sc) == ["linear", "spherical", "julian"]
    assert [sc.flame.xforms[0].var_weight[k] for k in range(3)] == [0.5, 0.25, 0.125]
    r, g, b = F32(0.2), F32(0.4), F32(0.9)
    assert sc.flame.xforms[0].color == float((r + g + b) / F32(3))
    # a weight on an index this engine lacks is refused, not dropped
    t["variations"] = [0.5, 0.0, 0.0, 0.0, 0.0, 0.0, 0.7]
    assert load_err(cfg).status == _abi.ERR_UNSUPPORTED


def test_finals_legacy_singular_and_attachments():
    cfg = minimal()
    tr = cfg["flame"]["transforms"]
    tr.append(dict(tr[0], a=0.3))
    fin = {"a": 2, "b": 0, "c": 0, "d": 2, "e": 0, "f": 0, "weight": 1, "color": 0,
           "variations": {"spherical": 1.0}}
    cfg["flame"]["final_transform"] = fin
    sc = pyfae.load(cfg)
    fd = sc.flame
    assert fd.n_finals == 1 and fd.per_transform_finals == 1
    assert [(fd.xforms[i].n_final_attach, fd.xforms[i].final_attach[0]) for i in range(2)] == \
        [(1, 0), (1, 0)]
    del cfg["flame"]["final_transform"]
    cfg["flame"]["final_transforms"] = [fin, dict(fin, a=3)]
    tr[0]["final_attachments"] = [1, 0]
    sc = pyfae.load(cfg)
    fd = sc.flame
    assert fd.n_finals == 2
    assert list(fd.xforms[0].final_attach[:fd.xforms[0].n_final_attach]) == [1, 0]
    assert fd.xforms[1].n_final_attach == 0
    tr[0]["final_attachments"] = [2]
    assert load_err(cfg).status == _abi.ERR_PARSE


def test_post_affine_view_and_tonemap_fields():
    cfg = minimal(zoom=2.5, pan_x=0.1, pan_y=-0.3, rotation=0.5, exposure=1.5, gamma=2.0,
                  gamma_threshold=7, brightness=3, vibrancy=0.5, white_level=150,
                  saturation=1.2, hue_shift=30, background_color=[0.1, 0.2, 0.3],
                  tonemap_mode="linear", highlight_mode="max_norm", levels_enabled=False,
                  filter_radius=1.5, filter_blur_edges=0.25, max_iterations=123456789,
                  reset_orbits=True, deterministic_rng=True, image_size=[640, 480],
                  density_scale=2.0)
    t = cfg["flame"]["transforms"][0]
    t.update(post_affine_enabled=True, post_a=0.9, post_b=0.1, post_c=-0.1, post_d=0.8,
             post_e=0.2, post_f=-0.2, color_speed=0.3, opacity=0.7)
    sc = pyfae.load(cfg)
    x = sc.flame.xforms[0]
    assert x.post_enabled == 1
    assert [x.post_a, x.post_b, x.post_c, x.post_d, x.post_e, x.post_f] == \
        [f32(v) for v in (0.9, 0.1, -0.1, 0.8, 0.2, -0.2)]
    assert (x.color_speed, x.opacity) == (f32(0.3), f32(0.7))
    assert (sc.zoom, sc.rotation) == (2.5, 0.5) and (sc.pan_x, sc.pan_y) == (0.1, -0.3)
    tp = sc.tonemap
    assert (tp.tonemap_mode, tp.highlight_mode, tp.levels_enabled) == (0, 1, 0)
    assert (tp.exposure, tp.gamma, tp.gamma_threshold, tp.density_scale) == (1.5, 2.0, 7.0, 2.0)
    assert list(tp.background) == [f32(0.1), f32(0.2), f32(0.3)]
    j = pyfae.scene_job(sc, 320, 200, iterations=1000)
    assert (j.filter_radius, j.filter_blur_edges) == (1.5, 0.25)
    assert j.target_samples == 1000 and j.persistent_orbits == 0 and j.seed_base == 12345
    assert (j.view.pan_x, j.view.pan_y) == (f32(0.1), f32(-0.3))
    assert (sc.image_width, sc.image_height) == (640, 480)
    assert pyfae.scene_job(sc, 320, 200).target_samples == 123456789


def test_palette_forms():
    sc = pyfae.load(minimal(palette={"name": "p", "indexed_colors": "FF0000,00FF80,0000FF"}))
    np.testing.assert_array_equal(pyfae.scene_stops(sc), F32(
        [[0, 1, 0, 0], [1 / F32(255), 0, 1, F32(128) / F32(255)], [2 / F32(255), 0, 0, 1]]))
    sc = pyfae.load(minimal(palette=None))  # null: the default
    assert sc.n_stops == 5
    assert load_err(minimal(palette={"stops": []})).status == _abi.ERR_PARSE
    assert load_err(minimal(palette={"name": "p", "indexed_colors": "FF00"})).status == \
        _abi.ERR_PARSE
    sc = pyfae.load(minimal(palette_rotation=0.25, palette_squeeze=2, palette_reverse=True,
                            palette_size=1024))
    xf = sc.palette_xf
    assert (xf.squeeze_mode, xf.squeeze_factor, xf.rotation, xf.reverse) == (0, 2.0, 0.25, 1)
    assert sc.palette_size == 1024
    sc = pyfae.load(minimal(palette_squeeze_mode="geometric", palette_squeeze_falloff=0.4))
    assert (sc.palette_xf.squeeze_mode, sc.palette_xf.squeeze_factor) == (1, f32(0.4))


@pytest.mark.parametrize("extra,what", [
    ({"render_mode": "escape", "version": 7}, "escape"),
    ({"color_mode": "speed"}, "color mode"),
    ({"highlight_mode": "agx"}, "highlight"),
    ({"density_effects": [{"effect_type": "bloom", "enabled": True}]}, "density effect 'bloom'"),
    ({"color_effects": [{"effect_type": "my_rhai_effect", "enabled": True}]}, "my_rhai_effect"),
    ({"auto_exposure": True}, "auto exposure"),
    ({"use_curve": True, "tonemap_curve": {"points": [{"x": 0, "y": 0}, {"x": 0.7, "y": 0.7},
                                                      {"x": 0.5, "y": 1}]}}, "tone curve"),
])
def test_unsupported_config_features_are_refused(extra, what):
    e = load_err(minimal(**extra))
    assert e.status == _abi.ERR_UNSUPPORTED and what in str(e)


def test_unsupported_flame_features_are_refused():
    cfg = minimal()
    cfg["flame"]["transforms"][0]["variations"] = {"linear": 1.0, "spiral": 0.5}
    e = load_err(cfg)
    assert e.status == _abi.ERR_UNSUPPORTED and "spiral" in str(e)
    cfg["flame"]["transforms"][0]["variations"] = {"linear": 1.0, "spiral": 0.0}
    assert vars_of(pyfae.load(cfg)) == ["linear"]  # a weight of 0 draws nothing
    cfg = minimal()
    tr = cfg["flame"]["transforms"]
    tr.append(dict(tr[0], a=0.3))
    tr.append(dict(tr[0], a=0.2))
    # Flame::has_xaos: some weight differs from 1 by more than 1e-6; short
    # rows read 1 (get_xaos)
    cfg["flame"]["xaos"] = [[1.0, 1.0000001], [1.0]]
    assert pyfae.load(cfg).flame.has_xaos == 0
    cfg["flame"]["xaos"] = [[0.5, 0.0], [1.0], [2.0, 3.0, 0.25, 7.0]]
    fd = pyfae.load(cfg).flame
    assert fd.has_xaos == 1
    got = [[fd.xaos[i][j] for j in range(3)] for i in range(3)]
    assert got == [[0.5, 0.0, 1.0], [1.0, 1.0, 1.0], [2.0, 3.0, 0.25]]
    cfg["flame"]["transforms"] = [tr[0]] * 33
    cfg["flame"]["xaos"] = [[0.5]]
    assert load_err(cfg).status == _abi.ERR_TOO_LARGE
    cfg = minimal()
    cfg["flame"]["transforms"][0]["linked_attachments"] = [0]
    assert "linked" in str(load_err(cfg))
    cfg = minimal()
    cfg["flame"]["subflames"] = [{"name": "s", "transforms": []}]
    assert "subflames" in str(load_err(cfg))


def test_json_arrays_syntax_errors_and_required_fields():
    two = [minimal(), minimal(zoom=3.0)]
    assert pyfae.scene_count(two) == 2
    assert pyfae.load(two, 1).zoom == 3.0
    assert load_err(two, 2).status == _abi.ERR_INVALID_ARGUMENT
    assert load_err('{"flame": {').status == _abi.ERR_PARSE
    assert load_err('{"flame": {}} x').status == _abi.ERR_PARSE
    assert load_err("42").status == _abi.ERR_PARSE
    cfg = minimal()
    del cfg["flame"]["transforms"][0]["weight"]
    e = load_err(cfg)
    assert e.status == _abi.ERR_PARSE and "weight" in str(e)
    cfg = minimal()
    cfg["flame"]["transforms"][0]["a"] = "1"
    assert load_err(cfg).status == _abi.ERR_PARSE
    # strings with escapes and non-ASCII names parse
    cfg = minimal()
    cfg["flame"]["name"] = 'café "\\n" \U0001F525'
    assert pyfae.load(json.dumps(cfg)).name.decode() == cfg["flame"]["name"]


# ------------------------------------------------------------------- XML

PAL = "".join("%02X%02X%02X" % (i, 255 - i, (7 * i) % 256) for i in range(256))

FLAME = f"""<?xml version="1.0"?>
<!-- written for these tests -->
<flames>
<flame name="two &amp; one" version="Apophysis 2.09" size="640 480" center="0.25 -0.5"
       scale="300" rotate="30" oversample="1" filter="0.75" quality="50"
       background="51 102 153" brightness="3.5" gamma="2.5" gamma_threshold="0.01"
       vibrancy="0.8" estimator_radius="9">
  <xform weight="0.6" color="0.1 0" symmetry="0.25" linear="0.5" julian="0.3"
         spherical="0.2" julian_power="3" julian_dist="-1" var_color="1"
         coefs="0.5 0.1 -0.2 0.6 0.3 -0.4" opacity="0.9"/>
  <xform weight="0.4" color="0.9" coefs="1 0 0 1 0 0" post="0.9 0.05 -0.05 0.8 0.1 0.2"
         swirl="1" chaos="1 1"/>
  <finalxform color="0" symmetry="1" coefs="1.5 0 0 1.5 0 0" spherical="1" weight="0.5"/>
  <finalxform color="0" symmetry="1" coefs="1 0 0 1 0.1 0" linear="1"/>
  <palette count="256" format="RGB">
{PAL[:700]}
{PAL[700:]}
  </palette>
</flame>
<flame name="second" scale="200"><xform weight="1" color="0" coefs="1 0 0 1 0 0" linear="1"/></flame>
</flames>
"""


def test_xml_flame_mapping():
    assert pyfae.scene_count(FLAME) == 2
    sc = pyfae.load(FLAME)
    assert sc.name == b"two &amp; one"  # attribute values as written
    assert sc.render_3d == 1 and sc.deterministic_rng == 0
    fd = sc.flame
    assert (fd.n_xforms, fd.n_finals) == (2, 2)
    # attribute order gives the local order; then the second transform and the finals
    assert vars_of(sc) == ["linear", "julian", "spherical", "swirl"]
    x0, x1 = fd.xforms[0], fd.xforms[1]
    # coefs are written "a c b d e f"
    assert [x0.a, x0.b, x0.c, x0.d, x0.e, x0.f] == [f32(v) for v in (0.5, -0.2, 0.1, 0.6, 0.3, -0.4)]
    assert (x0.weight, x0.color, x0.color_speed, x0.opacity) == (f32(0.6), f32(0.1), 0.25, f32(0.9))
    assert [x0.var_weight[k] for k in range(4)] == [0.5, f32(0.3), f32(0.2), 0.0]
    assert list(x0.var_param[1][:2]) == [3.0, -1.0]
    assert x0.post_enabled == 0 and x1.post_enabled == 1
    assert [x1.post_a, x1.post_b, x1.post_c, x1.post_d, x1.post_e, x1.post_f] == \
        [f32(v) for v in (0.9, -0.05, 0.05, 0.8, 0.1, 0.2)]
    # Transform::new for what the XML leaves out: colour speed 0, opacity 1
    assert (x1.color_speed, x1.opacity) == (0.0, 1.0)
    # every <finalxform>, in order, chained after every transform
    assert fd.per_transform_finals == 1
    for x in (x0, x1):
        assert list(x.final_attach[:x.n_final_attach]) == [0, 1]
    assert fd.finals[0].a == 1.5 and fd.finals[1].e == f32(0.1)
    # the view: zoom = scale / 200, pan = centre, rotation in radians
    assert sc.zoom == f32(F32(300) / F32(200))
    assert (sc.pan_x, sc.pan_y) == (0.25, -0.5)
    assert sc.rotation == f32(F32(30) * F32(math.pi) / F32(180))
    assert (sc.image_width, sc.image_height) == (640, 480)
    assert sc.filter_radius == 0.75 and sc.filter_blur_edges == 0.0
    tp = sc.tonemap
    assert (tp.brightness, tp.gamma, tp.vibrancy) == (3.5, 2.5, f32(0.8))
    assert tp.gamma_threshold == f32(F32(0.01) * F32(2000) + F32(50))
    assert list(tp.background) == [f32(F32(v) / F32(255)) for v in (51, 102, 153)]
    assert (tp.exposure, tp.levels_enabled, tp.levels_high, tp.levels_gamma) == (1.0, 1, 10, 0.5)
    st = pyfae.scene_stops(sc)
    assert st.shape == (256, 4)
    np.testing.assert_array_equal(st[:, 0], F32(np.arange(256)) / F32(255))
    np.testing.assert_array_equal(st[7, 1:], F32([7, 248, 49]) / F32(255))
    assert sc.max_iterations == 1_000_000_000
    second = pyfae.load(FLAME, 1)
    assert second.name == b"second" and second.zoom == 1.0
    # no <palette>: the grayscale ramp
    np.testing.assert_array_equal(pyfae.scene_stops(second), F32([[0, 0, 0, 0], [1, 1, 1, 1]]))


def test_xml_chaos_rows_become_xaos():
    doc = """<flame><xform weight="1" color="0" coefs="1 0 0 1 0 0" linear="1" chaos="0 2"/>
      <xform weight="1" color="0" coefs="1 0 0 1 0 0" linear="1"/>
      <xform weight="1" color="0" coefs="1 0 0 1 0 0" linear="1" chaos="1 1 0.5 9"/></flame>"""
    fd = pyfae.load(doc).flame
    assert fd.has_xaos == 1
    got = [[fd.xaos[i][j] for j in range(3)] for i in range(3)]
    # rows without chaos stay 1; short rows are 1 past their end; extra
    # entries past the transform count are dropped
    assert got == [[0.0, 2.0, 1.0], [1.0, 1.0, 1.0], [1.0, 1.0, 0.5]]
    plain = doc.replace('chaos="0 2"', 'chaos="1"').replace('chaos="1 1 0.5 9"', 'chaos="1 1 1"')
    assert pyfae.load(plain).flame.has_xaos == 0


def test_xml_jwf_layers_take_the_first():
    doc = """<jwf-flame name="L" scale="100" cam_zoom="2">
      <layer><xform weight="1" color="0" coefs="1 0 0 1 0 0" linear="1"/></layer>
      <layer><xform weight="1" color="0" coefs="1 0 0 1 0 0" spiral="1"/></layer>
    </jwf-flame>"""
    sc = pyfae.load(doc)
    assert sc.flame.n_xforms == 1 and sc.zoom == 1.0


@pytest.mark.parametrize("change,status", [
    ('linear="1"', None),
    ('linear="1" spiral="0.5"', _abi.ERR_UNSUPPORTED),   # a variation this engine lacks
    ('linear="1" spiral="0"', None),                      # ... at weight 0 draws nothing
    ('linear="1" chaos="0 1"', None),                     # xaos
    ('linear="1" chaos="1 1 1"', None),
    ('linear="1" mod_gamma="0.5" material="2"', None),    # JWildfire colouring, inert
    ('linear="1" linear_fx_priority="1"', _abi.ERR_UNSUPPORTED),
])
def test_xml_xform_attributes(change, status):
    doc = f'<flame><xform weight="1" color="0" coefs="1 0 0 1 0 0" {change}/><xform weight="1" color="0" coefs="1 0 0 1 0 0" linear="1"/></flame>'
    if status is None:
        pyfae.load(doc)
    else:
        assert load_err(doc).status == status


@pytest.mark.parametrize("attrs,status", [
    ('cam_pitch="0.3"', _abi.ERR_UNSUPPORTED),
    ('cam_perspective="0.2"', _abi.ERR_UNSUPPORTED),
    ('preserve_z="1"', _abi.ERR_UNSUPPORTED),
    ('post_symmetry_type="POINT"', _abi.ERR_UNSUPPORTED),
    ('soloxform="0"', _abi.ERR_UNSUPPORTED),
    ('sld_render_enabled="1"', _abi.ERR_UNSUPPORTED),
    ('curves="0 0 1 0.25 0.25 1 0.75 0.75 1 1 1 1 ' + "0 0 1 0.25 0.25 1 0.75 0.75 1 1 1 1 " * 3 + '"', None),
    ('curves="0 0 1 0.25 0.5 1 0.75 0.9 1 1 1 1 ' + "0 0 1 0.25 0.25 1 0.75 0.75 1 1 1 1 " * 3 + '"', None),
    ('preserve_z="0" cam_pitch="0"', None),
])
def test_xml_flame_attributes(attrs, status):
    doc = f'<flame {attrs}><xform weight="1" color="0" coefs="1 0 0 1 0 0" linear="1"/></flame>'
    if status is None:
        pyfae.load(doc)
    else:
        assert load_err(doc).status == status


def test_xml_errors():
    assert load_err("<flames></flames>").status == _abi.ERR_PARSE  # no <flame>
    assert load_err("<flame><xform></flame>").status == _abi.ERR_PARSE  # mismatched end
    assert load_err('<flame a="1" a="2"></flame>').status == _abi.ERR_PARSE
    short = '<flame><xform weight="1" color="0" coefs="1 0 0 1 0 0" linear="1"/><palette>FF00FF</palette></flame>'
    assert load_err(short).status == _abi.ERR_PARSE
    # unparseable numbers fall back to the reference's defaults
    sc = pyfae.load('<flame scale="x" gamma="bad"><xform weight="w" color="0" coefs="1 0 0 1 0 0" linear="1"/></flame>')
    assert sc.zoom == 0.5 and sc.tonemap.gamma == 4.0 and sc.flame.xforms[0].weight == 1.0


# --------------------------------------------------- palette table transform

def xf(mode=0, factor=1.0, log=0.0, rot=0.0, rev=0):
    return _abi.PaletteXf(mode, factor, log, rot, rev)


def test_palette_index_map_closed_forms():
    n = 256
    i = np.arange(n)
    assert (O.palette_index_map(n) == i).all()
    assert (O.palette_index_map(n, xf()) == i).all()
    assert (O.palette_index_map(n, xf(rev=1)) == i[::-1]).all()
    # rotation: a cyclic shift by round(rotation * size)
    assert (O.palette_index_map(n, xf(rot=0.25)) == (i + 64) % n).all()
    assert (O.palette_index_map(n, xf(rot=-0.1)) == (i - 26) % n).all()
    # linear squeeze k: the palette k times over
    assert (O.palette_index_map(n, xf(factor=2.0)) == (2 * i) % n).all()
    assert (O.palette_index_map(n, xf(factor=0.5)) == i // 2).all()
    # geometric r = 1/2: the first half maps onto the whole palette, the next
    # quarter again, ...
    g = O.palette_index_map(n, xf(mode=1, factor=0.5))
    assert (g[:128] == 2 * i[:128]).all() and (g[128:192] == 4 * (i[128:192] - 128)).all()
    # exponential redistribution: monotone, fixes 0, bunches toward the end
    lg = O.palette_index_map(n, xf(log=3.0))
    assert lg[0] == 0 and (np.diff(lg) >= 0).all() and lg[128] < 64
    ln = O.palette_index_map(n, xf(log=-3.0))
    assert (np.diff(ln) >= 0).all() and ln[128] > 192
    # the stages compose in the reference's order: squeeze, log, rotate, reverse
    both = O.palette_index_map(n, xf(factor=2.0, rot=0.25, rev=1))
    assert (both == ((2 * ((i[::-1] + 64) % n)) % n)).all()


# ------------------------------------------- FractalConfig writer and PNG metadata

from pyfae import formats  # noqa: E402


def _roundtrip(config, index=0):
    sc = pyfae.load(config, index)
    text = formats.scene_to_json(sc)
    back = pyfae.load(text)
    assert formats.scene_summary(back) == formats.scene_summary(sc)
    # the same slots in the same order: the descriptor itself is identical
    assert bytes(back.flame) == bytes(sc.flame)
    assert json.loads(text)["version"] == 7
    return sc, back


def test_scene_to_config_round_trips():
    _roundtrip(os.path.join(GOLDEN, "ref_simple_linear.fflame"))
    _roundtrip(FLAME, 0)
    _roundtrip(FLAME, 1)
    t = lambda v, **kw: dict({"a": 0.7, "b": 0.1, "c": -0.1, "d": 0.6, "e": 0.2, "f": -0.3,
                              "weight": 0.5, "color": 0.3, "variations": v}, **kw)
    cfg = {"version": 7, "flame": {"name": "mix", "transforms": [
        t({"julian": 0.6, "pre_blur": 0.2, "linear": 0.4}, variation_params={"julian.power": 3},
          final_attachments=[1]),
        t({"curl": 1.0, "blob": 0.3}, variation_order=["blob", "curl"],
          variation_params={"curl.c1": 0.3, "blob.waves": 4}, post_affine_enabled=True,
          post_a=0.9, post_b=0.1, post_c=0.0, post_d=1.1, post_e=0.0, post_f=0.2),
        t({"spherical": 1.0}, final_attachments=[0, 1])],
        "final_transforms": [t({"linear": 1.0}), t({"bubble": 0.5, "noise": 0.5})],
        "xaos": [[1, 0.5, 0], [2, 1, 1], [0, 1, 3]]},
        "palette_squeeze_mode": "geometric", "palette_squeeze_falloff": 0.4, "palette_reverse": True,
        "palette_rotation": 0.2, "palette_size": 1024, "highlight_mode": "reinhard",
        "tonemap_mode": "linear", "reset_orbits": True, "deterministic_rng": True,
        "pan_x": 0.123456789012345, "zoom": 1.7, "filter_radius": 0.8}
    sc, back = _roundtrip(cfg)
    assert back.flame.has_xaos == 1 and back.palette_xf.squeeze_mode == 1
    assert back.pan_x == 0.123456789012345  # pan stays f64


def test_png_metadata_round_trip():
    from PIL import Image as PILImage
    import io
    sc = pyfae.load(FLAME)
    rgba = np.random.default_rng(3).integers(0, 256, (12, 17, 4), dtype=np.uint8)
    img = pyfae.Image(rgba, {"width": 17, "height": 12, "iterations": 123456, "ms": 45.678,
                             "iterations_per_thread": 256, "dispatches": 3}, scene=sc)
    data = img.png()
    pil = PILImage.open(io.BytesIO(data))
    np.testing.assert_array_equal(np.asarray(pil.convert("RGBA")), rgba)
    assert pil.text["Resolution"] == "17x12" and pil.text["Iterations"] == "123456"
    md = formats.read_png_metadata(data)
    assert (md["width"], md["height"], md["total_iterations"]) == (17, 12, 123456)
    assert md["render_time_ms"] == 45.68 and md["iterations_per_thread"] == 256
    assert md["render_mode"] == "3D" and md["version"] == "0.5.0+mi355x"
    import hashlib
    assert md["config_checksum"] == hashlib.sha256(md["config_json"].encode()).hexdigest()
    # the PNG reloads as its flame
    assert formats.scene_summary(pyfae.load(data)) == formats.scene_summary(sc)
    # a name outside Latin-1 travels in an iTXt chunk
    cfg = json.loads(formats.scene_to_json(sc))
    cfg["flame"]["name"] = "flamme ☃"
    sc2 = pyfae.load(cfg)
    data2 = pyfae.Image(rgba, dict(img.info), scene=sc2).png()
    assert pyfae.load(data2).name.decode() == "flamme ☃"
    bare = img.png(strip_metadata=True)
    assert formats.png_text(bare) == {}
    with pytest.raises(ValueError):
        formats.config_from_png(bare)
