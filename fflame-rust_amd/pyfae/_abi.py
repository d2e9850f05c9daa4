"""ctypes mirror of include/ffm.h (plain C structs, no torch types)."""
import ctypes as C

MAX_XFORMS = 128
MAX_FINALS = 4
MAX_VARS = 16
MAX_VAR_PARAMS = 4
MAX_XAOS = 32
MIN_PALETTE = 256
MAX_PALETTE = 4096
MAX_STOPS = 1024
SCENE_NAME = 128
RENDER_3D_AS_2D = 1
RENDER_TRANSPARENT = 2
RENDER_PREMULTIPLIED = 4
RENDER_SUPERSAMPLE_2X = 8
RENDER_FILL_GPU = 16

OK = 0
ERR_INVALID_ARGUMENT = -1
ERR_OUT_OF_MEMORY = -2
ERR_TOO_LARGE = -3
ERR_CANCELLED = -4
ERR_PIXEL_READ_FAILED = -5
ERR_NO_PALETTE = -6
ERR_DEVICE = -7
ERR_NO_FLAME = -8
ERR_PARSE = -9
ERR_UNSUPPORTED = -10

SPLAT_U32_GLOBAL = 0
SPLAT_U32_XCD = 1
SPLAT_U32_LDS = 2
SPLAT_F32_GLOBAL = 3
SPLAT_F32_LDS = 4

# Reference registry indices (src/variations/defs/mod.rs:442-501).
VARIATIONS = {
    "linear": 0, "sinusoidal": 1, "spherical": 2, "swirl": 3, "horseshoe": 4,
    "polar": 5, "disc": 8, "hyperbolic": 10, "julia": 13, "waves": 15,
    "julian": 24, "blob": 25, "bubble": 27, "noise": 29, "blur": 30, "gaussian_blur": 31,
    "curl": 47, "pre_blur": 54,
}
VARIATION_NAMES = {v: k for k, v in VARIATIONS.items()}
# User parameters per variation, in slot order (julian's slot 2 is derived).
VARIATION_PARAMS = {"julian": ("power", "dist"), "curl": ("c1", "c2"),
                    "blob": ("high", "low", "waves")}
PARAM_DEFAULTS = {"julian": (2.0, 1.0), "curl": (1.0, 0.0), "blob": (1.0, 1.0, 6.0)}


class Xform(C.Structure):
    _fields_ = [
        ("a", C.c_float), ("b", C.c_float), ("c", C.c_float),
        ("d", C.c_float), ("e", C.c_float), ("f", C.c_float),
        ("post_a", C.c_float), ("post_b", C.c_float), ("post_c", C.c_float),
        ("post_d", C.c_float), ("post_e", C.c_float), ("post_f", C.c_float),
        ("post_enabled", C.c_int32),
        ("weight", C.c_float), ("color", C.c_float), ("color_speed", C.c_float),
        ("opacity", C.c_float),
        ("var_weight", C.c_float * MAX_VARS),
        ("var_param", (C.c_float * MAX_VAR_PARAMS) * MAX_VARS),
        ("n_final_attach", C.c_int32),
        ("final_attach", C.c_int32 * MAX_FINALS),
    ]


class FlameDesc(C.Structure):
    _fields_ = [
        ("n_xforms", C.c_int32), ("n_finals", C.c_int32), ("n_vars", C.c_int32),
        ("var_id", C.c_int32 * MAX_VARS),
        ("xforms", Xform * MAX_XFORMS),
        ("finals", Xform * MAX_FINALS),
        ("per_transform_finals", C.c_int32),
        ("has_xaos", C.c_int32),
        ("xaos", (C.c_float * MAX_XAOS) * MAX_XAOS),
    ]


class View(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("zoom", C.c_float),
                ("pan_x", C.c_float), ("pan_y", C.c_float), ("rotation", C.c_float)]


MAX_CURVE_POINTS = 16


class TonemapParams(C.Structure):
    _fields_ = [
        ("tonemap_mode", C.c_int32), ("highlight_mode", C.c_int32),
        ("exposure", C.c_float), ("gamma", C.c_float), ("gamma_threshold", C.c_float),
        ("brightness", C.c_float), ("vibrancy", C.c_float), ("white_level", C.c_float),
        ("density_scale", C.c_float), ("saturation", C.c_float), ("hue_shift", C.c_float),
        ("alpha_blend_low", C.c_float), ("alpha_blend_high", C.c_float),
        ("levels_enabled", C.c_int32),
        ("levels_low", C.c_float), ("levels_high", C.c_float), ("levels_gamma", C.c_float),
        ("background", C.c_float * 3),
        ("transparent_mode", C.c_int32),
        ("use_curve", C.c_int32), ("curve_n", C.c_int32),
        ("curve_x", C.c_float * MAX_CURVE_POINTS), ("curve_y", C.c_float * MAX_CURVE_POINTS),
    ]


class SoftJob(C.Structure):
    """ffm_soft_job: one differentiable soft-splat fit."""
    _fields_ = [("view", View), ("walkers", C.c_uint64), ("ipt", C.c_uint32),
                ("burn_in", C.c_uint32), ("seed", C.c_uint32), ("window", C.c_uint32),
                ("truncate", C.c_uint32), ("density_scale", C.c_float),
                ("walk_f64", C.c_uint32)]


class Job(C.Structure):
    _fields_ = [
        ("view", View),
        ("walkers", C.c_uint64),
        ("ipt", C.c_uint32), ("burn_in", C.c_uint32),
        ("seed_base", C.c_uint32), ("seed_stride", C.c_uint32),
        ("target_samples", C.c_uint64),
        ("batch_size", C.c_uint32),
        ("splat_mode", C.c_int32),
        ("persistent_orbits", C.c_int32),
        ("filter_radius", C.c_float), ("filter_blur_edges", C.c_float),
        ("de_max_radius", C.c_float), ("de_min_radius", C.c_float), ("de_curve", C.c_float),
        ("tonemap", TonemapParams),
    ]


class PaletteXf(C.Structure):
    """ffm_palette_xf: the palette lookup-table transform."""
    _fields_ = [("squeeze_mode", C.c_int32), ("squeeze_factor", C.c_float),
                ("log_strength", C.c_float), ("rotation", C.c_float), ("reverse", C.c_int32)]


MAX_EFFECTS = 16
MAX_EFFECT_PARAMS = 12
# ffm_effect_type: registry name -> (type, chain, parameter names in order)
EFFECTS = {
    "density_blur": (1, "density", ("radius", "threshold", "falloff")),
    "bilateral_blur": (2, "density", ("radius", "sigma_spatial", "sigma_range")),
    "sharpen": (3, "density", ("amount", "radius")),
    "vignette": (16, "color", ("intensity", "radius", "softness", "blend_mode")),
    "film_grain": (17, "color", ("intensity", "seed", "blend_mode")),
    "chromatic_aberration": (18, "color", ("amount", "radial", "intensity", "blend_mode")),
    "hue_shift": (19, "color", ("offset", "intensity", "blend_mode")),
    "sobel_edges": (20, "color", ("intensity", "threshold", "glow", "preserve_color",
                                  "blend_mode")),
    "kaleidoscope": (21, "color", ("segments", "rotation", "zoom", "intensity", "blend_mode",
                                   "square_mode", "edge_offset")),
    "plasma": (22, "color", ("intensity", "scale", "speed", "time", "blend_mode", "direction")),
    "tunnel": (23, "color", ("speed", "rotation_speed", "distortion", "time", "intensity",
                             "blend_mode")),
    "julia": (24, "color", ("mode", "julia_cx", "julia_cy", "max_iter", "zoom", "center_x",
                            "center_y", "color_mode", "intensity", "blend_mode")),
    "worley_noise": (25, "color", ("intensity", "scale", "time", "mode", "color_mode",
                                   "blend_mode", "direction")),
    "simplex_noise": (26, "color", ("intensity", "scale", "octaves", "time", "mode",
                                    "blend_mode", "direction")),
    "domain_warp": (27, "color", ("intensity", "scale", "octaves", "time", "blend_mode",
                                  "direction")),
    "bloom": (28, "color", ("threshold", "softness", "intensity", "spread")),
    "distance_field": (29, "color", ("threshold", "spread", "mode", "intensity", "band_count")),
    "trails": (30, "color", ("persistence", "intensity")),
}
BLEND_MODES = ("Normal", "Add", "Multiply", "Screen", "Overlay", "Soft Light", "Hard Light",
               "Color Dodge", "Color Burn", "Hue", "Saturation", "Color", "Luminosity")


class Effect(C.Structure):
    """ffm_effect: one EffectInstance (type, enabled, params in order)."""
    _fields_ = [("type", C.c_int32), ("enabled", C.c_int32),
                ("params", C.c_float * MAX_EFFECT_PARAMS)]


class Effects(C.Structure):
    """ffm_effects: a config's density and colour chains."""
    _fields_ = [("n_density", C.c_int32), ("n_color", C.c_int32),
                ("density", Effect * MAX_EFFECTS), ("color", Effect * MAX_EFFECTS),
                ("pixel_scale", C.c_float)]


class Scene(C.Structure):
    """ffm_scene: one config (.fflame JSON or .flame XML) as this path reads it."""
    _fields_ = [
        ("flame", FlameDesc),
        ("n_stops", C.c_int32),
        ("stops", (C.c_float * 4) * MAX_STOPS),
        ("palette_size", C.c_int32),
        ("palette_xf", PaletteXf),
        ("zoom", C.c_float), ("rotation", C.c_float),
        ("pan_x", C.c_double), ("pan_y", C.c_double),
        ("filter_radius", C.c_float), ("filter_blur_edges", C.c_float),
        ("max_iterations", C.c_uint64),
        ("reset_orbits", C.c_int32),
        ("deterministic_rng", C.c_int32),
        ("render_3d", C.c_int32),
        ("image_width", C.c_uint32), ("image_height", C.c_uint32),
        ("tonemap", TonemapParams),
        ("name", C.c_char * SCENE_NAME),
        ("effects", Effects),
    ]


class RenderStats(C.Structure):
    _fields_ = [("total_samples", C.c_uint64), ("dispatches", C.c_uint64),
                ("render_time_ms", C.c_double), ("iterate_ms", C.c_double),
                ("walker_kernel_ms", C.c_double), ("walker_launches", C.c_uint64),
                ("tile_accum_ms", C.c_double)]


def tonemap_defaults():
    """The reference's tonemap defaults (src/config/defaults.rs:19-47,100-101)."""
    t = TonemapParams()
    t.tonemap_mode = 1
    t.highlight_mode = 0
    t.exposure = 1.0
    t.gamma = 4.0
    t.gamma_threshold = 5.0
    t.brightness = 4.0
    t.vibrancy = 1.0
    t.white_level = 200.0
    t.density_scale = 1.0
    t.saturation = 1.0
    t.hue_shift = 0.0
    t.alpha_blend_low = 0.3
    t.alpha_blend_high = 0.8
    t.levels_enabled = 1
    t.use_curve = 1  # the linear curve (curve_n 0), on by default
    t.levels_low = 0.0
    t.levels_high = 10.0
    t.levels_gamma = 0.5
    return t


# Exported symbols of libffm.so: (name, restype, argtypes).
_vp = C.c_void_p
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_fp = C.POINTER(C.c_float)
_u8p = C.POINTER(C.c_uint8)
SIGNATURES = [
    ("ffm_abi_version", C.c_int, []),
    ("ffm_ctx_create", C.c_int, [C.c_int, C.POINTER(_vp)]),
    ("ffm_ctx_destroy", C.c_int, [_vp]),
    ("ffm_last_error", C.c_char_p, [_vp]),
    ("ffm_cancel", C.c_int, [_vp]),
    ("ffm_flame_set", C.c_int, [_vp, C.POINTER(FlameDesc)]),
    ("ffm_palette_set", C.c_int, [_vp, _fp, C.c_int, C.c_int]),
    ("ffm_render", C.c_int, [_vp, C.POINTER(Job), _u8p, C.POINTER(RenderStats)]),
    ("ffm_render_to_device", C.c_int, [_vp, C.POINTER(Job), _vp, C.POINTER(RenderStats)]),
    ("ffm_render_accum", C.c_int, [_vp, C.POINTER(Job), _fp, C.POINTER(RenderStats)]),
    ("ffm_iterate", C.c_int, [_vp, C.POINTER(View), C.c_uint64, C.c_uint32, C.c_uint32,
                              C.c_uint32, _vp, C.c_int, _vp, C.c_int, _vp, _vp, _vp]),
    ("ffm_iterate_noalloc", C.c_int, [_vp, C.POINTER(View), C.c_uint64, C.c_uint32,
                                      C.c_uint32, C.c_uint32, _vp, C.c_int, _vp, C.c_int, _vp,
                                      _vp, _vp]),
    ("ffm_trace", C.c_int, [_vp, C.POINTER(View), C.c_uint32, C.c_uint32, C.c_uint32,
                            C.c_uint32, _vp, _vp]),
    ("ffm_blur", C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, C.c_float, C.c_float, _vp]),
    ("ffm_density_estimate", C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, C.c_float, C.c_float,
                                       C.c_float, _vp]),
    ("ffm_accumulate", C.c_int, [_vp, _vp, C.c_int, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int,
                                 C.c_float, _vp]),
    ("ffm_tonemap", C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, C.c_float, C.c_double,
                              C.POINTER(TonemapParams), _vp]),
    ("ffm_fold_xcd", C.c_int, [_vp, _vp, C.c_int, C.c_int, _vp]),
    ("ffm_downsample2x", C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, _vp]),
    ("ffm_batch_render", C.c_int, [_vp, C.c_int, C.POINTER(FlameDesc), C.POINTER(_fp),
                                   C.POINTER(C.c_int), C.POINTER(Job), C.POINTER(_u8p),
                                   C.POINTER(RenderStats)]),
    ("ffm_batch_render_serial", C.c_int, [_vp, C.c_int, C.POINTER(FlameDesc), C.POINTER(_fp),
                                          C.POINTER(C.c_int), C.POINTER(Job), C.POINTER(_u8p),
                                          C.POINTER(RenderStats)]),
    ("ffm_batch_render_to_device", C.c_int, [_vp, C.c_int, C.POINTER(FlameDesc), C.POINTER(_fp),
                                             C.POINTER(C.c_int), C.POINTER(Job), C.POINTER(_u8p),
                                             C.POINTER(RenderStats)]),
    ("ffm_batch_render_serial_to_device", C.c_int,
     [_vp, C.c_int, C.POINTER(FlameDesc), C.POINTER(_fp), C.POINTER(C.c_int), C.POINTER(Job),
      C.POINTER(_u8p), C.POINTER(RenderStats)]),
    ("ffm_softsplat_param_count", C.c_int, [C.POINTER(FlameDesc)]),
    ("ffm_softsplat_pack", C.c_int, [C.POINTER(FlameDesc), _fp]),
    ("ffm_softsplat_set", C.c_int, [_vp, C.c_int, C.POINTER(FlameDesc), C.POINTER(SoftJob)]),
    ("ffm_softsplat_sizes", C.c_int, [_vp, _u64p, _u64p]),
    ("ffm_softsplat_forward", C.c_int, [_vp, _vp, _vp, _vp]),
    ("ffm_softsplat_adjoint", C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    ("ffm_scene_count", C.c_int, [C.c_char_p, C.c_size_t, C.POINTER(C.c_int), C.c_char_p,
                                  C.c_size_t]),
    ("ffm_scene_parse", C.c_int, [C.c_char_p, C.c_size_t, C.c_int, C.POINTER(Scene),
                                  C.c_char_p, C.c_size_t]),
    ("ffm_palette_set_xf", C.c_int, [_vp, _fp, C.c_int, C.c_int, C.POINTER(PaletteXf)]),
    ("ffm_scene_job", C.c_int, [C.POINTER(Scene), C.c_int, C.c_int, C.c_uint64, C.c_uint32,
                                C.c_uint32, C.POINTER(Job), C.c_char_p, C.c_size_t]),
    ("ffm_render_config", C.c_int, [_vp, C.c_char_p, C.c_size_t, C.c_int, C.c_int, C.c_uint64,
                                    C.c_uint32, C.c_uint32, _u8p, C.POINTER(RenderStats)]),
    ("ffm_effect_defaults", C.c_int, [C.c_int32, C.POINTER(Effect)]),
    ("ffm_effects_set", C.c_int, [_vp, C.POINTER(Effects)]),
    ("ffm_density_effect", C.c_int, [_vp, C.POINTER(Effect), _vp, _vp, C.c_int, C.c_int,
                                     C.c_float, C.c_float, _vp]),
    ("ffm_color_effect", C.c_int, [_vp, C.POINTER(Effect), _vp, _vp, C.c_int, C.c_int,
                                   C.c_float, _vp]),
]


def bind(lib):
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib
