"""ctypes binding of the CPU oracle (libffm_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker, never as the product.
"""
import ctypes as C
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "fflame-rust_amd"))
from pyfae import _abi  # noqa: E402
from pyfae._abi import FlameDesc, Job, SoftJob, TonemapParams, View  # noqa: E402

LIB = os.path.join(HERE, "libffm_oracle.so")


class Walker(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("color_index", C.c_float),
                ("fuse", C.c_uint32), ("prev", C.c_uint32)]


class Rng(C.Structure):
    _fields_ = [("state", C.c_uint32), ("inc", C.c_uint32)]


class OrbitFuses(C.Structure):
    _fields_ = [("left", C.POINTER(C.c_uint32)), ("len", C.c_uint32)]


def build():
    subprocess.run(["make", "-s", "-B", "-C", HERE], check=True)


_lib = None
_vp, _fp, _u32p, _u8p, _u64p = (C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_uint32),
                                C.POINTER(C.c_uint8), C.POINTER(C.c_uint64))
_dp = C.POINTER(C.c_double)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        sig = {
            "orc_pcg_hash": (C.c_uint32, [C.c_uint32]),
            "orc_rng_init": (None, [C.POINTER(Rng), C.c_uint32, C.c_uint32]),
            "orc_rng_next": (C.c_uint32, [C.POINTER(Rng)]),
            "orc_rng_nextf": (C.c_float, [C.POINTER(Rng)]),
            "orc_rng_stream": (None, [C.c_uint32, C.c_uint32, C.c_uint32, _u32p]),
            "orc_select": (C.c_int, [C.POINTER(FlameDesc), C.c_float]),
            "orc_select_xaos": (C.c_int, [C.POINTER(FlameDesc), C.c_float, C.c_int]),
            "orc_xaos_refresh_period": (C.c_uint32, [C.POINTER(FlameDesc)]),
            "orc_variation": (None, [C.c_int, _fp, C.POINTER(_abi.Xform), C.c_float, C.c_float,
                                     C.POINTER(Rng), _fp, _fp]),
            "orc_apply_xform": (None, [C.POINTER(FlameDesc), C.POINTER(_abi.Xform), _fp, _fp,
                                       C.POINTER(Rng)]),
            "orc_flame_prepare": (None, [C.POINTER(FlameDesc)]),
            "orc_palette_sample": (None, [_fp, C.c_int, C.c_float, _fp]),
            "orc_f32_to_f16": (C.c_uint16, [C.c_float]),
            "orc_f16_to_f32": (C.c_float, [C.c_uint16]),
            "orc_palette_linear": (None, [_fp, C.c_int, C.c_int, _fp]),
            "orc_palette_texel": (C.c_int, [C.c_float, C.c_int]),
            "orc_world_to_pixel": (None, [C.POINTER(View), C.c_float, C.c_float,
                                          C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
            "orc_iterate": (C.c_uint64, [C.POINTER(FlameDesc), _fp, C.c_int, C.POINTER(View),
                                         C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                         C.c_uint32, C.POINTER(Walker), C.c_int, C.c_uint32,
                                         _u32p, _u64p]),
            "orc_iterate_mt": (C.c_uint64, [C.POINTER(FlameDesc), _fp, C.c_int, C.POINTER(View),
                                            C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _u32p,
                                            C.c_int]),
            "orc_trace": (None, [C.POINTER(FlameDesc), _fp, C.c_int, C.POINTER(View), C.c_uint32,
                                 C.c_uint32, C.c_uint32, C.c_uint32, _u32p]),
            "orc_step_from": (C.c_int, [C.POINTER(FlameDesc), _fp, C.c_int, C.POINTER(View),
                                        C.c_uint32, C.POINTER(Rng), C.POINTER(Walker), _u32p]),
            "orc_orbit_fuses_dispatch": (C.c_uint64, [C.POINTER(OrbitFuses), C.c_uint32,
                                                      C.c_uint32, C.c_uint32, C.c_uint32,
                                                      C.c_uint32]),
            "orc_orbit_fuses_free": (None, [C.POINTER(OrbitFuses)]),
            "orc_blur_pass": (None, [_u32p, _u32p, _u32p, C.c_int, C.c_int, C.c_float, C.c_float,
                                     C.c_int]),
            "orc_blur": (None, [_u32p, _u32p, C.c_int, C.c_int, C.c_float, C.c_float]),
            "orc_blur_density_sigma": (C.c_float, [C.c_double, C.c_int, C.c_int, C.c_float]),
            "orc_accumulate": (None, [_vp, C.c_int, _fp, _fp, _u32p, C.c_int, C.c_int, C.c_int,
                                      C.c_float]),
            "orc_density_estimate": (None, [_fp, _fp, C.c_int, C.c_int, C.c_float, C.c_float,
                                            C.c_float]),
            "orc_round_f16": (C.c_float, [C.c_float]),
            "orc_curve_eval": (C.c_float, [_fp, _fp, C.c_int, C.c_float]),
            "orc_curve_lut": (None, [C.c_void_p, _fp]),
            "orc_curve_sample": (C.c_float, [_fp, C.c_float]),
            "orc_step_term_scale": (C.c_float, []),
            "orc_downsample2x": (None, [_u8p, _u8p, C.c_int, C.c_int]),
            "orc_use_libm": (None, [C.c_int]),
            "orc_fm_atan2": (C.c_float, [C.c_float, C.c_float]),
            "orc_tonemap": (None, [_fp, _u8p, C.c_int, C.c_int, C.c_float, C.c_double,
                                   C.POINTER(TonemapParams)]),
            "orc_render": (C.c_int, [C.POINTER(FlameDesc), _fp, C.c_int, C.c_int,
                                     C.POINTER(_abi.PaletteXf), C.POINTER(Job), C.c_uint64, _fp,
                                     _u8p, _u64p, C.c_int, C.POINTER(_abi.Effects)]),
            "orc_density_effect": (None, [C.POINTER(_abi.Effect), _fp, _fp, C.c_int, C.c_int,
                                          C.c_float, C.c_float]),
            "orc_color_effect": (None, [C.POINTER(_abi.Effect), _u8p, _u8p, C.c_int, C.c_int,
                                        C.c_float]),
            "orc_apply_blend": (None, [_fp, _fp, C.c_int, C.c_float, _fp]),
            "orc_fx_sincos": (None, [C.c_float, _fp, _fp]),
            "orc_palette_index_map": (None, [C.c_int, C.POINTER(_abi.PaletteXf),
                                             C.POINTER(C.c_int32)]),
            "orc_palette_linear_xf": (None, [_fp, C.c_int, C.c_int, C.POINTER(_abi.PaletteXf),
                                             _fp]),
            "orc_soft_param_count": (C.c_int, [C.POINTER(FlameDesc)]),
            "orc_soft_forward": (C.c_int, [C.POINTER(FlameDesc), _dp, C.POINTER(SoftJob), _dp,
                                           C.c_int]),
            "orc_soft_adjoint": (C.c_int, [C.POINTER(FlameDesc), _dp, C.POINTER(SoftJob), _dp,
                                           _dp, C.c_int]),
            "orc_soft_variation": (C.c_int, [C.c_int, _fp, C.c_double, C.c_double, _dp, _dp]),
            "orc_microbench": (C.c_double, [C.c_int, _fp, C.c_uint64, C.c_double, _fp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def fp(a):
    return a.ctypes.data_as(_fp)


def u32p(a):
    return a.ctypes.data_as(_u32p)


# ------------------------------------------------------------ conveniences

def prepared(fd):
    """A copy of the flame with derived variation slots filled."""
    f2 = FlameDesc()
    C.memmove(C.byref(f2), C.byref(fd), C.sizeof(FlameDesc))
    lib().orc_flame_prepare(C.byref(f2))
    return f2


def palette_linear(stops, size=256):
    s = np.ascontiguousarray(stops, dtype=np.float32)
    out = np.zeros(3 * size, dtype=np.float32)
    lib().orc_palette_linear(fp(s), s.shape[0], size, fp(out))
    return out


def iterate(fd, stops, view, walkers, ipt, burn_in, seed, *, walker0=0, pal_size=256,
            orbits=None, fresh=True, refresh_period=0):
    """(hist u32 (H, W, 4), plotted, inframe) of one dispatch. `orbits` is a
    Walker array (persistent orbits) or None."""
    f2 = prepared(fd)
    pal = palette_linear(stops, pal_size)
    hist = np.zeros((view.height, view.width, 4), dtype=np.uint32)
    infr = C.c_uint64(0)
    orb = C.cast(orbits, C.POINTER(Walker)) if orbits is not None else None
    plotted = lib().orc_iterate(C.byref(f2), fp(pal), pal_size, C.byref(view), walker0, walkers,
                                ipt, burn_in, seed & 0xFFFFFFFF, orb, int(bool(fresh)),
                                refresh_period, u32p(hist), C.byref(infr))
    return hist, plotted, infr.value


def select_xaos(fd, u, prev):
    return lib().orc_select_xaos(C.byref(fd), u, prev)


def xaos_refresh_period(fd):
    return lib().orc_xaos_refresh_period(C.byref(fd))


def iterate_mt(fd, stops, view, walkers, ipt, burn_in, seed, threads, pal_size=256):
    f2 = prepared(fd)
    pal = palette_linear(stops, pal_size)
    hist = np.zeros((view.height, view.width, 4), dtype=np.uint32)
    plotted = lib().orc_iterate_mt(C.byref(f2), fp(pal), pal_size, C.byref(view), walkers, ipt,
                                   burn_in, seed & 0xFFFFFFFF, u32p(hist), threads)
    return hist, plotted


def trace(fd, stops, view, walker, seed, burn_in, steps, pal_size=256):
    f2 = prepared(fd)
    pal = palette_linear(stops, pal_size)
    out = np.zeros((steps, 8), dtype=np.uint32)
    lib().orc_trace(C.byref(f2), fp(pal), pal_size, C.byref(view), walker, seed & 0xFFFFFFFF,
                    burn_in, steps, u32p(out))
    return out


def replay_steps(fd, stops, view, walker, seed, burn_in, trace_rows, pal_size=256,
                 with_scale=False):
    """Apply one oracle step to each state of a (device) trace: row t of the
    result is the oracle's step from the trace's state after step t-1 (the
    walker's start state for t = 0). Returns an (n, 8) record array; with
    `with_scale`, also each step's term scale (the largest magnitude among
    the terms the step sums: the point, the affine point, the weighted
    variation outputs, the post-affine products)."""
    f2 = prepared(fd)
    pal = palette_linear(stops, pal_size)
    L = lib()
    r = Rng()
    L.orc_rng_init(C.byref(r), walker, seed & 0xFFFFFFFF)
    inc = r.inc
    w = Walker()
    two, one = np.float32(2.0), np.float32(1.0)
    w.x = float(np.float32(L.orc_rng_nextf(C.byref(r))) * two - one)
    w.y = float(np.float32(L.orc_rng_nextf(C.byref(r))) * two - one)
    w.color_index = 0.0
    w.fuse = burn_in
    if fd.has_xaos:  # the unconditioned first pick draws once more
        w.prev = L.orc_select(C.byref(f2), L.orc_rng_nextf(C.byref(r)))
    fuse = burn_in
    out = np.zeros((len(trace_rows), 8), dtype=np.uint32)
    rec = (C.c_uint32 * 8)()
    scales = np.zeros(len(trace_rows), dtype=np.float64)
    for t, row in enumerate(trace_rows):
        L.orc_step_from(C.byref(f2), fp(pal), pal_size, C.byref(view), burn_in, C.byref(r),
                        C.byref(w), rec)
        out[t] = list(rec)
        scales[t] = L.orc_step_term_scale()
        # continue from the device's state after this step
        flags = int(row[7])
        fuse = burn_in if flags & 2 else (fuse - 1 if flags & 4 else fuse)
        r.state, r.inc = int(row[0]), inc
        w.x = np.uint32(row[2]).view(np.float32).item()
        w.y = np.uint32(row[3]).view(np.float32).item()
        w.color_index = np.uint32(row[4]).view(np.float32).item()
        w.fuse = fuse
        w.prev = int(row[1])  # with xaos, the pick conditions the next one
    return (out, scales) if with_scale else out


def render(fd, stops, job, pal_size=256, threads=1, palette_xf=None, exact_after=1 << 18,
           effects=None):
    """(rgba uint8 (H, W, 4), accum float32 (H, W, 4), total samples).
    `exact_after`: the accumulate pass from which densities count exactly;
    `effects`: an ffm_effects applied as the render's tail."""
    w, h = job.view.width, job.view.height
    rgba = np.zeros((h, w, 4), dtype=np.uint8)
    acc = np.zeros((h, w, 4), dtype=np.float32)
    tot = C.c_uint64(0)
    s = np.ascontiguousarray(stops, dtype=np.float32)
    xf = C.byref(palette_xf) if palette_xf is not None else None
    lib().orc_render(C.byref(fd), fp(s), s.shape[0], pal_size, xf, C.byref(job), exact_after,
                     fp(acc), rgba.ctypes.data_as(_u8p), C.byref(tot), threads,
                     C.byref(effects) if effects is not None else None)
    return rgba, acc, tot.value


def density_effect(fx, img, density_mean=1.0, pixel_scale=1.0):
    """One density effect over an (H, W, 4) float32 image."""
    a = np.ascontiguousarray(img, dtype=np.float32)
    out = np.empty_like(a)
    lib().orc_density_effect(C.byref(fx), fp(a), fp(out), a.shape[1], a.shape[0], density_mean,
                             pixel_scale)
    return out


def color_effect(fx, img, pixel_scale=1.0):
    """One colour effect over an (H, W, 4) uint8 image."""
    a = np.ascontiguousarray(img, dtype=np.uint8)
    out = np.empty_like(a)
    lib().orc_color_effect(C.byref(fx), a.ctypes.data_as(_u8p), out.ctypes.data_as(_u8p),
                           a.shape[1], a.shape[0], pixel_scale)
    return out


def apply_blend(base, effect, mode, amount):
    """The blend library's apply_blend on one rgb triple (float32)."""
    b = np.ascontiguousarray(base, dtype=np.float32)
    e = np.ascontiguousarray(effect, dtype=np.float32)
    out = np.zeros(3, np.float32)
    lib().orc_apply_blend(fp(b), fp(e), int(mode), float(amount), fp(out))
    return out


def palette_index_map(size, palette_xf=None):
    """The texel each entry of the transformed palette table shows."""
    out = np.zeros(size, dtype=np.int32)
    xf = C.byref(palette_xf) if palette_xf is not None else None
    lib().orc_palette_index_map(size, xf, out.ctypes.data_as(C.POINTER(C.c_int32)))
    return out


def accumulate(hist, prev, count=None, mode=0, blend=0.0):
    h, w = hist.shape[:2]
    out = np.zeros((h, w, 4), dtype=np.float32)
    is_f32 = hist.dtype == np.float32
    lib().orc_accumulate(hist.ctypes.data_as(_vp), int(is_f32), fp(prev), fp(out),
                         u32p(count) if count is not None else None, w, h, mode, blend)
    return out


def round_f16(v):
    return lib().orc_round_f16(float(v))


def curve_eval(points, x):
    """The tone curve through `points` [(x, y), ...] at x."""
    px = np.array([p[0] for p in points], dtype=np.float32)
    py = np.array([p[1] for p in points], dtype=np.float32)
    return lib().orc_curve_eval(fp(px), fp(py), len(points), float(x))


def curve_lut(tp):
    out = np.zeros(256, dtype=np.float32)
    lib().orc_curve_lut(C.byref(tp), fp(out))
    return out


def curve_sample(lut, c):
    """The curve table `lut` (256 float32) read by the tone map at channel c."""
    lut = np.ascontiguousarray(lut, dtype=np.float32)
    return lib().orc_curve_sample(fp(lut), float(c))


def downsample2x(img):
    """(2h, 2w, 4) uint8 -> (h, w, 4): the export's 2x box with the firefly clamp."""
    img = np.ascontiguousarray(img, dtype=np.uint8)
    h, w = img.shape[0] // 2, img.shape[1] // 2
    out = np.zeros((h, w, 4), dtype=np.uint8)
    lib().orc_downsample2x(img.ctypes.data_as(_u8p), out.ctypes.data_as(_u8p), w, h)
    return out


def tonemap(acc, zoom, samples, tp):
    h, w = acc.shape[:2]
    out = np.zeros((h, w, 4), dtype=np.uint8)
    lib().orc_tonemap(fp(acc), out.ctypes.data_as(_u8p), w, h, zoom, float(samples), C.byref(tp))
    return out


def blur(hist, sigma, density_sigma):
    h, w = hist.shape[:2]
    hh = np.ascontiguousarray(hist, dtype=np.uint32).copy()
    scratch = np.zeros_like(hh)
    lib().orc_blur(u32p(hh), u32p(scratch), w, h, sigma, density_sigma)
    return hh


def density_estimate(acc, max_r, min_r, curve):
    h, w = acc.shape[:2]
    out = np.zeros_like(acc)
    lib().orc_density_estimate(fp(np.ascontiguousarray(acc)), fp(out), w, h, max_r, min_r, curve)
    return out


# ---------------------------------------------------------------- soft splat

def _dptr(a):
    return a.ctypes.data_as(_dp)


def soft_forward(fd, theta, job, threads=8):
    """FP64 soft-splat density (H, W) at theta (float64, the ABI's layout)."""
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    out = np.zeros((job.view.height, job.view.width), np.float64)
    st = lib().orc_soft_forward(C.byref(fd), _dptr(theta), C.byref(job), _dptr(out), threads)
    if st:
        raise ValueError("orc_soft_forward: %d" % st)
    return out


def soft_adjoint(fd, theta, job, grad_density, threads=8):
    """FP64 dL/dtheta for dL/d density = grad_density."""
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    g = np.ascontiguousarray(grad_density, dtype=np.float64)
    P = lib().orc_soft_param_count(C.byref(fd))
    if P < 0:
        raise ValueError("flame cannot be fitted")
    out = np.zeros(P, np.float64)
    st = lib().orc_soft_adjoint(C.byref(fd), _dptr(theta), C.byref(job), _dptr(g), _dptr(out),
                                threads)
    if st:
        raise ValueError("orc_soft_adjoint: %d" % st)
    return out


def soft_variation(var_id, x, y, params=(0.0, 0.0, 0.0, 0.0)):
    """(V[2], J[2x2]) of one variation in FP64."""
    p = (C.c_float * 4)(*params)
    V = np.zeros(2)
    J = np.zeros(4)
    lib().orc_soft_variation(var_id, p, x, y, _dptr(V), _dptr(J))
    return V, J.reshape(2, 2)



# ------------------------------------------------ the reference's CPU benches

# The first transform of the first three presets (assets/presets.fae):
# a b c d e f, colour, colour speed, linear weight (None: no linear variation).
PRESET_FIRST_XFORMS = {
    "Grand": (1.0, 0.0, 0.0, 1.0, 0.0, 0.0, 1.0, 0.0, None),
    "Julian": (0.87169206, 0.38354817, -0.38360286, 0.8721385, 0.0, 0.0, 0.0, 0.9, None),
    "Plastic": (1.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, -0.8, None),
}

# (Criterion id, benchmark, params, the reference's published mean in ns on
# one core: benchmark_results/last_run.json:5-75). Inputs as in
# benches/flame_bench.rs: affine a..g = 0.7 0.2 -0.2 0.8 0.1 -0.1 0.05 at
# (0.5, 0.5); point helpers at (0.7, 0.3); palette position 0.3, colour 0.7.
_AFF = (0.7, 0.2, -0.2, 0.8, 0.1, -0.1)
MICROBENCHES = [
    ("affine_transform/affine_2x3_matrix", 0, _AFF + (0.5, 0.5, 0, 0, 0.05), 2.7385),
    ("affine_transform/affine_full", 1, _AFF + (0.5, 0.5, 0, 0, 0.05), 3.1707),
    ("point_helpers/r_squared", 2, (0,) * 6 + (0.7, 0.3), 0.66151),
    ("point_helpers/r", 3, (0,) * 6 + (0.7, 0.3), 0.74991),
    ("point_helpers/theta", 4, (0,) * 6 + (0.7, 0.3), 10.223),
    ("point_helpers/phi", 5, (0,) * 6 + (0.7, 0.3), 11.307),
    ("color_blending/palette_position/0", 6, (0,) * 6 + (0.3, 0.7, 0.0), 0.63141),
    ("color_blending/palette_position/0.5", 6, (0,) * 6 + (0.3, 0.7, 0.5), 0.64191),
    ("color_blending/palette_position/1", 6, (0,) * 6 + (0.3, 0.7, 1.0), 0.62891),
] + [
    (f"preset_iteration/single_transform_iter/{k}", 7,
     v[:6] + (0.5, 0.5, v[7], v[6], v[8] or 0.0, 0.0 if v[8] is None else 1.0), pub)
    for k, pub in (("Grand", 19.716), ("Julian", 162.53), ("Plastic", 20.303))
    for v in (PRESET_FIRST_XFORMS[k],)
]


def microbench(which, params, iters=0, min_seconds=0.1):
    """ns per iteration of one restated Criterion benchmark, and the last
    (x, y, colour) values it computed."""
    p = np.zeros(12, np.float32)
    p[:len(params)] = params
    st = np.zeros(3, np.float32)
    ns = lib().orc_microbench(which, fp(p), iters, min_seconds, fp(st))
    return ns, st


def microbench_suite(min_seconds=0.1):
    """Every restated benchmark: {id: {"ns": ours, "published_ns": reference}}."""
    return {name: {"ns": round(microbench(w, p, 0, min_seconds)[0], 4), "published_ns": pub}
            for name, w, p, pub in MICROBENCHES}
