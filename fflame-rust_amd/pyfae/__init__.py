"""pyfae (MI355X build): the flame path of the Fractal Art Editor's Python API,
over libffm.so -- hand-written gfx950 HIP kernels behind a C ABI.

Mirrors the reference's Python surface for this path (python/pyfae/__init__.py:
`Engine.render(config, width, height, iterations=..., iterations_per_thread=...)`
returning an image) and exposes the granular kernels for device buffers
(torch tensors or raw device pointers) plus the differentiable soft-splat
mode as a torch.autograd.Function (pyfae.diff).

There is no CPU fallback: without libffm.so (or without a GPU) every call
raises EngineError.
"""
import ctypes as C
import json
import os

import numpy as np

from . import _abi, formats
from ._abi import (FlameDesc, Job, PaletteXf, RenderStats, Scene, SoftJob,  # noqa: F401
                   TonemapParams, View, Xform, VARIATIONS, tonemap_defaults)

__version__ = "0.5.0+mi355x"

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libffm.so")


class EngineError(RuntimeError):
    """A failed engine call: `status` is the ffm_status code."""

    def __init__(self, status, message):
        super().__init__(f"{message} (status {status})")
        self.status = status


_lib = None


def library():
    """The loaded libffm.so (raises EngineError when it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise EngineError(_abi.ERR_DEVICE,
                              f"{LIB_PATH} is not built: run fflame-rust_amd/build.py")
        _lib = _abi.bind(C.CDLL(LIB_PATH))
    return _lib


def _ptr(x):
    """Device pointer of a torch tensor, an int, or None."""
    if x is None:
        return None
    if isinstance(x, int):
        return C.c_void_p(x)
    if hasattr(x, "data_ptr"):
        return C.c_void_p(x.data_ptr())
    raise TypeError(f"expected a tensor or device pointer, got {type(x)}")


# ---------------------------------------------------------------- flames

def xform(a=1.0, b=0.0, c=0.0, d=1.0, e=0.0, f=0.0, *, weight=1.0, color=0.0, color_speed=0.5,
          opacity=1.0, variations=None, params=None, post=None, final_attachments=None):
    """A transform as the reference's Transform holds it (src/scene/transforms.rs:337-530):
    `variations` {name: weight}, `params` {"julian.power": 3.0, ...},
    `post` (a, b, c, d, e, f) or None, `final_attachments` the indices of the
    finals it applies (None: every final, in order)."""
    return {"a": a, "b": b, "c": c, "d": d, "e": e, "f": f, "weight": weight, "color": color,
            "color_speed": color_speed, "opacity": opacity,
            "variations": dict(variations or {"linear": 1.0}), "params": dict(params or {}),
            "post": tuple(post) if post is not None else None,
            "final_attachments": list(final_attachments) if final_attachments is not None
            else None}


def flame_desc(transforms, finals=(), xaos=None):
    """Pack transforms (dicts from `xform`) into an ffm_flame_desc. Active
    variations get local indices in first-appearance order over the
    transforms, then the finals (src/scene/transforms.rs:30-85). `xaos`
    (n x n, row = the transform applied last) scales the next pick's weights
    (Flame::xaos, src/scene/transforms.rs:2040-2044)."""
    if len(transforms) > _abi.MAX_XFORMS:
        raise EngineError(_abi.ERR_TOO_LARGE,
                          f"{len(transforms)} transforms: at most {_abi.MAX_XFORMS} fit")
    order = []
    for t in list(transforms) + list(finals):
        for name in t["variations"]:
            if name not in VARIATIONS:
                raise EngineError(_abi.ERR_INVALID_ARGUMENT, f"unknown variation {name!r}")
            if name not in order:
                order.append(name)
    if len(order) > _abi.MAX_VARS:
        raise EngineError(_abi.ERR_TOO_LARGE, f"{len(order)} active variations")
    fd = FlameDesc()
    fd.n_xforms = len(transforms)
    fd.n_finals = len(finals)
    fd.n_vars = len(order)
    for k, name in enumerate(order):
        fd.var_id[k] = VARIATIONS[name]

    def fill(dst, t):
        for key in ("a", "b", "c", "d", "e", "f", "weight", "color", "color_speed", "opacity"):
            setattr(dst, key, float(t[key]))
        post = t.get("post")
        if post is not None:
            dst.post_enabled = 1
            dst.post_a, dst.post_b, dst.post_c, dst.post_d, dst.post_e, dst.post_f = map(float, post)
        else:
            dst.post_enabled = 0
            dst.post_a, dst.post_d = 1.0, 1.0
        for k, name in enumerate(order):
            dst.var_weight[k] = float(t["variations"].get(name, 0.0))
            names = _abi.VARIATION_PARAMS.get(name, ())
            defaults = _abi.PARAM_DEFAULTS.get(name, ())
            for q, pname in enumerate(names):
                dst.var_param[k][q] = float(t["params"].get(f"{name}.{pname}", defaults[q]))

    for i, t in enumerate(transforms):
        fill(fd.xforms[i], t)
    for i, t in enumerate(finals):
        fill(fd.finals[i], t)
    # per-transform final chains when any transform names its own
    # (a transform that names none then applies none, as the reference's)
    if any(t.get("final_attachments") is not None for t in transforms):
        fd.per_transform_finals = 1
        for i, t in enumerate(transforms):
            att = t.get("final_attachments") or []
            if len(att) > _abi.MAX_FINALS:
                raise EngineError(_abi.ERR_TOO_LARGE, f"transform {i} attaches {len(att)} finals")
            fd.xforms[i].n_final_attach = len(att)
            for j, k in enumerate(att):
                fd.xforms[i].final_attach[j] = int(k)
    if xaos is not None:
        m = np.asarray(xaos, dtype=np.float32)
        n = len(transforms)
        if m.shape != (n, n):
            raise EngineError(_abi.ERR_INVALID_ARGUMENT, f"xaos must be {n} x {n}")
        if n > _abi.MAX_XAOS:
            raise EngineError(_abi.ERR_TOO_LARGE, f"xaos on {n} transforms: at most {_abi.MAX_XAOS}")
        fd.has_xaos = 1
        for i in range(n):
            for j in range(n):
                fd.xaos[i][j] = float(m[i, j])
    return fd


def palette_stops(stops):
    """(n, 4) float32 array of (position, r, g, b) stops, sRGB in 0..1."""
    a = np.ascontiguousarray(np.asarray(stops, dtype=np.float32).reshape(-1, 4))
    return a


def effect(name, enabled=True, **params):
    """An ffm_effect: the built-in effect `name` (its registry id, e.g.
    "density_blur", "vignette") with the registry's defaults, overridden by
    `params` by name (src/effects/mod.rs:319-336 EffectInstance)."""
    if name not in _abi.EFFECTS:
        raise ValueError(f"effect '{name}' is not one this engine runs "
                         f"({', '.join(sorted(_abi.EFFECTS))})")
    t, _, names = _abi.EFFECTS[name]
    fx = _abi.Effect()
    st = library().ffm_effect_defaults(t, C.byref(fx))
    if st != 0:
        raise EngineError(st, f"no defaults for effect {name}")
    fx.enabled = 1 if enabled else 0
    for k, v in params.items():
        if k not in names:
            raise ValueError(f"effect '{name}' has no parameter '{k}' ({', '.join(names)})")
        fx.params[names.index(k)] = float(v)
    return fx


def _as_effect(e):
    if isinstance(e, _abi.Effect):
        return e
    if isinstance(e, str):
        return effect(e)
    if isinstance(e, dict):  # a config's EffectInstance
        return effect(e["effect_type"], e.get("enabled", True), **(e.get("params") or {}))
    name, params = e
    return effect(name, **params)


def effects(density=(), color=(), pixel_scale=1.0):
    """An ffm_effects: the density chain (before the tone map) and the
    colour chain (after it), each item an ffm_effect, a name, a
    (name, params) pair or a config's {"effect_type", "enabled", "params"}."""
    fx = _abi.Effects()
    density, color = [_as_effect(e) for e in density], [_as_effect(e) for e in color]
    if len(density) > _abi.MAX_EFFECTS or len(color) > _abi.MAX_EFFECTS:
        raise ValueError(f"at most {_abi.MAX_EFFECTS} effects a chain")
    for i, e in enumerate(density):
        if _abi.EFFECTS[effect_name(e.type)][1] != "density":
            raise ValueError(f"{effect_name(e.type)} is a colour effect")
        fx.density[i] = e
    for i, e in enumerate(color):
        if _abi.EFFECTS[effect_name(e.type)][1] != "color":
            raise ValueError(f"{effect_name(e.type)} is a density effect")
        fx.color[i] = e
    fx.n_density, fx.n_color = len(density), len(color)
    fx.pixel_scale = float(pixel_scale)
    return fx


def effect_name(t):
    for k, (tt, _, _) in _abi.EFFECTS.items():
        if tt == t:
            return k
    raise ValueError(f"no effect type {t}")


def make_job(width, height, *, walkers=8192, ipt=256, burn_in=20, seed_base=12345,
             seed_stride=2654435761, target_samples=0, batch_size=4, splat_mode=0,
             persistent_orbits=1, zoom=1.0, pan_x=0.0, pan_y=0.0, rotation=0.0,
             filter_radius=0.0, filter_blur_edges=0.5, de=None, tonemap=None):
    """An ffm_job. Defaults follow RenderJob::new and the render loop
    (src/renderer/render.rs:111-138,350-351) and the deterministic seed
    schedule 12345 + k*2654435761 (src/renderer/compute_kernel.rs:3541-3554)."""
    j = Job()
    j.view = View(width, height, zoom, pan_x, pan_y, rotation)
    j.walkers = walkers
    j.ipt = ipt
    j.burn_in = burn_in
    j.seed_base = seed_base & 0xFFFFFFFF
    j.seed_stride = seed_stride & 0xFFFFFFFF
    j.target_samples = target_samples
    j.batch_size = batch_size
    j.splat_mode = splat_mode
    j.persistent_orbits = persistent_orbits
    j.filter_radius = filter_radius
    j.filter_blur_edges = filter_blur_edges
    if de is not None:
        j.de_max_radius, j.de_min_radius, j.de_curve = de
    j.tonemap = tonemap if tonemap is not None else tonemap_defaults()
    return j


def soft_job(width, height, *, walkers=1 << 20, ipt=64, burn_in=20, seed=1, window=8,
             truncate=False, density_scale=None, walk_f64=True, zoom=1.0, pan_x=0.0, pan_y=0.0,
             rotation=0.0):
    """An ffm_soft_job. density_scale defaults to W*H / (walkers * (ipt - burn_in)),
    so an attractor that fills the frame has mean density 1. walk_f64 walks
    (and differentiates) in f64; False uses the chaos game's f32 arithmetic."""
    j = SoftJob()
    j.view = View(width, height, zoom, pan_x, pan_y, rotation)
    j.walkers = walkers
    j.ipt = ipt
    j.burn_in = burn_in
    j.seed = seed & 0xFFFFFFFF
    j.window = window
    j.truncate = int(bool(truncate))
    if density_scale is None:
        density_scale = width * height / float(walkers * max(ipt - burn_in, 1))
    j.density_scale = density_scale
    j.walk_f64 = int(bool(walk_f64))
    return j


def soft_theta(fd):
    """The soft-splat parameter vector of a flame (float64): per transform
    (a, b, c, d, e, f), then per transform its n_vars variation weights."""
    n, v = fd.n_xforms, fd.n_vars
    th = np.zeros(n * (6 + v), dtype=np.float64)
    for i in range(n):
        x = fd.xforms[i]
        th[6 * i:6 * i + 6] = (x.a, x.b, x.c, x.d, x.e, x.f)
        th[6 * n + v * i:6 * n + v * (i + 1)] = [x.var_weight[k] for k in range(v)]
    return th


# ---------------------------------------------------------------- configs

def _text(x):
    """Config text from a dict / list (JSON), bytes, JSON or XML text, a
    metadata PNG (its `Config` chunk), or a path to any of these
    (python/pyfae/__init__.py:49-60, with XML and PNG besides)."""
    if isinstance(x, (dict, list)):
        return json.dumps(x)
    if isinstance(x, os.PathLike):
        x = os.fspath(x)
    if isinstance(x, str) and not x.lstrip().startswith(("{", "[", "<")) and os.path.exists(x):
        with open(x, "rb") as f:
            x = f.read()
    if isinstance(x, (bytes, bytearray)):
        if bytes(x[:8]) == formats.PNG_SIGNATURE:
            return formats.config_from_png(bytes(x))
        return bytes(x).decode("utf-8")
    if not isinstance(x, str):
        raise TypeError(f"expected a config (dict, JSON / XML text, PNG or a path), got {type(x)}")
    return x


def scene_count(config):
    """How many configs a text holds: 1 for a JSON object, its length for a
    JSON array, the number of <flame> elements of a .flame document."""
    raw = _text(config).encode("utf-8")
    n, err = C.c_int(), C.create_string_buffer(512)
    st = library().ffm_scene_count(raw, len(raw), C.byref(n), err, 512)
    if st != 0:
        raise EngineError(st, err.value.decode(errors="replace"))
    return n.value


def load(config, index=0):
    """Config `index` of a .fflame / .fae JSON text or a .flame XML document
    (text, dict or path) as an ffm_scene: the flame descriptor, palette,
    view, tone map and budget the reference would render it with. Raises
    EngineError with status ERR_UNSUPPORTED for a feature this engine does
    not render (the message names it)."""
    raw = _text(config).encode("utf-8")
    sc = Scene()
    err = C.create_string_buffer(512)
    st = library().ffm_scene_parse(raw, len(raw), int(index), C.byref(sc), err, 512)
    if st != 0:
        raise EngineError(st, err.value.decode(errors="replace"))
    return sc


def scene_stops(sc):
    """The scene's palette stops as an (n, 4) float32 array."""
    return np.ctypeslib.as_array(sc.stops)[:sc.n_stops].astype(np.float32).copy()


def scene_job(sc, width, height, *, iterations=None, iterations_per_thread=None, as_2d=False,
              transparent=False, premultiplied=False, supersample=False, fill_gpu=False):
    """The ffm_job the reference's export builds for a scene at width x height
    (with `supersample`, the 2x render's job: twice the size, 4x the budget;
    with `fill_gpu`, a GPU-filling walker count a dispatch)."""
    j = Job()
    err = C.create_string_buffer(512)
    flags = ((_abi.RENDER_3D_AS_2D if as_2d else 0)
             | (_abi.RENDER_TRANSPARENT if transparent else 0)
             | (_abi.RENDER_PREMULTIPLIED if premultiplied else 0)
             | (_abi.RENDER_SUPERSAMPLE_2X if supersample else 0)
             | (_abi.RENDER_FILL_GPU if fill_gpu else 0))
    st = library().ffm_scene_job(C.byref(sc), int(width), int(height), int(iterations or 0),
                                 int(iterations_per_thread or 0), flags, C.byref(j), err, 512)
    if st != 0:
        raise EngineError(st, err.value.decode(errors="replace"))
    return j


class Image:
    """A rendered frame (python/pyfae/__init__.py:76-113): RGBA8 pixels,
    `width` x `height`, and its numbers (`info`: iterations, ms,
    iterations_per_thread, dispatches)."""

    def __init__(self, rgba, info, scene=None):
        self.rgba = rgba
        self.info = info
        self.width = info["width"]
        self.height = info["height"]
        self.scene = scene

    def to_numpy(self):
        """An (height, width, 4) uint8 array."""
        return self.rgba

    def to_pil(self):
        from PIL import Image as PILImage
        return PILImage.fromarray(self.rgba, "RGBA")

    def png(self, strip_metadata=False):
        """The PNG the reference's export writes for this frame: its config
        (as FractalConfig JSON) and numbers in tEXt chunks
        (src/png_metadata.rs:144-260), so it reloads as its flame
        (`load(png)`); `strip_metadata` writes the pixels alone."""
        text = ()
        if self.scene is not None and not strip_metadata:
            text = formats.metadata_chunks(self.width, self.height, self.info["iterations"],
                                           self.info["ms"], self.info["iterations_per_thread"],
                                           formats.scene_to_json(self.scene), self.scene)
        return formats.encode_png(self.rgba, text)

    def save(self, path):
        with open(path, "wb") as f:
            f.write(self.png())

    def __repr__(self):
        return (f"<pyfae.Image {self.width}x{self.height}, {self.info['iterations']:,} "
                f"iterations, {self.info['ms']:.0f} ms>")


# ---------------------------------------------------------------- engine

class Engine:
    """One render context on one GPU (FlameRenderer, src/renderer/compute_kernel.rs:374)."""

    def __init__(self, device=0):
        self.lib = library()
        ctx = C.c_void_p()
        st = self.lib.ffm_ctx_create(int(device), C.byref(ctx))
        if st != 0:
            raise EngineError(st, f"no usable GPU {device}")
        self.ctx = ctx
        self.device = device

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.ffm_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st != 0:
            msg = self.lib.ffm_last_error(self.ctx)
            raise EngineError(st, msg.decode() if msg else "error")

    # -- state
    def set_flame(self, flame):
        fd = flame if isinstance(flame, FlameDesc) else flame_desc(*flame)
        self._check(self.lib.ffm_flame_set(self.ctx, C.byref(fd)))
        return fd

    def set_palette(self, stops, size=256):
        s = palette_stops(stops)
        self._check(self.lib.ffm_palette_set(self.ctx, s.ctypes.data_as(C.POINTER(C.c_float)),
                                             int(s.shape[0]), int(size)))

    def cancel(self):
        self._check(self.lib.ffm_cancel(self.ctx))

    def set_scene(self, sc):
        """Set a scene's flame and palette (with its lookup-table transform)."""
        self._check(self.lib.ffm_flame_set(self.ctx, C.byref(sc.flame)))
        self._check(self.lib.ffm_palette_set_xf(
            self.ctx, C.cast(C.addressof(sc.stops), C.POINTER(C.c_float)), sc.n_stops,
            sc.palette_size, C.byref(sc.palette_xf)))

    def set_effects(self, fx=None, **chains):
        """The tail later renders apply: an ffm_effects, or density=[...],
        color=[...] (see `effects`); None clears it."""
        if fx is None and chains:
            fx = effects(**chains)
        self._check(self.lib.ffm_effects_set(self.ctx, C.byref(fx) if fx is not None else None))

    def density_effect(self, fx, src, dst, density_mean=1.0, pixel_scale=1.0, stream=None):
        """One density effect between (H, W, 4) float32 device tensors."""
        fx = _as_effect(fx)
        self._check(self.lib.ffm_density_effect(self.ctx, C.byref(fx), _ptr(src), _ptr(dst),
                                                int(src.shape[1]), int(src.shape[0]),
                                                float(density_mean), float(pixel_scale),
                                                _ptr(stream)))

    def color_effect(self, fx, src, dst, pixel_scale=1.0, stream=None):
        """One colour effect between (H, W, 4) uint8 device tensors."""
        fx = _as_effect(fx)
        self._check(self.lib.ffm_color_effect(self.ctx, C.byref(fx), _ptr(src), _ptr(dst),
                                              int(src.shape[1]), int(src.shape[0]),
                                              float(pixel_scale), _ptr(stream)))

    # -- whole renders
    def render(self, config, width=None, height=None, *, iterations=None,
               iterations_per_thread=None, transparent=False, premultiplied=False, as_2d=False,
               supersample=False, fill_gpu=False):
        """Render a config (python/pyfae/__init__.py:201-221): `config` is a
        .fflame / .fae JSON or .flame XML text, a dict or a path, rendered at
        `width` x `height` with the config's flame, palette, view and tone map
        (`iterations`: the sample budget, default the config's own;
        `iterations_per_thread`: default 1024; `transparent`: alpha instead
        of the background, straight or `premultiplied`). Returns an Image.
        `as_2d` renders a 3D-mode config (every .flame import) whose camera
        is flat through the 2D chaos game. `supersample` is the export's 2x
        antialiasing (twice the size, 4x the samples, boxed down with the
        firefly clamp); `fill_gpu` dispatches a GPU-filling walker count
        instead of the reference's 8192 (the same picture within the noise
        for a flame that mixes within the burn-in; not for one whose
        picture is its walkers' start points).

        Given an ffm_job instead, renders the current flame and returns
        (pixels, stats)."""
        if isinstance(config, Job):
            return self._render_job(config)
        if width is None or height is None:
            raise TypeError("render(config, width, height): width and height are required")
        text = _text(config)
        raw = text.encode("utf-8")
        out = np.zeros((int(height), int(width), 4), dtype=np.uint8)
        stats = RenderStats()
        ipt = int(iterations_per_thread or 0)
        flags = ((_abi.RENDER_3D_AS_2D if as_2d else 0)
                 | (_abi.RENDER_TRANSPARENT if transparent else 0)
                 | (_abi.RENDER_PREMULTIPLIED if premultiplied else 0)
                 | (_abi.RENDER_SUPERSAMPLE_2X if supersample else 0)
                 | (_abi.RENDER_FILL_GPU if fill_gpu else 0))
        self._check(self.lib.ffm_render_config(
            self.ctx, raw, len(raw), int(width), int(height), int(iterations or 0), ipt, flags,
            out.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(stats)))
        return Image(out, {"width": int(width), "height": int(height),
                           "iterations": int(stats.total_samples), "ms": stats.render_time_ms,
                           "iterations_per_thread": ipt or 1024,
                           "dispatches": int(stats.dispatches)}, scene=load(text))

    def _render_job(self, job):
        """RGBA8 (H, W, 4) uint8 pixels of the current flame; returns (pixels, stats)."""
        w, h = job.view.width, job.view.height
        out = np.zeros((h, w, 4), dtype=np.uint8)
        stats = RenderStats()
        self._check(self.lib.ffm_render(self.ctx, C.byref(job),
                                        out.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(stats)))
        return out, stats

    def render_to_device(self, job, rgba):
        """Render into a device buffer (a (H, W, 4) uint8 tensor); returns stats."""
        stats = RenderStats()
        self._check(self.lib.ffm_render_to_device(self.ctx, C.byref(job), _ptr(rgba),
                                                  C.byref(stats)))
        return stats

    def render_accum(self, job):
        """The RGBA32F accumulator (mean colour, count) after the render."""
        w, h = job.view.width, job.view.height
        out = np.zeros((h, w, 4), dtype=np.float32)
        stats = RenderStats()
        self._check(self.lib.ffm_render_accum(self.ctx, C.byref(job),
                                              out.ctypes.data_as(C.POINTER(C.c_float)),
                                              C.byref(stats)))
        return out, stats

    # -- batched gallery
    def _batch_args(self, items, device=False):
        n = len(items)
        flames = (FlameDesc * n)()
        jobs = (Job * n)()
        stops = []
        outs = []
        for i, (fd, pal, job) in enumerate(items):
            C.memmove(C.byref(flames[i]), C.byref(fd), C.sizeof(FlameDesc))
            C.memmove(C.byref(jobs[i]), C.byref(job), C.sizeof(Job))
            stops.append(palette_stops(pal))
            shape = (job.view.height, job.view.width, 4)
            if device:
                import torch
                outs.append(torch.empty(shape, dtype=torch.uint8, device="cuda:%d" % self.device))
            else:
                outs.append(np.zeros(shape, dtype=np.uint8))
        sp = (C.POINTER(C.c_float) * n)(*[a.ctypes.data_as(C.POINTER(C.c_float)) for a in stops])
        ns = (C.c_int * n)(*[a.shape[0] for a in stops])
        ptrs = [C.cast(C.c_void_p(o.data_ptr()), C.POINTER(C.c_uint8)) if device
                else o.ctypes.data_as(C.POINTER(C.c_uint8)) for o in outs]
        op = (C.POINTER(C.c_uint8) * n)(*ptrs)
        return n, flames, sp, ns, jobs, op, outs, stops

    def batch_render(self, items, serial=False, device=False):
        """Render [(flame_desc, palette_stops, job), ...]: one persistent launch
        per wave in a hipGraph (serial=False) or one ffm_render per flame
        (serial=True). Returns ([rgba arrays], stats); with device=True the
        pictures stay on the GPU ([uint8 tensors on the engine's device])."""
        n, flames, sp, ns, jobs, op, outs, keep = self._batch_args(items, device)
        stats = RenderStats()
        if device:
            fn = (self.lib.ffm_batch_render_serial_to_device if serial
                  else self.lib.ffm_batch_render_to_device)
        else:
            fn = self.lib.ffm_batch_render_serial if serial else self.lib.ffm_batch_render
        self._check(fn(self.ctx, n, flames, sp, ns, jobs, op, C.byref(stats)))
        return outs, stats

    # -- differentiable soft splat (device tensors; see pyfae.diff for autograd)
    def softsplat_set(self, fits):
        """Set [(flame_desc, soft_job), ...]; returns (theta floats, density floats)."""
        n = len(fits)
        flames = (FlameDesc * n)()
        jobs = (SoftJob * n)()
        for i, (fd, job) in enumerate(fits):
            C.memmove(C.byref(flames[i]), C.byref(fd), C.sizeof(FlameDesc))
            C.memmove(C.byref(jobs[i]), C.byref(job), C.sizeof(SoftJob))
        self._check(self.lib.ffm_softsplat_set(self.ctx, n, flames, jobs))
        t, d = C.c_uint64(), C.c_uint64()
        self._check(self.lib.ffm_softsplat_sizes(self.ctx, C.byref(t), C.byref(d)))
        return t.value, d.value

    def softsplat_forward(self, theta, density, stream=None):
        self._check(self.lib.ffm_softsplat_forward(self.ctx, _ptr(theta), _ptr(density),
                                                   _ptr(stream)))

    def softsplat_adjoint(self, theta, grad_density, grad_theta, stream=None):
        self._check(self.lib.ffm_softsplat_adjoint(self.ctx, _ptr(theta), _ptr(grad_density),
                                                   _ptr(grad_theta), _ptr(stream)))

    # -- granular kernels on device buffers
    def iterate(self, view, walkers, ipt, burn_in, seed, hist, *, orbits=None, fresh=True,
                splat_mode=0, plotted=None, inframe=None, stream=None):
        self._check(self.lib.ffm_iterate(self.ctx, C.byref(view), int(walkers), int(ipt),
                                         int(burn_in), int(seed) & 0xFFFFFFFF, _ptr(orbits),
                                         int(bool(fresh)), _ptr(hist), int(splat_mode),
                                         _ptr(plotted), _ptr(inframe), _ptr(stream)))

    def iterate_noalloc(self, view, walkers, ipt, burn_in, seed, hist, *, orbits=None,
                        fresh=True, splat_mode=2, plotted=None, inframe=None, stream=None):
        """iterate() through ffm_iterate_noalloc: the splat's scratch is the
        library's per-GPU device array, nothing is allocated."""
        self._check(self.lib.ffm_iterate_noalloc(
            self.ctx, C.byref(view), int(walkers), int(ipt), int(burn_in),
            int(seed) & 0xFFFFFFFF, _ptr(orbits), int(bool(fresh)), _ptr(hist),
            int(splat_mode), _ptr(plotted), _ptr(inframe), _ptr(stream)))

    def downsample2x(self, src, dst, out_width, out_height, stream=None):
        """Box (2w, 2h) RGBA8 `src` down into (w, h) `dst` with the firefly
        clamp (device tensors)."""
        self._check(self.lib.ffm_downsample2x(self.ctx, _ptr(src), _ptr(dst), int(out_width),
                                              int(out_height), _ptr(stream)))

    def trace(self, view, walker, seed, burn_in, steps, out, stream=None):
        self._check(self.lib.ffm_trace(self.ctx, C.byref(view), int(walker), int(seed) & 0xFFFFFFFF,
                                       int(burn_in), int(steps), _ptr(out), _ptr(stream)))

    def blur(self, hist, scratch, width, height, sigma, density_sigma, stream=None):
        self._check(self.lib.ffm_blur(self.ctx, _ptr(hist), _ptr(scratch), width, height,
                                      sigma, density_sigma, _ptr(stream)))

    def density_estimate(self, acc_in, acc_out, width, height, max_r, min_r, curve, stream=None):
        self._check(self.lib.ffm_density_estimate(self.ctx, _ptr(acc_in), _ptr(acc_out), width,
                                                  height, max_r, min_r, curve, _ptr(stream)))

    def accumulate(self, hist, prev, out, width, height, *, hist_is_f32=False, count=None, mode=0,
                   blend=0.0, stream=None):
        self._check(self.lib.ffm_accumulate(self.ctx, _ptr(hist), int(hist_is_f32), _ptr(prev),
                                            _ptr(out), _ptr(count), width, height, mode, blend,
                                            _ptr(stream)))

    def tonemap(self, accum, rgba, width, height, zoom, samples_in_buffer, params=None,
                stream=None):
        tp = params if params is not None else tonemap_defaults()
        self._check(self.lib.ffm_tonemap(self.ctx, _ptr(accum), _ptr(rgba), width, height, zoom,
                                         float(samples_in_buffer), C.byref(tp), _ptr(stream)))
