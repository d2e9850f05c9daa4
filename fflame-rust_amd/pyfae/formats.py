"""Host-side formats around the render: the PNG the reference exports, with
its metadata text chunks (src/png_metadata.rs:144-300: the render's numbers
and the full config as JSON, so a picture reloads as its flame), and an
ffm_scene written back as FractalConfig JSON (the canonical `.fflame` form,
version 7, src/config/fractal_config.rs:1683-1740), which is how a `.flame`
import is saved and what the PNG's `Config` chunk holds.
"""
import ctypes as C
import hashlib
import json
import struct
import zlib

import numpy as np

from . import _abi

PNG_SIGNATURE = b"\x89PNG\r\n\x1a\n"
CONFIG_VERSION = 7
SOFTWARE = "Fractal Art Editor v0.5.0+mi355x"

_TONEMAP = ("linear", "logarithmic", "density")
_HIGHLIGHT = ("clip", "max_norm", "reinhard", "filmic")


# ------------------------------------------------------------------ PNG

def _chunk(tag, data):
    return (struct.pack(">I", len(data)) + tag + data
            + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF))


def _text_chunk(key, value):
    """tEXt (Latin-1) when the text is Latin-1, else an uncompressed iTXt."""
    k = key.encode("latin-1")
    try:
        return _chunk(b"tEXt", k + b"\x00" + value.encode("latin-1"))
    except UnicodeEncodeError:
        return _chunk(b"iTXt", k + b"\x00\x00\x00\x00\x00" + value.encode("utf-8"))


def encode_png(rgba, text=()):
    """An 8-bit RGBA PNG of an (h, w, 4) uint8 array with the given
    (keyword, text) chunks before the image data."""
    rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
    h, w = rgba.shape[:2]
    raw = b"".join(b"\x00" + rgba[y].tobytes() for y in range(h))
    out = [PNG_SIGNATURE, _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0))]
    out += [_text_chunk(k, v) for k, v in text]
    out += [_chunk(b"IDAT", zlib.compress(raw, 6)), _chunk(b"IEND", b"")]
    return b"".join(out)


def png_chunks(data):
    """[(type, payload)] of a PNG, CRCs checked."""
    if data[:8] != PNG_SIGNATURE:
        raise ValueError("not a PNG")
    out, i = [], 8
    while i + 12 <= len(data):
        n = struct.unpack(">I", data[i:i + 4])[0]
        tag = data[i + 4:i + 8]
        payload = data[i + 8:i + 8 + n]
        crc = struct.unpack(">I", data[i + 8 + n:i + 12 + n])[0]
        if zlib.crc32(tag + payload) & 0xFFFFFFFF != crc:
            raise ValueError(f"bad CRC in {tag!r} chunk")
        out.append((tag, payload))
        i += 12 + n
        if tag == b"IEND":
            break
    return out


def png_text(data):
    """{keyword: text} of a PNG's tEXt and uncompressed iTXt chunks."""
    out = {}
    for tag, p in png_chunks(data):
        if tag == b"tEXt":
            k, _, v = p.partition(b"\x00")
            out[k.decode("latin-1")] = v.decode("latin-1")
        elif tag == b"iTXt":
            k, _, rest = p.partition(b"\x00")
            if rest[:1] == b"\x00":  # not compressed
                lang_and_more = rest[2:]
                _, _, more = lang_and_more.partition(b"\x00")
                _, _, text = more.partition(b"\x00")
                out[k.decode("latin-1")] = text.decode("utf-8")
    return out


def read_png_metadata(data):
    """The metadata read_png_metadata (src/png_metadata.rs:264-340) returns."""
    t = png_text(data)
    sw = t.get("Software", "")
    version = "unknown"
    for pre in ("Fractal Art Editor v", "Fractal Flame Renderer v"):
        if sw.startswith(pre):
            version = sw[len(pre):]
    w, _, h = t.get("Resolution", "0x0").partition("x")

    def num(key, kind, dflt):
        try:
            return kind(t[key].rstrip("ms") if key == "RenderTime" else t[key])
        except (KeyError, ValueError):
            return dflt
    cs = t.get("ConfigChecksum", "")
    for pre in ("sha256:", "hash:"):
        if cs.startswith(pre):
            cs = cs[len(pre):]
    return {
        "version": version, "width": int(w) if w.isdigit() else 0,
        "height": int(h) if h.isdigit() else 0,
        "total_iterations": num("Iterations", int, 0),
        "render_time_ms": num("RenderTime", float, 0.0),
        "iterations_per_thread": num("IterationsPerThread", int, 256),
        "speed_factor": num("SpeedFactor", float, 1.0),
        "config_json": t.get("Config", ""), "config_checksum": cs,
        "use_tone_curve": t.get("ToneCurve") == "true",
        "render_mode": t.get("RenderMode", ""), "projection": t.get("Projection", ""),
    }


def metadata_chunks(width, height, iterations, render_ms, ipt, config_json, sc):
    """PngMetadata::to_text_chunks (src/png_metadata.rs:144-190) for a render
    of scene `sc` whose config is `config_json`."""
    tp = sc.tonemap
    bg = ",".join(repr(float(v)) for v in tp.background)
    return [
        ("Software", SOFTWARE),
        ("Resolution", f"{width}x{height}"),
        ("Iterations", str(int(iterations))),
        ("RenderTime", f"{render_ms:.2f}ms"),
        ("IterationsPerThread", str(int(ipt))),
        ("SpeedFactor", "1.00"),
        ("ConfigChecksum", "sha256:" + hashlib.sha256(config_json.encode("utf-8")).hexdigest()),
        ("Config", config_json),
        ("BackgroundColor", bg),
        ("Exposure", repr(float(tp.exposure))),
        ("Gamma", repr(float(tp.gamma))),
        ("ToneCurve", "false"),
        ("RenderMode", "3D" if sc.render_3d else "2D"),
        ("Projection", "Orthographic"),
        ("ShaderVariations", str(int(sc.flame.n_vars))),
    ]


def config_from_png(data):
    """The `Config` JSON a metadata PNG carries (how the app reopens one)."""
    cfg = png_text(data).get("Config")
    if not cfg:
        raise ValueError("the PNG carries no Config chunk")
    return cfg


# ----------------------------------------------------- scene -> FractalConfig

def _f(v):
    return float(np.float32(v))


def _xform_json(fd, x, final_att=None):
    # every slot of the flame, zero weights included, with the flame's order
    # as the transform's variation_order: the reloaded flame then has the
    # same slots in the same order (a zero weight is skipped, so it changes
    # nothing else)
    names = [_abi.VARIATION_NAMES[fd.var_id[k]] for k in range(fd.n_vars)]
    variations, params, order = {}, {}, []
    for k, nm in enumerate(names):
        variations[nm] = _f(x.var_weight[k])
        order.append(nm)
        for q, pn in enumerate(_abi.VARIATION_PARAMS.get(nm, ())):
            params[f"{nm}.{pn}"] = _f(x.var_param[k][q])
    t = {"a": _f(x.a), "b": _f(x.b), "c": _f(x.c), "d": _f(x.d), "e": _f(x.e), "f": _f(x.f),
         "g": 0.0, "weight": _f(x.weight), "variations": variations, "variation_params": params,
         "variation_order": order, "color": _f(x.color), "color_speed": _f(x.color_speed),
         "opacity": _f(x.opacity)}
    if x.post_enabled:
        t.update(post_affine_enabled=True, post_a=_f(x.post_a), post_b=_f(x.post_b),
                 post_c=_f(x.post_c), post_d=_f(x.post_d), post_e=_f(x.post_e),
                 post_f=_f(x.post_f), post_g=0.0)
    if final_att is not None:
        t["final_attachments"] = final_att
    return t


def scene_to_config(sc, palette_name="Imported"):
    """A scene as FractalConfig JSON (a dict, version 7): every field this
    path renders, explicit; the reloaded flame has the same variation slots
    in the same order, so it renders the same pixels."""
    fd = sc.flame
    finals = [_xform_json(fd, fd.finals[i]) for i in range(fd.n_finals)]
    xs = []
    for i in range(fd.n_xforms):
        x = fd.xforms[i]
        att = None
        if fd.n_finals:
            att = (list(x.final_attach[:x.n_final_attach]) if fd.per_transform_finals
                   else list(range(fd.n_finals)))
        xs.append(_xform_json(fd, x, att))
    flame = {"name": sc.name.decode("utf-8", "replace"), "transforms": xs}
    if finals:
        flame["final_transforms"] = finals
    if fd.has_xaos:
        n = fd.n_xforms
        flame["xaos"] = [[_f(fd.xaos[i][j]) for j in range(n)] for i in range(n)]
    tp = sc.tonemap
    xf = sc.palette_xf
    stops = [{"position": _f(sc.stops[i][0]),
              "color": [_f(sc.stops[i][1]), _f(sc.stops[i][2]), _f(sc.stops[i][3])]}
             for i in range(sc.n_stops)]
    geometric = xf.squeeze_mode == 1
    return {
        "version": CONFIG_VERSION,
        "flame": flame,
        "render_mode": "3d" if sc.render_3d else "2d",
        "preserve_z": False,
        "reset_orbits": bool(sc.reset_orbits),
        "zoom": _f(sc.zoom), "pan_x": float(sc.pan_x), "pan_y": float(sc.pan_y),
        "rotation": _f(sc.rotation),
        "image_size": [int(sc.image_width), int(sc.image_height)],
        "filter_radius": _f(sc.filter_radius), "filter_blur_edges": _f(sc.filter_blur_edges),
        "density_scale": _f(tp.density_scale),
        "max_iterations": int(sc.max_iterations),
        "color_mode": "palette",
        "palette": {"name": palette_name, "stops": stops, "locked": False},
        "palette_size": int(sc.palette_size),
        "palette_squeeze_mode": "geometric" if geometric else "linear",
        "palette_squeeze": 1.0 if geometric else _f(xf.squeeze_factor),
        "palette_squeeze_falloff": _f(xf.squeeze_factor) if geometric else 0.5,
        "palette_log_strength": _f(xf.log_strength),
        "palette_rotation": _f(xf.rotation),
        "palette_reverse": bool(xf.reverse),
        "background_color": [_f(v) for v in tp.background],
        "tonemap_mode": _TONEMAP[tp.tonemap_mode],
        "highlight_mode": _HIGHLIGHT[tp.highlight_mode],
        "use_curve": bool(tp.use_curve),
        **({"tonemap_curve": {"points": [{"x": _f(tp.curve_x[i]), "y": _f(tp.curve_y[i])}
                                         for i in range(tp.curve_n)]}} if tp.curve_n else {}),
        "exposure": _f(tp.exposure), "gamma": _f(tp.gamma),
        "gamma_threshold": _f(tp.gamma_threshold), "brightness": _f(tp.brightness),
        "vibrancy": _f(tp.vibrancy), "white_level": _f(tp.white_level),
        "saturation": _f(tp.saturation), "hue_shift": _f(tp.hue_shift),
        "alpha_blend_low": _f(tp.alpha_blend_low), "alpha_blend_high": _f(tp.alpha_blend_high),
        "levels_enabled": bool(tp.levels_enabled), "levels_low": _f(tp.levels_low),
        "levels_high": _f(tp.levels_high), "levels_gamma": _f(tp.levels_gamma),
        "deterministic_rng": bool(sc.deterministic_rng),
        "density_effects": _effects_json(sc.effects.density, sc.effects.n_density),
        "color_effects": _effects_json(sc.effects.color, sc.effects.n_color),
    }


def _effects_json(lst, n):
    """An effect chain as the config's EffectInstance list (effect_type,
    enabled, params by name; src/effects/mod.rs:319-336)."""
    out = []
    for e in lst[:n]:
        name = next(k for k, (t, _, _) in _abi.EFFECTS.items() if t == e.type)
        names = _abi.EFFECTS[name][2]
        out.append({"effect_type": name, "enabled": bool(e.enabled),
                    "params": {k: _f(e.params[i]) for i, k in enumerate(names)}})
    return out


def scene_to_json(sc, **kw):
    """scene_to_config as compact JSON text."""
    return json.dumps(scene_to_config(sc, **kw), separators=(",", ":"))


def scene_summary(sc):
    """What a scene renders, independent of slot layout: per transform its
    affine, post, colour terms and {variation: (weight, params)} of non-zero
    weights in order; finals, attachments, xaos, palette, view, tone map.
    Two scenes with equal summaries render the same pixels."""
    fd = sc.flame

    def xs(x):
        names = [_abi.VARIATION_NAMES[fd.var_id[k]] for k in range(fd.n_vars)]
        v = [(nm, x.var_weight[k], tuple(x.var_param[k][:len(_abi.VARIATION_PARAMS.get(nm, ()))]))
             for k, nm in enumerate(names) if x.var_weight[k] != 0.0]
        post = (x.post_a, x.post_b, x.post_c, x.post_d, x.post_e, x.post_f) if x.post_enabled else None
        return ((x.a, x.b, x.c, x.d, x.e, x.f), post, x.weight, x.color, x.color_speed, x.opacity, v)
    att = [tuple(fd.xforms[i].final_attach[:fd.xforms[i].n_final_attach]) if fd.per_transform_finals
           else tuple(range(fd.n_finals)) for i in range(fd.n_xforms)]
    xa = ([[fd.xaos[i][j] for j in range(fd.n_xforms)] for i in range(fd.n_xforms)]
          if fd.has_xaos else None)
    tp = sc.tonemap
    return {
        "xforms": [xs(fd.xforms[i]) for i in range(fd.n_xforms)],
        "finals": [xs(fd.finals[i]) for i in range(fd.n_finals)],
        "attach": att if fd.n_finals else None, "xaos": xa,
        "stops": np.ctypeslib.as_array(sc.stops)[:sc.n_stops].tolist(),
        "palette": (sc.palette_size, bytes(sc.palette_xf)),
        "view": (sc.zoom, sc.pan_x, sc.pan_y, sc.rotation),
        "render": (sc.filter_radius, sc.filter_blur_edges, sc.max_iterations, sc.reset_orbits,
                   sc.deterministic_rng, sc.render_3d),
        "tonemap": bytes(C.string_at(C.addressof(tp), C.sizeof(tp))),
        "effects": bytes(C.string_at(C.addressof(sc.effects), C.sizeof(sc.effects))),
    }
