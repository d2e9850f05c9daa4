"""The five synthetic workloads of this path, generated deterministically.

Flame parameters come from a splitmix64 stream seeded with 1000 + the
workload index (walker streams are separate: rng_init(walker, seed_k) with
seed_k = 1 + k * 2654435761). Every flame uses only the reference's own
parameter API (transforms, variation weights and params, final transform,
palette stops) so the same descriptor drives the GPU and the oracle.

  plain      3 xforms, linear, 512x512, 2^16 walkers x 1000 iterations
  single4k   6 xforms + final, 3 variations each, 3840x2160, 2^24 x 256
  gallery    512 flames, 2..12 xforms, 128^2..1024^2, 200 samples/pixel
  fit        4 xforms {linear, spherical, sinusoidal}, 256x256 soft splat
  fit_batch  64 independent 3-xform soft-splat fits at 128x128
  stress     12 xforms incl. julian/curl/waves, 12288^2 f32 histogram
  blurs      the reference presets' other variations (pre_blur, blur, noise,
             gaussian_blur, blob) beside julian/bubble/spherical, 640x480
"""
import math

import numpy as np

from . import flame_desc, make_job, soft_job, tonemap_defaults, xform

SEED_STRIDE = 2654435761
C1_VARIATIONS = ["linear", "sinusoidal", "spherical", "swirl", "horseshoe", "polar", "julia",
                 "bubble", "hyperbolic", "disc"]
C4_VARIATIONS = C1_VARIATIONS + ["julian", "curl", "waves"]
FIT_VARIATIONS = ["linear", "spherical", "sinusoidal"]
BLUR_VARIATIONS = ["pre_blur", "blur", "noise", "gaussian_blur", "blob", "julian", "bubble",
                   "spherical", "linear"]


class SplitMix64:
    """Vigna's splitmix64 (public domain reference algorithm)."""

    def __init__(self, seed):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next() >> 11) * (1.0 / 9007199254740992.0))

    def normal(self, sigma=1.0):
        u1 = max(self.uniform(), 1e-300)
        u2 = self.uniform()
        return sigma * math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)

    def randint(self, lo, hi):
        """Uniform integer in lo..hi inclusive."""
        return lo + int(self.uniform() * (hi - lo + 1))


def _affine(rng, norm):
    a, b, c, d, e, f = (rng.uniform(-1.0, 1.0) for _ in range(6))
    s = np.linalg.svd(np.array([[a, b], [c, d]]), compute_uv=False)[0]
    k = norm / s if s > 0 else 1.0
    return a * k, b * k, c * k, d * k, e, f


def _variations(rng, pool, n=3):
    names = []
    while len(names) < n:
        v = pool[int(rng.uniform() * len(pool))]
        if v not in names:
            names.append(v)
    w = [rng.uniform() for _ in names]
    s = sum(w) or 1.0
    return {nm: wi / s for nm, wi in zip(names, w)}


def _params(rng, names):
    p = {}
    if "julian" in names:
        p["julian.power"] = float(rng.randint(2, 5)) * (1 if rng.uniform() < 0.8 else -1)
        p["julian.dist"] = rng.uniform(0.5, 1.5)
    if "curl" in names:
        p["curl.c1"] = rng.uniform(-1.0, 1.0)
        p["curl.c2"] = rng.uniform(-0.5, 0.5)
    if "blob" in names:
        p["blob.high"] = rng.uniform(0.5, 1.5)
        p["blob.low"] = rng.uniform(0.0, 0.8)
        p["blob.waves"] = float(rng.randint(2, 8))
    return p


def random_palette(rng, n=256):
    """n stops at i/(n-1), RGB uniform in [0, 1] (the XML import's layout,
    src/flame_xml.rs:1049-1067)."""
    return np.array([[np.float32(i) / np.float32(n - 1), rng.uniform(), rng.uniform(), rng.uniform()]
                     for i in range(n)], dtype=np.float32)


def random_flame(rng, n_xforms, pool, norm=0.8, final=True, post=True):
    w = [-math.log(max(rng.uniform(), 1e-300)) for _ in range(n_xforms)]  # Dirichlet(1)
    sw = sum(w)
    xs = []
    for i in range(n_xforms):
        var = _variations(rng, pool)
        pst = None
        if post:
            pst = (1.0 + rng.normal(0.1), rng.normal(0.1), rng.normal(0.1), 1.0 + rng.normal(0.1),
                   rng.normal(0.1), rng.normal(0.1))
        xs.append(xform(*_affine(rng, norm), weight=w[i] / sw,
                        color=i / max(n_xforms - 1, 1), color_speed=0.5, variations=var,
                        params=_params(rng, var), post=pst))
    fins = []
    if final:
        var = _variations(rng, pool)
        fins.append(xform(*_affine(rng, norm), variations=var, params=_params(rng, var)))
    return xs, fins


def log_tonemap(gamma=2.2):
    t = tonemap_defaults()
    t.gamma = gamma
    t.vibrancy = 1.0
    t.levels_enabled = 0
    return t


# ---------------------------------------------------------------- workloads

def plain(walkers=1 << 16, ipt=1000):
    """3 xforms, weights 1/3, affines U[-1,1] at spectral norm 0.7, linear
    weight 1, colour i/2, colour speed 0.5; 512x512, camera [-1.5,1.5]^2."""
    rng = SplitMix64(1000)
    xs = []
    for i in range(3):
        xs.append(xform(*_affine(rng, 0.7), weight=1.0 / 3.0, color=i / 2.0, color_speed=0.5,
                        variations={"linear": 1.0}))
    pal = random_palette(rng)
    job = make_job(512, 512, walkers=walkers, ipt=ipt, burn_in=20, seed_base=1,
                   seed_stride=SEED_STRIDE, target_samples=0, zoom=4.0 / 3.0,
                   tonemap=log_tonemap(2.2))
    return flame_desc(xs), pal, job


def single4k(walkers=1 << 24, ipt=256, width=3840, height=2160, splat_mode=0):
    """6 xforms + final, 3 of 10 variations each, post = I + N(0,0.1);
    3840x2160; density estimation radius 9 -> 0, curve 0.4."""
    rng = SplitMix64(1001)
    xs, fins = random_flame(rng, 6, C1_VARIATIONS, norm=0.8)
    pal = random_palette(rng)
    job = make_job(width, height, walkers=walkers, ipt=ipt, burn_in=20, seed_base=1,
                   seed_stride=SEED_STRIDE, target_samples=0, zoom=4.0 / 3.0,
                   splat_mode=splat_mode, de=(9.0, 0.0, 0.4), tonemap=log_tonemap(2.2))
    return flame_desc(xs, fins), pal, job


def blurs(walkers=1 << 16, ipt=256, width=640, height=480):
    """5 xforms + final over BLUR_VARIATIONS (3 each; pre_blur, when drawn,
    runs before the others whatever its local index)."""
    rng = SplitMix64(1006)
    xs, fins = random_flame(rng, 5, BLUR_VARIATIONS, norm=0.7)
    pal = random_palette(rng)
    job = make_job(width, height, walkers=walkers, ipt=ipt, burn_in=20, seed_base=1,
                   seed_stride=SEED_STRIDE, zoom=1.0, tonemap=log_tonemap(2.2))
    return flame_desc(xs, fins), pal, job


def gallery(n=512, spp=200.0, ipt=256):
    """n independent flames; flame k from splitmix64(1002 + k): 2..12
    xforms, sides log-uniform in 128..1024, `spp` samples per pixel."""
    out = []
    for k in range(n):
        rng = SplitMix64(1002 * 100003 + k)
        nx = rng.randint(2, 12)
        xs, fins = random_flame(rng, nx, C1_VARIATIONS, norm=0.8)
        side = int(round(math.exp(rng.uniform(math.log(128), math.log(1024)))))
        pal = random_palette(rng)
        walkers = max(64, int(math.ceil(spp * side * side / (ipt - 20))))
        job = make_job(side, side, walkers=walkers, ipt=ipt, burn_in=20,
                       seed_base=(1 + k) & 0xFFFFFFFF, seed_stride=SEED_STRIDE, zoom=4.0 / 3.0,
                       tonemap=log_tonemap(2.2))
        out.append((flame_desc(xs, fins), pal, job))
    return out


def _fit_flame(rng, n_xforms, perturb_rng, perturb):
    xs, ts = [], []
    for i in range(n_xforms):
        aff = list(_affine(rng, 0.8))
        var = {nm: rng.uniform() for nm in FIT_VARIATIONS}
        xs.append(xform(*aff, weight=1.0 / n_xforms, color=i / max(n_xforms - 1, 1),
                        color_speed=0.5, variations=var))
        taff = [v + perturb_rng.normal(perturb) for v in aff]
        ts.append(xform(*taff, weight=1.0 / n_xforms, color=i / max(n_xforms - 1, 1),
                        color_speed=0.5, variations=dict(var)))
    return flame_desc(xs), flame_desc(ts)


def fit(walkers=1 << 20, ipt=64, side=256, window=8, truncate=False, seed=1003):
    """Differentiable fit: 4 xforms over {linear, spherical, sinusoidal} with
    weights U[0,1], affines at spectral norm 0.8, selection weights 1/4;
    side^2 soft-splat density of walkers x ipt (burn-in 20), adjoint window K.
    The target is the same flame with affines + N(0, 0.05), rendered at
    walker seed 2 (the fit walks seed 1).
    Returns (flame, target_flame, job, target_job)."""
    fd, tfd = _fit_flame(SplitMix64(seed), 4, SplitMix64(seed + 7777), 0.05)
    kw = dict(walkers=walkers, ipt=ipt, burn_in=20, window=window, truncate=truncate,
              zoom=4.0 / 3.0)
    return fd, tfd, soft_job(side, side, seed=1, **kw), soft_job(side, side, seed=2, **kw)


def fit_batch(n=64, side=128, walkers=1 << 16, ipt=64, window=8):
    """n independent 3-xform fits at side^2 (fit k from splitmix64(1004 * 100003 + k)),
    each as `fit` builds them. Returns [(flame, target_flame, job, target_job)]."""
    out = []
    for k in range(n):
        fd, tfd = _fit_flame(SplitMix64(1004 * 100003 + k), 3,
                             SplitMix64(1004 * 100003 + k + 7777), 0.05)
        kw = dict(walkers=walkers, ipt=ipt, burn_in=20, window=window, zoom=4.0 / 3.0)
        out.append((fd, tfd, soft_job(side, side, seed=1 + k, **kw),
                    soft_job(side, side, seed=100001 + k, **kw)))
    return out


def stress(walkers=1 << 26, ipt=256, side=12288):
    """12 xforms over the 13 variations, side^2 f32 histogram."""
    rng = SplitMix64(1004)
    xs, fins = random_flame(rng, 12, C4_VARIATIONS, norm=0.8)
    pal = random_palette(rng)
    job = make_job(side, side, walkers=walkers, ipt=ipt, burn_in=20, seed_base=1,
                   seed_stride=SEED_STRIDE, zoom=4.0 / 3.0, splat_mode=3,
                   tonemap=log_tonemap(2.2))
    return flame_desc(xs, fins), pal, job
