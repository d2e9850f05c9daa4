"""Chaos-game throughput on one MI355X, with every metric of the workloads.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Default (the headline): a step is one whole render of the single4k workload
through libffm.so: 6 transforms + a final (3 variations each), 3840x2160,
2^24 walkers x 256 iterations (one dispatch), accumulate, density estimation
(radius 9 -> 0, curve 0.4), the log tonemap and the 2x box filter with the
firefly clamp down to 1920x1080, pixels left in HBM. `value`
is chaos-game samples (walker iterations) per second over exactly K timed
steps after W warmup steps.

After the timed steps (never inside them) the same JSON line carries the
other workloads' figures:
  splat_compare  the 4K render with the reference's global u32 atomics
                 against the LDS tile splat;
  plain          3 linear transforms, 512x512, 2^16 x 1000: samples/s;
  gallery        512 flames batched (one hipGraph) and one by one: flames/s
                 and the ratio;
  fit            the differentiable fit step (forward + loss + adjoint):
                 step ms, forward ms, step/forward, for the exact and the
                 truncated (K = 8) adjoint, each against FP64 central
                 differences of the oracle's loss at h = 1e-10 and 1e-4;
  cpu_baseline   the oracle's chaos game on this host's cores over a
                 bounded sample of the same flame and frame.
Progress goes to stderr; stdout is the ONE JSON line.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fflame-rust_amd"))

# Worker threads of the CPU legs: the cores this process may run on, capped
# at the share a one-GPU box grants a job (16).
CPU_SHARE = 16


def progress(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def host_cpu():
    model = platform.processor() or "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"model": model, "logical_cpus": os.cpu_count(), "usable_cpus": usable,
            "threads": max(1, min(usable, CPU_SHARE))}


def cpu_baseline(fd, pal, job, seconds=10.0, single_seconds=4.0):
    """The oracle (a scalar C restatement of the reference's step, built
    -O3 -march=x86-64-v3) timed on this host: calibrate on a small sample,
    then about `seconds` of work on every usable core (up to the box's
    share) and `single_seconds` on one. Beside it, the reference's Criterion
    microbenchmarks (benches/flame_bench.rs) restated in C, with the per-op
    means the reference publishes."""
    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    cpu = host_cpu()
    threads = cpu["threads"]
    v = job.view

    def timed(threads, seconds):
        n = 1 << 12
        t0 = time.perf_counter()
        O.iterate_mt(fd, pal, v, n * threads, job.ipt, job.burn_in, 1, threads)
        dt = time.perf_counter() - t0
        n = int(max(1 << 12, min(1 << 24, n * threads * seconds / max(dt, 1e-3))))
        n = max(threads, (n // threads) * threads)
        t0 = time.perf_counter()
        O.iterate_mt(fd, pal, v, n, job.ipt, job.burn_in, 1, threads)
        return n, time.perf_counter() - t0

    n, dt = timed(threads, seconds)
    n1, dt1 = timed(1, single_seconds)
    return {"value": n * job.ipt / dt, "unit": "samples/s", "cores": threads, "kind": "port",
            "sample": f"{n} walkers x {job.ipt} iterations of the same flame at "
                      f"{v.width}x{v.height} (iterate + splat), {dt:.1f} s, {threads} threads",
            "host": cpu, "build": "gcc -O3 -march=x86-64-v3 -ffp-contract=off",
            "single_core": {"value": n1 * job.ipt / dt1, "unit": "samples/s", "cores": 1,
                            "sample": f"{n1} walkers x {job.ipt} iterations, {dt1:.1f} s"},
            "microbench_ns": O.microbench_suite()}


def recorded_traffic(kernel, splat_mode, walkers):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, calibrated on a 1 GiB copy), when it
    was measured on this exact configuration; else None."""
    path = os.path.join(ROOT, "profiles", "bin_iterate_single4k_hbm_traffic.md")
    try:
        for line in open(path):
            if line.startswith("TRAFFIC_JSON "):
                d = json.loads(line[len("TRAFFIC_JSON "):])
                if (d["kernel"], d["splat_mode"], d["walkers"]) == (kernel, splat_mode, walkers):
                    return d["bytes_per_launch"]
    except OSError:
        pass
    return None


# ------------------------------------------------------------- one flame

def time_renders(eng, job, rgba, steps, warmup, out=None):
    """Render `warmup` then exactly `steps` times (each boxed down 2x into
    `out` when given); (seconds, [stats])."""
    import torch

    def step():
        st = eng.render_to_device(job, rgba)
        if out is not None:  # on the engine's stream, after the render's tonemap
            eng.downsample2x(rgba, out, out.shape[1], out.shape[0])
        return st

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for _ in range(steps):
        stats.append(step())
    torch.cuda.synchronize()
    return time.perf_counter() - t0, stats


def bench_flame(args, eng, fd, pal, job, workload, metric, supersample=False):
    import torch
    from pyfae import roofline
    W, H = job.view.width, job.view.height
    eng.set_flame(fd)
    eng.set_palette(pal, 256)
    rgba = torch.empty((H, W, 4), dtype=torch.uint8, device="cuda:0")
    out_px = (torch.empty((H // 2, W // 2, 4), dtype=torch.uint8, device="cuda:0")
              if supersample else None)
    if args.copy_calibration:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    dt, stats = time_renders(eng, job, rgba, args.steps, args.warmup, out_px)
    samples = job.walkers * job.ipt
    ms_per_step = dt / args.steps * 1e3
    value = samples / (ms_per_step / 1e3)
    # the dominant kernel: the walker kernel, timed per launch with HIP events
    # recorded around each launch on the stream it runs on
    launches = sum(s.walker_launches for s in stats)
    launch_ms = sum(s.walker_kernel_ms for s in stats) / launches
    samples_per_launch = samples * args.steps / launches
    achieved, frac, slots = roofline.iterate_roofline(fd, samples_per_launch, launch_ms / 1e3,
                                                      job.burn_in, job.ipt)
    iterate_ms = sum(s.iterate_ms for s in stats) / args.steps
    kname = "bin_iterate_kernel" if job.splat_mode in (2, 4) else "iterate_kernel"
    traffic = recorded_traffic(kname, job.splat_mode, job.walkers)
    accum_ms = sum(s.tile_accum_ms for s in stats) / args.steps
    out = {
        "metric": metric,
        "value": value,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": workload, "splat_mode": job.splat_mode,
                   "samples_per_step": samples,
                   "output": "%dx%d" % ((W // 2, H // 2) if supersample else (W, H))},
        "roofline": {"bound": "valu",
                     "kernel": kname,
                     "launch_ms": launch_ms, "samples_per_launch": samples_per_launch,
                     "achieved": achieved / 1e12, "peak": roofline.PEAK_SLOTS / 1e12,
                     "unit": "TFLOP/s", "frac": frac, "traffic": traffic,
                     "note": "VALU issue slots per lane (quarter-rate ops = 4); %.0f slots per "
                             "iteration; peak = 157.3 TFLOP/s FP32 / 4" % slots},
        "kernel_ms": {"walkers": launch_ms * launches / args.steps, "tile_accum": accum_ms,
                      "dispatch_phase": iterate_ms, "rest_of_step": ms_per_step - iterate_ms},
    }
    return out, rgba


def splat_compare(eng, job, rgba, mode):
    """One render with splat `mode` against the job's own (after a warmup of
    each), outside the timed steps: {ms of each, speedup of the job's}."""
    own = job.splat_mode
    dt_own, _ = time_renders(eng, job, rgba, 1, 1)
    job.splat_mode = mode
    try:
        dt_other, _ = time_renders(eng, job, rgba, 1, 1)
    finally:
        job.splat_mode = own
    return {"splat_mode": own, "ms": dt_own * 1e3, "other_splat_mode": mode,
            "other_ms": dt_other * 1e3, "speedup": dt_other / dt_own}


def plain_extra(eng, steps=3):
    """The plain workload: 3 linear transforms, 512x512, 2^16 walkers x 1000."""
    import torch
    from pyfae import configs
    fd, pal, job = configs.plain()
    job.splat_mode = 2
    eng.set_flame(fd)
    eng.set_palette(pal, 256)
    rgba = torch.empty((512, 512, 4), dtype=torch.uint8, device="cuda:0")
    dt, _ = time_renders(eng, job, rgba, steps, 1)
    samples = job.walkers * job.ipt
    return {"workload": "plain: 3 linear xforms, 512x512, 2^16 walkers x 1000 iterations, "
                        "log tonemap, LDS tile splat",
            "samples_per_s": samples * steps / dt, "ms_per_step": dt / steps * 1e3,
            "steps": steps}


# --------------------------------------------------------------- gallery

def gallery_run(eng, n, steps):
    """Batched flames/s (persistent launches in one hipGraph; the first call
    autotunes and builds the graph, untimed) against the same flames one by
    one, both with the pictures left in device memory; then both again with
    every picture read back into host memory (PCIe included)."""
    import torch
    from pyfae import configs
    items = configs.gallery(n=n)
    t0 = time.perf_counter()
    eng.batch_render(items, device=True)
    first = time.perf_counter() - t0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        outs, st = eng.batch_render(items, device=True)
        del outs
    dt = (time.perf_counter() - t0) / steps
    t1 = time.perf_counter()
    outs, st_s = eng.batch_render(items, serial=True, device=True)
    dt_s = time.perf_counter() - t1
    del outs
    t2 = time.perf_counter()
    eng.batch_render(items)
    dt_h = time.perf_counter() - t2
    t3 = time.perf_counter()
    eng.batch_render(items, serial=True)
    dt_sh = time.perf_counter() - t3
    px = sum(j.view.width * j.view.height for _, _, j in items)
    return {"workload": "gallery: %d flames, 2..12 xforms, sides log-uniform 128..1024, "
                        "200 samples/pixel, ipt 256, pictures left in device memory" % n,
            "flames": n, "pixels": px, "samples_per_step": st.total_samples,
            "batched_flames_per_s": n / dt, "batched_ms": dt * 1e3, "steps": steps,
            "batched_graph_gpu_ms": st.iterate_ms,
            "serial_flames_per_s": n / dt_s, "serial_ms": dt_s * 1e3,
            "speedup_vs_serial": dt_s / dt, "samples_per_s": st.total_samples / dt,
            "first_call_s": first,
            "host_buffers": {"note": "every picture read back into pageable host memory",
                             "batched_ms": dt_h * 1e3, "serial_ms": dt_sh * 1e3,
                             "speedup_vs_serial": dt_sh / dt_h}}


# ------------------------------------------------------------------- fits

def fit_step_timing(eng, items, steps, warmup):
    """One fit step = soft-splat forward, the log-density L2 loss in torch and
    the adjoint, through pyfae.diff's autograd Function, on a side stream.
    Returns (timings dict, last gradient, target density)."""
    import torch
    from pyfae.diff import SoftSplat, log_density_l2
    target = SoftSplat(eng, [(tfd, tjob) for _, tfd, _, tjob in items])
    with torch.no_grad():
        tau = target(target.theta0()).clone()
    splat = SoftSplat(eng, [(fd, job) for fd, _, job, _ in items])
    theta0 = splat.theta0()
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        def step():
            theta = theta0.clone().requires_grad_()
            loss = log_density_l2(splat(theta), tau)
            loss.backward()
            return theta.grad
        for _ in range(max(warmup, 1)):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            grad = step()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        # forward and adjoint kernels alone, HIP events on the launch stream
        dens = torch.empty(splat.density_floats, device="cuda")
        g = torch.randn(splat.density_floats, device="cuda")
        gt = torch.empty_like(theta0)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        n = 10
        ev[0].record(stream)
        for _ in range(n):
            eng.softsplat_forward(theta0, dens, stream=stream.cuda_stream)
        ev[1].record(stream)
        for _ in range(n):
            eng.softsplat_adjoint(theta0, g, gt, stream=stream.cuda_stream)
        ev[2].record(stream)
    torch.cuda.synchronize()
    fwd_ms = ev[0].elapsed_time(ev[1]) / n
    adj_ms = ev[1].elapsed_time(ev[2]) / n
    samples = sum(job.walkers * job.ipt for _, _, job, _ in items)
    return ({"step_ms": dt * 1e3, "forward_ms": fwd_ms, "adjoint_ms": adj_ms,
             "step_over_forward": dt * 1e3 / fwd_ms,
             "kernels_over_forward": (fwd_ms + adj_ms) / fwd_ms,
             "walk_samples_per_s": {"forward": samples / (fwd_ms / 1e3),
                                    "adjoint": samples / (adj_ms / 1e3)}},
            grad, tau)


def fd_reference(item, tau, params, hs=(1e-10, 1e-4)):
    """FP64 central differences of the oracle's loss L(theta) =
    sum (log1p(rho(theta)) - log1p(tau))^2 for the parameters `params`, at
    each h: {h: array}."""
    import numpy as np
    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    from pyfae import soft_theta
    fd, _, job, _ = item
    th = soft_theta(fd).astype(np.float32).astype(np.float64)
    tau = tau.reshape(job.view.height, job.view.width)
    threads = host_cpu()["threads"]

    def L(t):
        d = np.log1p(O.soft_forward(fd, t, job, threads=threads)) - np.log1p(tau)
        return float((d * d).sum())

    out = {}
    for h in hs:
        out[h] = np.array([(L(th + h * e) - L(th - h * e)) / (2 * h)
                           for e in np.eye(th.size)[list(params)]])
    return out


def fit_extra(eng, steps=5, warmup=2, fd_params_every=3, do_fd=True):
    """The fit workload with the exact and the truncated adjoint; the FD
    check on every `fd_params_every`-th parameter (FP64 oracle on the host,
    outside the timed region)."""
    import numpy as np
    from pyfae import configs
    res = {"workload": "fit: 4 xforms {linear, spherical, sinusoidal}, 256x256 soft splat, "
                       "2^20 walkers x 64 iterations, adjoint windows of K=8, f64 walk"}
    grads = {}
    tau = None
    for name, trunc in (("exact", False), ("truncated", True)):
        item = configs.fit(truncate=trunc)
        t, grad, tau = fit_step_timing(eng, [item], steps, warmup)
        t["steps"] = steps
        res[name] = t
        grads[name] = grad.cpu().numpy().astype(np.float64)
    if do_fd:
        params = list(range(0, grads["exact"].size, fd_params_every))
        t0 = time.perf_counter()
        num = fd_reference(configs.fit(), tau.cpu().numpy().astype(np.float64), params)
        for name, g in grads.items():
            for h, v in num.items():
                key = "rel_err_fd_h%g" % h
                res[name][key] = float(np.linalg.norm(g[params] - v) / np.linalg.norm(v))
        res["fd_check"] = {"params": params, "of": int(grads["exact"].size),
                           "seconds": time.perf_counter() - t0,
                           "threads": host_cpu()["threads"],
                           "note": "relative 2-norm error of the GPU gradient against FP64 "
                                   "central differences of the oracle's loss; h = 1e-10 is the "
                                   "converged derivative, h = 1e-4 is where the pathwise "
                                   "loss's kinks (samples crossing bilinear cells) dominate"}
    return res


# -------------------------------------------------------- single workloads

def bench_effects(args, eng):
    """Each built-in tail effect at its defaults on a 3840x2160 picture
    (density effects on a random f32 accumulator, colour effects on random
    RGBA8), HIP events around `steps` launches after `warmup`."""
    import torch
    import pyfae
    from pyfae import _abi
    H, W = 2160, 3840
    g = torch.Generator(device="cuda").manual_seed(1)
    src = torch.rand((H, W, 4), device="cuda", generator=g)
    dst = torch.empty_like(src)
    s8 = (torch.rand((H, W, 4), device="cuda", generator=g) * 255).to(torch.uint8)
    d8 = torch.empty_like(s8)
    ms = {}
    for name, (_, chain, _) in _abi.EFFECTS.items():
        fx = pyfae.effect(name)
        if chain == "density":
            run = lambda: eng.density_effect(fx, src, dst, density_mean=1.0)  # noqa: E731
        else:
            run = lambda: eng.color_effect(fx, s8, d8)  # noqa: E731
        for _ in range(max(args.warmup, 1)):
            run()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(args.steps):
            run()
        b.record()
        torch.cuda.synchronize()
        ms[name] = round(a.elapsed_time(b) / args.steps, 4)
    total = sum(ms.values())
    return {
        "metric": "tail effects ms at 3840x2160 (sum over the 18 built-ins)", "value": total,
        "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": total, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "effects: each built-in density and colour effect at its "
                               "defaults, 3840x2160"},
        "effect_ms": ms}


def bench_gallery(args, eng):
    r = gallery_run(eng, args.gallery_size, args.steps)
    return {"metric": "batched flames/s (gallery of %d)" % r["flames"],
            "value": r["batched_flames_per_s"], "unit": "flames/s", "n_gpus": 1,
            "steps": args.steps, "warmup": 1, "ms_per_step": r["batched_ms"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic", "config": {"workload": r["workload"]}, "gallery": r}


def bench_fit(args, eng):
    from pyfae import configs
    if args.workload == "fit":
        r = fit_extra(eng, args.steps, args.warmup, 1, not args.no_fd)
        value = r["exact"]["step_ms"]
        workload = r["workload"]
    else:
        items = configs.fit_batch()
        t, _, _ = fit_step_timing(eng, items, args.steps, args.warmup)
        r = {"fit_batch": t}
        value = t["step_ms"]
        workload = ("fit_batch: 64 independent 3-xform fits at 128x128, 2^16 walkers x 64 "
                    "iterations each, exact adjoint in windows of K=8, f64 walk, one launch")
    return {"metric": "adjoint fit step ms (%s)" % args.workload, "value": value, "unit": "ms",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": value,
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic", "config": {"workload": workload}, "fit": r}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="single4k",
                    choices=["single4k", "gallery", "plain", "fit", "fit_batch", "stress",
                             "effects"],
                    help="single4k (the headline, with every other workload's figures), "
                         "gallery, plain, fit, fit_batch (64 fits), stress (12288^2 f32, 2^34 "
                         "samples, LDS bins vs global atomics), effects (every tail effect at "
                         "3840x2160)")
    ap.add_argument("--no-fd", action="store_true", help="skip the finite-difference check")
    ap.add_argument("--no-extras", action="store_true",
                    help="single4k: the headline figures only")
    ap.add_argument("--gallery-size", type=int, default=512)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--splat-mode", type=int, default=2)
    ap.add_argument("--walkers-log2", type=int, default=24)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--variant", default="", choices=["", "nofinal", "linear"],
                    help="single4k instruction accounting (profiles), not a benchmark: drop the "
                         "final transform (nofinal), or also make every variation linear (linear)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="copy 1 GiB device-to-device first (a known-size kernel for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported by this build "
              "(the walkers would shard over GPUs by walker range, with one histogram "
              "reduce per batch)",
              file=sys.stderr)
        sys.exit(2)

    # torch first: its HIP runtime is the one libffm.so then binds to
    import torch
    import pyfae
    from pyfae import configs

    torch.cuda.init()
    eng = pyfae.Engine(0)
    if args.workload == "gallery":
        out = bench_gallery(args, eng)
    elif args.workload == "effects":
        out = bench_effects(args, eng)
    elif args.workload in ("fit", "fit_batch"):
        out = bench_fit(args, eng)
    elif args.workload == "stress":
        fd, pal, job = configs.stress(walkers=1 << 26)
        job.splat_mode = 4
        out, rgba = bench_flame(args, eng, fd, pal, job,
                                "stress: 12 xforms + final over 13 variations (incl. julian, "
                                "curl, waves), 12288x12288 float4 histogram, 2^26 walkers x 256 "
                                "iterations (2^34 samples), LDS tile bins",
                                "chaos-game samples/s at 12288^2 (stress)")
        out["splat_compare"] = splat_compare(eng, job, rgba, 3)
    elif args.workload == "plain":
        fd, pal, job = configs.plain()
        job.splat_mode = args.splat_mode
        out, _ = bench_flame(args, eng, fd, pal, job, "plain: 3 linear xforms, 512x512, 2^16 "
                             "walkers x 1000 iterations, log tonemap",
                             "chaos-game samples/s (plain 512x512)")
        if not args.no_cpu:
            out["cpu_baseline"] = cpu_baseline(fd, pal, job)
    else:
        fd, pal, job = configs.single4k(walkers=1 << args.walkers_log2,
                                        splat_mode=args.splat_mode)
        metric = "chaos-game samples/s at 4K"
        if args.variant:
            fd.n_finals = 0
            if args.variant == "linear":
                for k in range(fd.n_vars):
                    fd.var_id[k] = pyfae.VARIATIONS["linear"]
            metric += " (instruction-accounting variant %s, not the workload)" % args.variant
            args.no_extras = True
        progress(f"single4k: {args.warmup} warmup + {args.steps} timed renders")
        out, rgba = bench_flame(args, eng, fd, pal, job,
                                "single4k: 6 xforms + final xform, 3 of 10 variations each, "
                                "3840x2160, 2^%d walkers x 256 iterations, density estimation "
                                "9->0 curve 0.4, log tonemap, boxed down 2x to 1920x1080 "
                                "(firefly clamp)" % args.walkers_log2, metric, supersample=True)
        progress(f"single4k: {out['value']:.4g} samples/s")
        if not args.no_extras:
            progress("splat compare: global u32 atomics at 4K")
            out["splat_compare"] = splat_compare(eng, job, rgba, 0)
            del rgba
            progress("plain")
            out["plain"] = plain_extra(eng)
            progress("gallery (first call autotunes)")
            out["gallery"] = gallery_run(eng, 512, 2)
            progress("fit (exact and truncated adjoint; FD check on the host)")
            out["fit"] = fit_extra(eng, do_fd=not args.no_fd)
        if not args.no_cpu:
            progress("cpu baseline")
            out["cpu_baseline"] = cpu_baseline(fd, pal, job)
    eng.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
