"""Chaos-game throughput on one MI355X: the 4K single-flame workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one whole render of the single4k workload through libffm.so:
6 transforms + a final (3 variations each), 3840x2160, 2^24 walkers x 256
iterations (one dispatch), accumulate, density estimation (radius 9 -> 0,
curve 0.4) and the log tonemap, pixels left in HBM. `value` is chaos-game
samples (walker iterations) per second over the timed steps.

Prints ONE JSON line. The CPU baseline is the oracle's chaos game on this
host's cores over a bounded sample of the same flame and frame.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fflame-rust_amd"))


def cpu_baseline(fd, pal, job, seconds=10.0, single_seconds=4.0):
    """The oracle (a scalar C restatement of the reference's step) timed on
    this host's cores: calibrate on a small sample, then run about
    `seconds` of work on all cores and `single_seconds` on one. Beside it,
    the reference's Criterion microbenchmarks (benches/flame_bench.rs)
    restated in C, with the per-op means the reference publishes."""
    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    threads = max(1, min(os.cpu_count() or 1, 16))
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
                      f"{v.width}x{v.height} (iterate + splat), {dt:.1f} s",
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="single4k",
                    choices=["single4k", "gallery", "plain", "fit", "fit_batch", "stress",
                             "effects"],
                    help="single4k (the headline), gallery (batched vs one by one), plain, "
                         "fit (differentiable soft-splat fit step), fit_batch (64 fits), "
                         "stress (12288^2 f32, 2^34 samples, LDS bins vs global atomics), "
                         "effects (every tail effect at 3840x2160, defaults)")
    ap.add_argument("--no-fd", action="store_true", help="fit: skip the finite-difference check")
    ap.add_argument("--gallery-size", type=int, default=512)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--splat-mode", type=int, default=2)
    ap.add_argument("--walkers-log2", type=int, default=24)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="copy 1 GiB device-to-device first (a known-size kernel for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported by this build "
              "(the walkers would shard over GPUs by walker range, with one histogram "
              "reduce per batch)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import pyfae
    from pyfae import configs, roofline

    if args.workload == "gallery":
        return bench_gallery(args)
    if args.workload == "effects":
        return bench_effects(args)
    if args.workload in ("fit", "fit_batch"):
        return bench_fit(args)
    if args.workload == "stress":
        fd, pal, job = configs.stress(walkers=1 << 26)
        job.splat_mode = 4
        args.splat_mode = 4
        args.walkers_log2 = 26
        return bench_flame(args, fd, pal, job,
                           "stress: 12 xforms + final over 13 variations (incl. julian, curl, "
                           "waves), 12288x12288 float4 histogram, 2^26 walkers x 256 iterations "
                           "(2^34 samples), LDS tile bins", "chaos-game samples/s at 12288^2 (stress)",
                           compare_mode=3)
    if args.workload == "plain":
        fd, pal, job = configs.plain()
        job.splat_mode = args.splat_mode
        args.walkers_log2 = 16
        return bench_flame(args, fd, pal, job, "plain: 3 linear xforms, 512x512, 2^16 walkers x "
                           "1000 iterations, log tonemap", "chaos-game samples/s (plain 512x512)")
    fd, pal, job = configs.single4k(walkers=1 << args.walkers_log2, splat_mode=args.splat_mode)
    return bench_flame(args, fd, pal, job,
                       "single4k: 6 xforms + final xform, 3 of 10 variations each, 3840x2160, "
                       "2^%d walkers x 256 iterations, density estimation 9->0 curve 0.4, log "
                       "tonemap" % args.walkers_log2, "chaos-game samples/s at 4K")


def bench_effects(args):
    """Each built-in tail effect at its defaults on a 3840x2160 picture
    (density effects on a random f32 accumulator, colour effects on random
    RGBA8), HIP events around `steps` launches after `warmup`."""
    import torch
    import pyfae
    from pyfae import _abi
    eng = pyfae.Engine(0)
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
    print(json.dumps({
        "metric": "tail effects ms at 3840x2160 (sum over the 18 built-ins)", "value": total,
        "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": total, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "effects: each built-in density and colour effect at its "
                               "defaults, 3840x2160"},
        "effect_ms": ms}))


def bench_gallery(args):
    """Batched gallery: flames/s of one batched render (persistent launches in a
    hipGraph) against the same flames rendered one by one."""
    import pyfae
    from pyfae import configs
    items = configs.gallery(n=args.gallery_size)
    eng = pyfae.Engine(0)
    for _ in range(max(args.warmup, 1)):  # the first call autotunes and builds the graph
        eng.batch_render(items)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _, st = eng.batch_render(items)
    dt = (time.perf_counter() - t0) / args.steps
    t1 = time.perf_counter()
    _, st_s = eng.batch_render(items, serial=True)
    dt_s = time.perf_counter() - t1
    n = len(items)
    samples = st.total_samples
    px = sum(j.view.width * j.view.height for _, _, j in items)
    eng.close()
    out = {
        "metric": "batched flames/s (gallery of %d)" % n,
        "value": n / dt, "unit": "flames/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "gallery: %d flames, 2..12 xforms, sides log-uniform 128..1024, "
                               "200 samples/pixel, ipt 256" % n,
                   "pixels": px, "samples_per_step": samples},
        "serial_flames_per_s": n / dt_s, "serial_ms": dt_s * 1e3,
        "speedup_vs_serial": dt_s / dt, "samples_per_s": samples / dt,
    }
    print(json.dumps(out), flush=True)


def bench_fit(args):
    """Differentiable fit: one step = soft-splat forward, the log-density L2
    loss in torch, and the adjoint, through pyfae.diff's autograd Function.
    Reports the step time, the forward and adjoint kernel times and, for the
    single fit, the gradient against FP64 central finite differences of the
    oracle's loss (outside the timed region)."""
    import numpy as np
    import torch
    import pyfae
    from pyfae import configs
    from pyfae.diff import SoftSplat, log_density_l2
    eng = pyfae.Engine(0)
    if args.workload == "fit":
        items = [configs.fit()]
        workload = ("fit: 4 xforms {linear, spherical, sinusoidal}, 256x256 soft splat, 2^20 "
                    "walkers x 64 iterations, exact adjoint in windows of K=8, f64 walk")
    else:
        items = configs.fit_batch()
        workload = ("fit_batch: 64 independent 3-xform fits at 128x128, 2^16 walkers x 64 "
                    "iterations each, exact adjoint in windows of K=8, f64 walk, one launch")
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
        for _ in range(max(args.warmup, 1)):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            grad = step()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
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
    out = {
        "metric": "adjoint fit step ms (%s)" % args.workload,
        "value": dt * 1e3, "unit": "ms", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt * 1e3, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": workload, "fits": len(items), "params": int(theta0.numel()),
                   "samples_per_step": samples},
        "forward_ms": fwd_ms, "adjoint_ms": adj_ms,
        "step_over_forward": (fwd_ms + adj_ms) / fwd_ms,
        "walk_samples_per_s": {"forward": samples / (fwd_ms / 1e3),
                               "adjoint": samples / (adj_ms / 1e3)},
    }
    if args.workload == "fit" and not args.no_fd:
        out["fd_check"] = fd_check(items[0], tau.cpu().numpy().astype(np.float64),
                                   grad.cpu().numpy().astype(np.float64))
    eng.close()
    print(json.dumps(out), flush=True)


def fd_check(item, tau, grad):
    """Relative 2-norm error of the GPU gradient against FP64 central
    differences of the oracle's loss, at h = 1e-10 (converged) and at 1e-4
    (where the pathwise loss is rough at that scale)."""
    import numpy as np
    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    from pyfae import soft_theta
    fd, _, job, _ = item
    th = soft_theta(fd).astype(np.float32).astype(np.float64)
    tau = tau.reshape(job.view.height, job.view.width)
    threads = max(1, min(os.cpu_count() or 1, 16))

    def L(t):
        d = np.log1p(O.soft_forward(fd, t, job, threads=threads)) - np.log1p(tau)
        return float((d * d).sum())

    out = {}
    t0 = time.perf_counter()
    for name, h in (("h1e-10", 1e-10), ("h1e-4", 1e-4)):
        num = np.array([(L(th + h * e) - L(th - h * e)) / (2 * h) for e in np.eye(th.size)])
        out["rel_err_" + name] = float(np.linalg.norm(grad - num) / np.linalg.norm(num))
    out["seconds"] = time.perf_counter() - t0
    out["threads"] = threads
    return out


def bench_flame(args, fd, pal, job, workload, metric, compare_mode=None):
    import torch
    import pyfae
    from pyfae import roofline
    W, H = job.view.width, job.view.height
    eng = pyfae.Engine(0)
    eng.set_flame(fd)
    eng.set_palette(pal, 256)
    rgba = torch.empty((H, W, 4), dtype=torch.uint8, device="cuda:0")
    if args.copy_calibration:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    for _ in range(args.warmup):
        eng.render_to_device(job, rgba)
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        stats.append(eng.render_to_device(job, rgba))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
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
    kname = "bin_iterate_kernel" if args.splat_mode in (2, 4) else "iterate_kernel"
    traffic = recorded_traffic(kname, args.splat_mode, job.walkers)
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
        "config": {"workload": workload, "splat_mode": args.splat_mode,
                   "samples_per_step": samples},
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
    if compare_mode is not None:
        # the same render with another splat (outside the timed steps)
        mode = job.splat_mode
        job.splat_mode = compare_mode
        eng.render_to_device(job, rgba)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eng.render_to_device(job, rgba)
        torch.cuda.synchronize()
        cmp_ms = (time.perf_counter() - t0) * 1e3
        job.splat_mode = mode
        out["compare"] = {"splat_mode": compare_mode, "ms_per_step": cmp_ms,
                          "speedup": cmp_ms / ms_per_step}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(fd, pal, job)
    eng.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
