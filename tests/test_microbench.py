"""The restated Criterion microbenchmarks (benches/flame_bench.rs) that
bench.py times beside the GPU run: each runs and reports a finite time, and
computes what the Rust bench computes — checked against float32 numpy
restatements and closed forms, so a wrong operand or a dropped term shows."""
import math

import numpy as np

from oracle import oracle as O

f32 = np.float32


def test_every_bench_runs():
    names = [m[0] for m in O.MICROBENCHES]
    assert len(names) == len(set(names)) == 12
    for name, which, params, pub in O.MICROBENCHES:
        ns, _ = O.microbench(which, params, iters=20000)
        assert 0 < ns < 1e5, name
        assert pub > 0


def test_affine_and_helpers_compute_the_reference_values():
    aff = O.MICROBENCHES[0][2]
    _, st = O.microbench(0, aff, iters=10)
    a, b, c, d, e, f, x, y = (f32(v) for v in aff[:8])
    assert st[0] == a * x + b * y + e and st[1] == c * x + d * y + f
    assert abs(float(st[0]) - 0.55) < 1e-6 and abs(float(st[1]) - 0.2) < 1e-6
    pt = (0, 0, 0, 0, 0, 0, 0.7, 0.3)
    px, py = f32(0.7), f32(0.3)
    assert O.microbench(2, pt, iters=10)[1][0] == px * px + py * py
    assert O.microbench(3, pt, iters=10)[1][0] == np.sqrt(px * px + py * py)
    # theta = atan2(y, x), phi = atan2(x, y): operands not swapped
    th = float(O.microbench(4, pt, iters=10)[1][0])
    ph = float(O.microbench(5, pt, iters=10)[1][0])
    assert abs(th - math.atan2(0.3, 0.7)) < 1e-7 and abs(ph - math.atan2(0.7, 0.3)) < 1e-7
    for s in (0.0, 0.5, 1.0):
        got = O.microbench(6, (0,) * 6 + (0.3, 0.7, s), iters=10)[1][0]
        assert got == f32(0.3) * (f32(1) - f32(s)) + f32(0.7) * f32(s)


def walk(v, n):
    """numpy float32 restatement of single_transform_iter's loop body."""
    a, b, c, d, e, f, col, s = (f32(t) for t in v[:8])
    x, y, cp = f32(0.5), f32(0.5), f32(0.5)
    for _ in range(n):
        ax, ay = a * x + b * y + e, c * x + d * y + f
        x, y = (ax * f32(v[8]), ay * f32(v[8])) if v[8] is not None else (ax, ay)
        cp = cp * (f32(1) - s) + col * s
    return x, y, cp


def test_single_transform_iter_walks_the_preset_transform():
    with np.errstate(over="ignore"):
        for name, which, params, _ in O.MICROBENCHES[9:]:
            key = name.rsplit("/", 1)[1]
            for n in (1, 7, 3000):
                _, st = O.microbench(which, params, iters=n)
                want = walk(O.PRESET_FIRST_XFORMS[key], n)
                np.testing.assert_array_equal(st, np.array(want, np.float32), err_msg=f"{key} {n}")
    # Julian Disc's first transform contracts (|det| < 1): after 3000 steps the
    # point sits in the subnormal range, which is why that bench is slow
    _, st = O.microbench(7, O.MICROBENCHES[10][2], iters=3000)
    assert 0 < max(abs(st[0]), abs(st[1])) < np.finfo(np.float32).tiny
    # a linear weight scales the point
    v = (0.5, 0, 0, 0.5, 0.1, 0.2, 0.25, 0.5, 2.0)
    p = v[:6] + (0.5, 0.5, v[7], v[6], 2.0, 1.0)
    np.testing.assert_array_equal(O.microbench(7, p, iters=5)[1], np.array(walk(v, 5), np.float32))
