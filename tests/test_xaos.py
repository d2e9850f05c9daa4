"""Xaos in the oracle (CPU): select_transform_xaos
(shaders/core/utilities.wgsl:107-140), the extra start draw and the
stratified first pick of persistent orbits (shaders/core/main_template.wgsl:
63-71,125-181), and the orbit refresh rule (src/scene/xaos.rs:64-124;
src/renderer/compute_kernel.rs:46-60).

Pins: the all-ones matrix reduces to the pinned unconditioned selection;
zero entries are transitions that never happen; pick frequencies over a
fine grid of draws are the closed-form probabilities weight_j * xaos_ij /
row sum; the refresh rule on hand-solved reachability graphs."""
import ctypes as C

import numpy as np

import pyfae
from oracle import oracle as O


def lin_flame(n, xaos, weights=None):
    xs = []
    for i in range(n):
        ang = 2 * np.pi * i / n
        xs.append(pyfae.xform(0.5, 0.0, 0.0, 0.5, 0.6 * np.cos(ang), 0.6 * np.sin(ang),
                              weight=1.0 if weights is None else weights[i], color=i / max(n - 1, 1),
                              variations={"linear": 1.0}))
    return pyfae.flame_desc(xs, xaos=xaos)


def test_all_ones_matrix_is_the_plain_pick():
    w = [0.3, 1.7, 0.25, 2.0, 0.5]
    fd = lin_flame(5, np.ones((5, 5)), w)
    for u in np.linspace(0, 1, 4097, dtype=np.float32)[:-1]:
        want = O.lib().orc_select(C.byref(fd), float(u))
        for prev in range(5):
            assert O.select_xaos(fd, float(u), prev) == want


def test_pick_frequencies_are_the_transition_probabilities():
    w = np.array([0.3, 1.7, 0.25, 2.0], np.float32)
    m = np.array([[0, 1, 2, 1], [1, 0, 0, 3], [0.5, 0.5, 0.5, 0.5], [4, 0, 1, 0]], np.float32)
    fd = lin_flame(4, m, w)
    grid = (np.arange(1 << 14, dtype=np.float64) + 0.5) / (1 << 14)
    for prev in range(4):
        picks = np.array([O.select_xaos(fd, float(u), prev) for u in grid])
        p = w * m[prev]
        p = p / p.sum()
        freq = np.bincount(picks, minlength=4) / len(grid)
        np.testing.assert_allclose(freq, p, atol=2.0 / len(grid))
        # a zero entry is a transition that never happens
        assert all(m[prev, j] > 0 for j in set(picks.tolist()))
    # a row of zeros: total 0, threshold 0, the first transform
    z = lin_flame(3, [[0, 0, 0], [1, 1, 1], [1, 1, 1]])
    assert {O.select_xaos(z, float(u), 0) for u in (0.0, 0.3, 0.99)} == {0}


def test_start_draws_once_more_and_chain_respects_the_matrix():
    n = 4
    fd1 = lin_flame(n, np.ones((n, n)))
    fd0 = lin_flame(n, None)
    view = pyfae.View(64, 64, 1.0, 0.0, 0.0, 0.0)
    stops = np.array([[0, 1, 0, 0], [1, 0, 0, 1]], np.float32)
    pal = O.palette_linear(stops)
    L = O.lib()
    for walker in (0, 77, 4097):
        want = O.trace(fd1, stops, view, walker, 9, 20, 200)
        # the same walk without xaos, from a stream one draw further on
        r = O.Rng()
        L.orc_rng_init(C.byref(r), walker, 9)
        w = O.Walker()
        two, one = np.float32(2), np.float32(1)
        w.x = float(np.float32(L.orc_rng_nextf(C.byref(r))) * two - one)
        w.y = float(np.float32(L.orc_rng_nextf(C.byref(r))) * two - one)
        w.fuse = 20
        L.orc_rng_nextf(C.byref(r))  # the xaos start pick's draw
        rec = (C.c_uint32 * 8)()
        got = []
        f0 = O.prepared(fd0)
        for _ in range(200):
            L.orc_step_from(C.byref(f0), O.fp(pal), 256, C.byref(view), 20, C.byref(r),
                            C.byref(w), rec)
            got.append(list(rec))
        np.testing.assert_array_equal(np.array(got, np.uint32), want)
    # a chain 0 -> 1 -> 2 -> 0 only
    ring = np.zeros((3, 3))
    ring[0, 1] = ring[1, 2] = ring[2, 0] = 1.0
    tr = O.trace(lin_flame(3, ring), stops, view, 5, 1, 20, 300)
    xi = tr[:, 1].astype(int)
    assert ((xi[1:] - xi[:-1]) % 3 == 1).all()


def test_refresh_period_rule():
    one = np.ones((3, 3))
    assert O.xaos_refresh_period(lin_flame(3, None)) == 0
    assert O.xaos_refresh_period(lin_flame(3, one)) == 0  # one closed group
    # two isolated self-loops: each first pick stays in its own group
    assert O.xaos_refresh_period(lin_flame(2, [[1, 0], [0, 1]])) == 0
    # 0 can fall into {1} or {2}: the walk decides the group
    assert O.xaos_refresh_period(lin_flame(3, [[1, 1, 1], [0, 1, 0], [0, 0, 1]])) == 4
    # ... unless 0 is never a first pick and never reached (weight 0)
    assert O.xaos_refresh_period(lin_flame(3, [[1, 1, 1], [0, 1, 0], [0, 0, 1]],
                                           [0.0, 1.0, 1.0])) == 0
    # a row with no way out goes to transform 0 (the reference's fallback)
    assert O.xaos_refresh_period(lin_flame(3, [[0, 1, 1], [0, 0, 0], [0, 0, 1]])) == 0
