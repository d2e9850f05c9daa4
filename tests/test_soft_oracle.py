"""CPU checks of the FP64 soft-splat oracle (oracle/ffm_soft.c) and of the
host-side soft-splat helpers. No GPU.

The oracle defines the differentiable mode (the reference has none), so it is
pinned by what the mathematics fixes:
  * each variation's Jacobian equals central differences of its value;
  * the adjoint equals central finite differences of the forward's loss;
  * the truncated adjoint equals the exact one when its window spans the walk;
  * the forward's mass equals density_scale x the plotted-sample count of the
    chaos-game oracle (same RNG stream, same transform picks).
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle
from pyfae import _abi, configs, flame_desc, soft_job, soft_theta, xform

FIT_IDS = ["linear", "sinusoidal", "spherical", "swirl", "horseshoe", "polar", "disc",
           "hyperbolic", "julia", "julian", "bubble", "curl"]


def loss_grad(rho, tau):
    d = np.log1p(rho) - np.log1p(tau)
    return float((d * d).sum()), 2.0 * d / (1.0 + rho)


@pytest.mark.parametrize("name", FIT_IDS)
def test_variation_jacobian_matches_central_differences(name):
    vid = _abi.VARIATIONS[name]
    params = (3.0, 1.2, 1.2 / 3.0 / 2.0, 0.0) if name == "julian" else (0.4, -0.2, 0.0, 0.0)
    rng = np.random.default_rng(vid)
    h = 1e-6
    for _ in range(40):
        x, y = rng.uniform(-1.5, 1.5, 2)
        if x * x + y * y < 0.05 or (name in ("julia", "julian") and y * y < 1e-3 and x < 0):
            continue  # away from the singular origin and the atan2 branch cut
        V, J = oracle.soft_variation(vid, x, y, params)
        Vxp, _ = oracle.soft_variation(vid, x + h, y, params)
        Vxm, _ = oracle.soft_variation(vid, x - h, y, params)
        Vyp, _ = oracle.soft_variation(vid, x, y + h, params)
        Vym, _ = oracle.soft_variation(vid, x, y - h, params)
        fd = np.stack([(Vxp - Vxm) / (2 * h), (Vyp - Vym) / (2 * h)], axis=1)
        scale = max(1.0, np.abs(fd).max())
        assert np.abs(J - fd).max() <= 1e-6 * scale, (name, x, y, J, fd)


def small_fit(walkers=512, ipt=24, side=40, burn_in=4):
    fd, tfd, _, _ = configs.fit()
    job = soft_job(side, side, walkers=walkers, ipt=ipt, burn_in=burn_in, seed=1, window=8,
                   zoom=4.0 / 3.0)
    tjob = soft_job(side, side, walkers=walkers, ipt=ipt, burn_in=burn_in, seed=2, window=8,
                    zoom=4.0 / 3.0)
    return fd, tfd, job, tjob


def test_adjoint_matches_finite_differences():
    fd, tfd, job, tjob = small_fit()
    theta = soft_theta(fd)
    tau = oracle.soft_forward(tfd, soft_theta(tfd), tjob)
    rho = oracle.soft_forward(fd, theta, job)
    _, g = loss_grad(rho, tau)
    adj = oracle.soft_adjoint(fd, theta, job, g)
    h = 1e-7
    num = np.zeros_like(theta)
    for q in range(theta.size):
        tp, tm = theta.copy(), theta.copy()
        tp[q] += h
        tm[q] -= h
        num[q] = (loss_grad(oracle.soft_forward(fd, tp, job), tau)[0] -
                  loss_grad(oracle.soft_forward(fd, tm, job), tau)[0]) / (2 * h)
    err = np.linalg.norm(adj - num) / np.linalg.norm(num)
    assert err < 1e-4, (err, adj, num)
    # every parameter is reached (affines of all 4 transforms, all weights)
    assert np.all(np.abs(num) > 0)


def test_truncated_adjoint_limits():
    fd, tfd, job, tjob = small_fit(walkers=256)
    theta = soft_theta(fd)
    tau = oracle.soft_forward(tfd, soft_theta(tfd), tjob)
    _, g = loss_grad(oracle.soft_forward(fd, theta, job), tau)
    exact = oracle.soft_adjoint(fd, theta, job, g)
    job.truncate = 1
    job.window = job.ipt
    assert np.array_equal(oracle.soft_adjoint(fd, theta, job, g), exact)
    job.window = 1
    one = oracle.soft_adjoint(fd, theta, job, g)
    job.window = 8
    eight = oracle.soft_adjoint(fd, theta, job, g)
    # longer windows approach the exact gradient
    assert np.linalg.norm(eight - exact) < np.linalg.norm(one - exact)


def test_forward_mass_equals_plotted_samples():
    fd, _, job, _ = small_fit(walkers=1024, ipt=40, side=64)
    job.view.zoom = 0.25  # the whole attractor well inside the frame
    rho = oracle.soft_forward(fd, soft_theta(fd), job)
    stops = configs.random_palette(configs.SplitMix64(5))
    _, plotted, inframe = oracle.iterate(fd, stops, job.view, job.walkers, job.ipt, job.burn_in,
                                         job.seed)
    assert inframe == plotted == job.walkers * (job.ipt - job.burn_in)
    assert rho.sum() == pytest.approx(job.density_scale * plotted, rel=1e-9)


def test_forward_is_thread_count_independent():
    fd, _, job, _ = small_fit(walkers=300)
    th = soft_theta(fd)
    a = oracle.soft_forward(fd, th, job, threads=1)
    b = oracle.soft_forward(fd, th, job, threads=7)
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-15)


def test_theta_packing_matches_library():
    """pyfae.soft_theta restates ffm_softsplat_pack (host-only call)."""
    from pyfae import library
    fd, _, _, _ = configs.fit()
    P = library().ffm_softsplat_param_count(C.byref(fd))
    assert P == 4 * (6 + 3)
    out = (C.c_float * P)()
    assert library().ffm_softsplat_pack(C.byref(fd), out) == 0
    np.testing.assert_array_equal(np.array(out[:], np.float32), soft_theta(fd).astype(np.float32))


def test_unfittable_flames_are_rejected():
    from pyfae import library
    fd = flame_desc([xform(variations={"waves": 1.0})])
    assert library().ffm_softsplat_param_count(C.byref(fd)) < 0
    fd = flame_desc([xform()], finals=[xform()])
    assert library().ffm_softsplat_param_count(C.byref(fd)) < 0
    assert oracle.lib().orc_soft_param_count(C.byref(fd)) < 0
