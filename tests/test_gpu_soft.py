"""GPU checks of the differentiable soft splat (softsplat.hip) against the
FP64 oracle (oracle/ffm_soft.c), whose adjoint is itself pinned by finite
differences on the CPU.

Bars (same RNG stream and transform picks as the oracle):
  * f64 walk: density within 1e-6 relative L1 of the oracle's; gradients
    within 1e-5 relative (2-norm) of the oracle's adjoint, and within 1e-3 of
    FP64 central finite differences of the oracle's loss (h = 1e-10, where
    the differences have converged on the full-size fit);
  * f32 walk (the chaos game's arithmetic): density within 1e-4; truncated
    gradients within 1e-3; exact gradients within 5e-3, because the pathwise
    gradient amplifies f32 rounding of the positions (rounding only the
    oracle's positions to f32 moves its own exact gradient by 3e-3);
  * batched fits bit-identical to the same fits run alone; autograd through
    pyfae.diff equal to the direct adjoint call.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle  # noqa: E402
from pyfae import configs, soft_job, soft_theta  # noqa: E402

pytestmark = pytest.mark.gpu


def loss_grad(rho, tau):
    d = np.log1p(rho) - np.log1p(tau)
    return float((d * d).sum()), 2.0 * d / (1.0 + rho)


def gpu_forward(engine, fits, theta32):
    T, D = engine.softsplat_set(fits)
    th = torch.from_numpy(np.ascontiguousarray(theta32, np.float32)).cuda()
    dens = torch.empty(D, device="cuda", dtype=torch.float32)
    engine.softsplat_forward(th, dens)
    torch.cuda.synchronize()
    return th, dens


def gpu_adjoint(engine, th, g):
    gd = torch.from_numpy(np.ascontiguousarray(g, np.float32)).cuda()
    out = torch.empty_like(th)
    engine.softsplat_adjoint(th, gd, out)
    torch.cuda.synchronize()
    return out.cpu().numpy().astype(np.float64)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def small(walkers=1 << 14, ipt=64, side=128, window=8, truncate=False, f64=True):
    fd, tfd, _, _ = configs.fit()
    kw = dict(walkers=walkers, ipt=ipt, burn_in=20, window=window, truncate=truncate,
              walk_f64=f64, zoom=4.0 / 3.0)
    return fd, tfd, soft_job(side, side, seed=1, **kw), soft_job(side, side, seed=2, **kw)


@pytest.mark.parametrize("f64,tol", [(True, 1e-6), (False, 1e-4)])
def test_soft_forward_matches_oracle(engine, f64, tol):
    fd, _, job, _ = small(f64=f64)
    print("soft forward", f64, job.walkers, job.ipt, job.view.width, flush=True)
    th32 = soft_theta(fd).astype(np.float32)
    _, dens = gpu_forward(engine, [(fd, job)], th32)
    got = dens.cpu().numpy().astype(np.float64).reshape(job.view.height, job.view.width)
    want = oracle.soft_forward(fd, th32.astype(np.float64), job)
    assert np.abs(got - want).sum() <= tol * want.sum(), np.abs(got - want).sum() / want.sum()
    assert got.sum() == pytest.approx(want.sum(), rel=1e-5)


@pytest.mark.parametrize("f64,truncate,window,tol", [
    (True, False, 8, 1e-5), (True, True, 8, 1e-5), (True, False, 5, 1e-5), (True, True, 16, 1e-5),
    (False, True, 8, 1e-3), (False, False, 8, 5e-3)])
def test_soft_adjoint_matches_oracle(engine, f64, truncate, window, tol):
    fd, tfd, job, tjob = small(window=window, truncate=truncate, f64=f64)
    print("soft adjoint", f64, truncate, window, flush=True)
    th32 = soft_theta(fd).astype(np.float32)
    th, _ = gpu_forward(engine, [(fd, job)], th32)
    tau = oracle.soft_forward(tfd, soft_theta(tfd), tjob)
    rho = oracle.soft_forward(fd, th32.astype(np.float64), job)
    _, g = loss_grad(rho, tau)
    got = gpu_adjoint(engine, th, g)
    want = oracle.soft_adjoint(fd, th32.astype(np.float64), job, g)
    print("rel", rel(got, want), flush=True)
    assert rel(got, want) < tol, (rel(got, want), got, want)
    # deterministic: a second call gives the same bits
    assert np.array_equal(gpu_adjoint(engine, th, g), got)


@pytest.mark.timeout(600)
def test_fit_gradient_matches_fp64_finite_differences(engine):
    """The fit workload's full size (2^20 walkers x 64, 256^2, K = 8 exact, f64 walk).
    h = 1e-10: the pathwise loss is rough at coarser steps (kinks of the
    bilinear taps crossed by samples whose positions are very sensitive), and
    its central differences converge to the oracle's adjoint (to 5e-7) only
    from about h = 1e-9 down."""
    fd, tfd, job, tjob = configs.fit()
    print("fit FD check", job.walkers, job.ipt, flush=True)
    th32 = soft_theta(fd).astype(np.float32)
    th64 = th32.astype(np.float64)
    _, dens = gpu_forward(engine, [(tfd, tjob)], soft_theta(tfd).astype(np.float32))
    tau = dens.cpu().numpy().astype(np.float64).reshape(256, 256)
    th, dens = gpu_forward(engine, [(fd, job)], th32)
    rho = dens.cpu().numpy().astype(np.float64).reshape(256, 256)
    _, g = loss_grad(rho, tau)
    got = gpu_adjoint(engine, th, g)

    def L(t):
        return loss_grad(oracle.soft_forward(fd, t, job, threads=16), tau)[0]

    h = 1e-10
    num = np.zeros_like(th64)
    for q in range(th64.size):
        e = np.zeros_like(th64)
        e[q] = h
        num[q] = (L(th64 + e) - L(th64 - e)) / (2 * h)
    print("rel to FD", rel(got, num), flush=True)
    assert rel(got, num) < 1e-3, (rel(got, num), got, num)


def test_autograd_matches_direct_adjoint(engine):
    from pyfae.diff import SoftSplat, log_density_l2
    fd, tfd, job, tjob = small(walkers=1 << 13)
    target_splat = SoftSplat(engine, [(tfd, tjob)])
    with torch.no_grad():
        tau = target_splat(target_splat.theta0())
    splat = SoftSplat(engine, [(fd, job)])
    theta = splat.theta0().requires_grad_()
    loss = log_density_l2(splat(theta), tau)
    loss.backward()
    torch.cuda.synchronize()
    rho = splat(theta.detach())
    g = (2.0 * (torch.log1p(rho) - torch.log1p(tau)) / (1.0 + rho)).cpu().numpy()
    direct = gpu_adjoint(engine, theta.detach().contiguous(), g)
    np.testing.assert_allclose(theta.grad.cpu().numpy(), direct, rtol=1e-5, atol=1e-6 * np.abs(direct).max())


def test_batched_fits_equal_single_fits(engine):
    items = configs.fit_batch(n=6, side=64, walkers=1 << 12, ipt=32)
    fits = [(fd, job) for fd, _, job, _ in items]
    thetas = [soft_theta(fd).astype(np.float32) for fd, _ in fits]
    print("batched fits", len(fits), flush=True)
    th, dens = gpu_forward(engine, fits, np.concatenate(thetas))
    rng = np.random.default_rng(3)
    g = rng.standard_normal(dens.numel())
    got = gpu_adjoint(engine, th, g)
    dens = dens.cpu().numpy()
    off_p = off_d = 0
    for (fd, job), t in zip(fits, thetas):
        npx = job.view.width * job.view.height
        th1, d1 = gpu_forward(engine, [(fd, job)], t)
        np.testing.assert_allclose(dens[off_d:off_d + npx], d1.cpu().numpy(), rtol=1e-5,
                                   atol=1e-6 * float(d1.max()))
        g1 = gpu_adjoint(engine, th1, g[off_d:off_d + npx])
        assert np.array_equal(got[off_p:off_p + t.size], g1)
        off_p += t.size
        off_d += npx


def test_soft_errors_are_status_codes(engine):
    from pyfae import EngineError, flame_desc, xform
    fd = flame_desc([xform(variations={"waves": 1.0})])
    with pytest.raises(EngineError):
        engine.softsplat_set([(fd, soft_job(16, 16, walkers=64, ipt=8))])
    fd, _, job, _ = small(walkers=256)
    job.window = 0
    with pytest.raises(EngineError):
        engine.softsplat_set([(fd, job)])
    job.window = 8
    engine.softsplat_set([(fd, job)])
    with pytest.raises(EngineError):
        engine.softsplat_forward(None, None)
