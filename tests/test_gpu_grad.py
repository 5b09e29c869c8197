"""Gradient of S(theta, m) = phi^dag (D^dag D)^-1 phi (no reference
counterpart; D is operators.cpp:184-239 with U = exp(i theta)).

The device computes chi = (D^dag D)^-1 phi, psi = D chi and the per-link
force in one pass (glx_dev_staggered_grad).  Checks, at the test size and at
the 512^2 workload (m = 0.05, theta ~ N(0, 0.4^2), phi complex Gaussian):
- against the FP64 CPU gradient of the oracle (its own CG solve, oracle
  force); both solves to 1e-12, so the bound is kappa * 1e-12 sized;
- against central finite differences (h = 1e-6, 32 seeded links and the
  mass) of the estimator 2 Re phi^dag x - x^dag D^dag D x, whose error is
  second order in the solve residual; expected agreement ~1e-5 relative
  (rounding of S / h);
- the FP32-bulk variant (mixed-precision solve to 1e-10) against the FP64
  gradient, with its error stated;
- the torch.autograd wrapper: its backward equals the device gradient and a
  random directional derivative matches finite differences.
"""
import numpy as np
import pytest

from helpers import relerr

pytestmark = pytest.mark.gpu


def inputs(L, seed=21):
    rng = np.random.default_rng(seed)
    theta = rng.normal(0.0, 0.4, 2 * L * L)
    phi = rng.normal(size=L * L) + 1j * rng.normal(size=L * L)
    return theta, np.exp(1j * theta), phi


@pytest.mark.parametrize("L", [64, 512])
def test_grad_matches_cpu_fp64(glx, orc, L):
    m = 0.05
    theta, links, phi = inputs(L)
    lat = glx.Lattice(links, L, L)
    g = glx.staggered_grad(lat, m, glx.DeviceArray(phi), precision=0, eps=1e-12)
    chi, info = orc.solve("cg", "staggered_normal", phi, None, links=links, mass=m, lx=L, ly=L, max_iter=100000,
                          eps=1e-12)
    psi = orc.apply("staggered", chi, links, m, L, L)
    gc, dmc = orc.force(chi, psi, links, m, L, L)
    assert g.info.success and info.success
    err = relerr(g.grad, gc)
    print(f"L={L}: |grad - grad_cpu| / |grad_cpu| = {err:.2e}, dS/dm {g.dS_dm:.10e} vs {dmc:.10e}")
    # measured 2e-14 (64^2) and 4e-13 (512^2)
    assert err < 1e-11
    assert abs(g.dS_dm - dmc) <= 1e-11 * abs(dmc)
    S = np.vdot(phi, chi).real
    assert abs(g.S - S) <= 1e-9 * abs(S)


@pytest.mark.parametrize("L", [64, 512])
def test_grad_matches_finite_differences(glx, L):
    m, h = 0.05, 1e-6
    theta, links, phi = inputs(L)
    dphi = glx.DeviceArray(phi)
    lat = glx.Lattice(links, L, L)
    g = glx.staggered_grad(lat, m, dphi, eps=1e-12)

    def S(th, mm=m):
        lat.update(np.exp(1j * th))
        return glx.staggered_grad(lat, mm, dphi, eps=1e-13).S_robust

    picks = np.random.default_rng(99).choice(2 * L * L, 32, replace=False)
    fd = np.empty(32)
    for n, j in enumerate(picks):
        tp = theta.copy()
        tp[j] += h
        tm = theta.copy()
        tm[j] -= h
        fd[n] = (S(tp) - S(tm)) / (2 * h)
    err = np.max(np.abs(fd - g.grad[picks])) / np.max(np.abs(g.grad[picks]))
    fdm = (S(theta, m + h) - S(theta, m - h)) / (2 * h)
    errm = abs(fdm - g.dS_dm) / abs(g.dS_dm)
    print(f"L={L}: FD max error / max |grad| over 32 links = {err:.2e}; dS/dm relative error {errm:.2e}")
    # measured 5e-8 (64^2) and 6e-6 (512^2); dS/dm 4e-10
    assert err < 1e-4
    assert errm < 1e-6


@pytest.mark.parametrize("prec", [1, 2])
def test_grad_fp32_bulk_variant(glx, prec):
    L, m = 512, 0.05
    theta, links, phi = inputs(L)
    lat = glx.Lattice(links, L, L)
    dphi = glx.DeviceArray(phi)
    g64 = glx.staggered_grad(lat, m, dphi, precision=0, eps=1e-12)
    gmx = glx.staggered_grad(lat, m, dphi, precision=prec, eps=1e-10)
    err = relerr(gmx.grad, g64.grad)
    errm = abs(gmx.dS_dm - g64.dS_dm) / abs(g64.dS_dm)
    print(f"FP32 bulk ({'FP32' if prec == 1 else 'FP16'} links), solve to 1e-10: gradient relative error {err:.2e}, "
          f"dS/dm {errm:.2e}, {gmx.info.iter} iterations, {gmx.info.reliable_updates} reliable updates")
    assert gmx.info.success
    # chi has relative error <= kappa * 1e-10 (kappa = (m^2 + 2) / m^2 ~ 800);
    # measured 2.9e-10 with FP32 and with FP16 links
    assert err < 1e-8
    assert errm < 1e-8


def test_torch_autograd_wrapper(glx):
    torch = pytest.importorskip("torch")
    from glx.torch_check import StaggeredAction

    L, m, h = 64, 0.05, 1e-6
    theta, links, phi = inputs(L, seed=22)
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    mt = torch.tensor(m, dtype=torch.float64, requires_grad=True)
    ph = torch.from_numpy(phi)
    S = StaggeredAction.apply(th, mt, ph, L, L)
    S.backward()
    g = glx.staggered_grad(glx.Lattice(links, L, L), m, glx.DeviceArray(phi), eps=1e-12)
    assert relerr(th.grad.numpy(), g.grad) < 1e-12
    assert abs(mt.grad.item() - g.dS_dm) <= 1e-12 * abs(g.dS_dm)
    v = np.random.default_rng(5).normal(size=theta.size)
    Sp = StaggeredAction.apply(torch.from_numpy(theta + h * v), mt.detach(), ph, L, L).item()
    Sm = StaggeredAction.apply(torch.from_numpy(theta - h * v), mt.detach(), ph, L, L).item()
    dd, want = (Sp - Sm) / (2 * h), float(g.grad @ v)
    print(f"directional derivative: FD {dd:.10e}, gradient {want:.10e}")
    assert abs(dd - want) <= 1e-5 * abs(want)
