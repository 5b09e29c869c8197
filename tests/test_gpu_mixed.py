"""Mixed-precision CG with reliable updates (no reference counterpart).

FP32 Krylov vectors with FP32 or FP16 (+ per-site scale) links; the FP64
true residual r = b - A x is recomputed at every reliable update and the
solve stops on it.  Checks: the returned FP64 true relative residual is
<= tol; the solution is within ENVELOPE times the error of the oracle's FP64
CG solve at the same tolerance, both measured against the oracle's solve at
1e-13 (helpers.cg_envelope); the device's reported residual equals the
oracle's recomputation.
"""
import numpy as np
import pytest

from helpers import ENVELOPE, cg_envelope, lattice_inputs, relerr

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("link_prec", [1, 2])
@pytest.mark.parametrize("L,mass,beta", [(64, 0.05, 4.0), (48, 0.1, 6.25), (130, 0.05, 4.0)])
def test_mixed_cg_staggered(glx, orc, L, mass, beta, link_prec):
    links, bc, _ = lattice_inputs(orc, L, L, beta=beta, seed=31 + L)
    lat = glx.Lattice(links, L, L)
    db = glx.DeviceArray(bc)
    dx = glx.DeviceArray(np.zeros_like(bc))
    info = glx.dev_cg_mixed(lat, "staggered_normal", mass, dx, db, max_iter=20000, eps=1e-10, delta=0.1,
                            link_prec=link_prec)
    x = dx.download()
    assert info.success
    assert info.reliable_updates >= 1
    # independent check of the true residual with the oracle's FP64 operator
    r = bc - orc.apply("staggered_normal", x, links, mass, L, L)
    relres = np.linalg.norm(r) / np.linalg.norm(bc)
    assert relres <= 1e-10 * 1.001, relres
    assert abs(np.sqrt(info.resSq) / np.linalg.norm(bc) - relres) <= 1e-3 * relres
    xs, err_o, xo, io = cg_envelope(orc, "staggered_normal", bc, links, mass, L, L)
    err = relerr(x, xs)
    print(f"{L}^2 m={mass} links {link_prec}: |x - x*|/|x*| {err:.2e}, FP64 CG at the same tolerance {err_o:.2e}, "
          f"|x - x_o|/|x_o| {relerr(x, xo):.2e}")
    assert err <= ENVELOPE * err_o, (err, err_o)
    # iteration inflation of the FP32 Krylov space stays modest
    assert info.iter <= 1.5 * io.iter + 20, (info.iter, io.iter)


def test_mixed_cg_real_laplace(glx, orc):
    L = 64
    g = orc.Rng(4)
    b = g.gaussian_real(L * L)
    dx = glx.DeviceArray(np.zeros_like(b))
    db = glx.DeviceArray(b)
    lat = glx.Lattice(np.ones(2 * L * L, np.complex128), L, L)
    info = glx.dev_cg_mixed(lat, "laplace_real", 0.01, dx, db, max_iter=20000, eps=1e-10)
    x = dx.download()
    r = b - orc.apply("laplace_real", x, None, 0.01, L, L)
    assert info.success and np.linalg.norm(r) / np.linalg.norm(b) <= 1e-10 * 1.001


def test_mixed_cg_max_iter(glx, orc):
    """An unconverged mixed solve reports success = false and its true residual."""
    L = 32
    links, bc, _ = lattice_inputs(orc, L, L, beta=4.0, seed=2)
    lat = glx.Lattice(links, L, L)
    db = glx.DeviceArray(bc)
    dx = glx.DeviceArray(np.zeros_like(bc))
    info = glx.dev_cg_mixed(lat, "staggered_normal", 0.02, dx, db, max_iter=10, eps=1e-10)
    x = dx.download()
    assert not info.success
    r = bc - orc.apply("staggered_normal", x, links, 0.02, L, L)
    assert abs(np.sqrt(info.resSq) - np.linalg.norm(r)) <= 1e-6 * np.linalg.norm(r)


@pytest.mark.parametrize("L,mass,op", [(64, 0.05, "staggered_normal"), (130, 0.1, "staggered_normal"),
                                        (64, 0.01, "laplace_real")])
@pytest.mark.parametrize("graph", [True, False])
def test_fused_fp64_cg(glx, orc, L, mass, op, graph):
    """FP64 single-pass CG: same algorithm as the reference CG with the
    residual norm from the recurrence; iteration count within 2 of the
    oracle's, true residual below tolerance, solution within the accuracy the
    tolerance allows."""
    links, bc, br = lattice_inputs(orc, L, L, beta=4.0, seed=51 + L)
    b = br if op == "laplace_real" else bc
    lat = glx.Lattice(links, L, L)
    glx.set_tuning(group=8, use_graph=graph)
    db = glx.DeviceArray(b)
    dx = glx.DeviceArray(np.zeros_like(b))
    info = glx.dev_cg_fused(lat, op, mass, dx, db, max_iter=20000, eps=1e-10)
    glx.set_tuning()
    x = dx.download()
    xs, err_o, xo, io = cg_envelope(orc, op, b, links, mass, L, L)
    assert info.success and abs(info.iter - io.iter) <= 2, (info.iter, io.iter)
    r = b - orc.apply(op, x, links, mass, L, L)
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1.05e-10
    assert relerr(x, xs) <= ENVELOPE * err_o, (relerr(x, xs), err_o)
