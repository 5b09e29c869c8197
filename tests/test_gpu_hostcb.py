"""Two routes of the drop-in boundary against the oracle:

  * an operator the library does not know (a caller's own host function,
    here numpy code following operators.cpp:184-239, :372-453 and
    tests/multishift/multishift.cpp:634-672 term by term): the solver keeps
    its vectors and reductions on the GPU and calls the function on host
    copies for each application (generic_cg.cpp:278-382 and the other
    solvers take their operator as such a callback);
  * the dispatcher minv_unpreconditioned (generic_inverter.cpp:18-188) with
    CG, GCR and BiCGStab, plain and restarted, on registered device
    operators and on the host callback.

CG and GCR: the oracle's iteration count, ops_count, success and name, x
within 1e-11 (1e-13 for the real Laplacian).  BiCGStab is not stable under
rounding (generic_bicgstab.cpp:205-341): it must lie inside the oracle's own
spread under 1-ulp perturbations of b.
"""
import ctypes as C

import numpy as np
import pytest

from helpers import lattice_inputs, relerr
from test_gpu_configs import bicgstab_envelope, check_in_envelope

pytestmark = pytest.mark.gpu

MATVEC = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.c_void_p)


def np_staggered(psi, links, m, lx, ly, dagger=False):
    """D (or D^dag) with the reference's terms in its order: -U psi(+x),
    +U*(x-1) psi(-x), -eta U_y psi(+y), +eta U_y*(y-1) psi(-y), x 1/2, + m psi."""
    p = psi.reshape(ly, lx)
    u = links.reshape(ly, lx, 2)
    ux, uy = u[..., 0], u[..., 1]
    eta = (1.0 - 2.0 * (np.arange(lx) % 2))[None, :]
    h = np.zeros_like(p)
    h = h - ux * np.roll(p, -1, axis=1)
    h = h + np.conj(np.roll(ux, 1, axis=1)) * np.roll(p, 1, axis=1)
    h = h - (eta * uy) * np.roll(p, -1, axis=0)
    h = h + (eta * np.conj(np.roll(uy, 1, axis=0))) * np.roll(p, 1, axis=0)
    if dagger:
        h = -h
    return (0.5 * h + m * p).reshape(-1)


def np_laplace_real(v, m2, lx, ly):
    p = v.reshape(ly, lx)
    h = np.zeros_like(p)
    h = h - np.roll(p, -1, axis=1)
    h = h - np.roll(p, 1, axis=1)
    h = h - np.roll(p, -1, axis=0)
    h = h - np.roll(p, 1, axis=0)
    return (h + (4.0 + m2) * p).reshape(-1)


class HostOp:
    """A host callback computing `op` in numpy; counts its calls."""

    def __init__(self, op, links, m, lx, ly):
        self.n = lx * ly
        self.calls = 0
        real = op == "laplace_real"

        def fn(lhs, rhs, extra):
            self.calls += 1
            if real:
                x = np.ctypeslib.as_array((C.c_double * self.n).from_address(rhs))
                out = np.ctypeslib.as_array((C.c_double * self.n).from_address(lhs))
                out[:] = np_laplace_real(x.copy(), m, lx, ly)
                return
            x = np.ctypeslib.as_array((C.c_double * (2 * self.n)).from_address(rhs)).view(np.complex128)
            out = np.ctypeslib.as_array((C.c_double * (2 * self.n)).from_address(lhs)).view(np.complex128)
            if op == "staggered":
                out[:] = np_staggered(x.copy(), links, m, lx, ly)
            else:
                out[:] = np_staggered(np_staggered(x.copy(), links, m, lx, ly), links, m, lx, ly, dagger=True)

        self.fn = MATVEC(fn)
        self.addr = C.cast(self.fn, C.c_void_p).value


def test_numpy_operators_are_the_oracles(orc):
    """The callbacks compute the reference's operators (to rounding)."""
    lx, ly = 16, 12
    links, bc, br = lattice_inputs(orc, lx, ly, beta=3.0, seed=41)
    assert relerr(np_staggered(bc, links, 0.1, lx, ly), orc.apply("staggered", bc, links, 0.1, lx, ly)) < 1e-15
    want = orc.apply("staggered_normal", bc, links, 0.1, lx, ly)
    got = np_staggered(np_staggered(bc, links, 0.1, lx, ly), links, 0.1, lx, ly, dagger=True)
    assert relerr(got, want) < 1e-15
    assert relerr(np_laplace_real(br, 0.01, lx, ly), orc.apply("laplace_real", br, None, 0.01, lx, ly)) < 1e-15


CASES = [
    # solver, op, lx, ly, mass, restart_freq
    ("cg", "staggered_normal", 16, 16, 0.05, 0),
    ("cg", "staggered_normal", 24, 10, 0.1, 9),
    ("cg", "laplace_real", 32, 32, 0.01, 0),
    ("gcr", "staggered", 16, 16, 0.1, 0),
    ("gcr", "staggered", 16, 16, 0.1, 12),
]


def _oracle_kw(op, links, mass, lx, ly, max_iter=2000):
    return dict(links=None if op == "laplace_real" else links, mass=mass, lx=lx, ly=ly, max_iter=max_iter, eps=1e-10)


@pytest.mark.parametrize("solver,op,lx,ly,mass,rf", CASES)
def test_host_callback_matches_oracle(glx, orc, solver, op, lx, ly, mass, rf):
    links, bc, br = lattice_inputs(orc, lx, ly, beta=6.25, seed=300 + lx + rf)
    b = br if op == "laplace_real" else bc
    kw = _oracle_kw(op, links, mass, lx, ly)
    xo, io = orc.solve(solver, op, b, None, restart_freq=rf, verb_level=3, prefix="[H]: ", **kw)
    h = HostOp(op, links, mass, lx, ly)
    xg, ig = glx.solve(solver, op, b, None, restart_freq=rf, verb_level=3, prefix="[H]: ", callback=h.addr,
                       extra=None, max_iter=kw["max_iter"], eps=kw["eps"])
    assert (ig.iter, ig.ops_count, ig.success, ig.name) == (io.iter, io.ops_count, io.success, io.name)
    assert relerr(xg, xo) < (1e-13 if op == "laplace_real" else 1e-11), relerr(xg, xo)
    assert h.calls >= ig.iter  # every application went through the caller's function
    gl, ol = ig.text.splitlines(), io.text.splitlines()
    assert len(gl) == len(ol) and all(a.split("RelRes")[0] == o.split("RelRes")[0] for a, o in zip(gl, ol))


def test_host_callback_bicgstab_inside_envelope(glx, orc):
    lx = ly = 16
    links, bc, _ = lattice_inputs(orc, lx, ly, beta=6.25, seed=77)
    kw = _oracle_kw("staggered", links, 0.1, lx, ly)
    x0, i0, its, dx = bicgstab_envelope(orc, "staggered", bc, kw)
    h = HostOp("staggered", links, 0.1, lx, ly)
    xg, ig = glx.solve("bicgstab", "staggered", bc, None, callback=h.addr, extra=None, max_iter=2000, eps=1e-10)
    assert ig.success and ig.ops_count == 2 * ig.iter + 2
    check_in_envelope(xg, ig, x0, its, np.maximum(dx, 1e-14), "host-callback BiCGStab 16^2")


DISPATCH = [
    # solver, op, restart
    ("cg", "staggered_normal", False),
    ("cg", "staggered_normal", True),
    ("cg", "laplace_real", False),
    ("gcr", "staggered", False),
    ("gcr", "staggered", True),
]


@pytest.mark.parametrize("host", [False, True])
@pytest.mark.parametrize("solver,op,restart", DISPATCH)
def test_dispatcher_matches_oracle(glx, orc, solver, op, restart, host):
    lx, ly, mass, rf = 16, 16, 0.08, 11
    links, bc, br = lattice_inputs(orc, lx, ly, beta=6.25, seed=500 + int(restart))
    b = br if op == "laplace_real" else bc
    kw = _oracle_kw(op, links, mass, lx, ly)
    xo, io = orc.solve(solver, op, b, None, restart_freq=rf if restart else 0, **kw)
    extra = {}
    if host:
        h = HostOp(op, links, mass, lx, ly)
        extra = dict(callback=h.addr, extra=None)
    xg, ig = glx.minv_unpreconditioned(solver, op, b, None, links=links, mass=mass, lx=lx, ly=ly, tol=1e-10,
                                       max_iters=2000, restart=restart, restart_freq=rf, **extra)
    assert (ig.iter, ig.ops_count, ig.success, ig.name) == (io.iter, io.ops_count, io.success, io.name)
    assert relerr(xg, xo) < (1e-13 if op == "laplace_real" else 1e-11)


@pytest.mark.parametrize("restart", [False, True])
def test_dispatcher_bicgstab_inside_envelope(glx, orc, restart):
    lx, ly, mass, rf = 16, 16, 0.2, 25
    links, bc, _ = lattice_inputs(orc, lx, ly, beta=6.25, seed=5)
    kw = _oracle_kw("staggered", links, mass, lx, ly)
    x0, i0, its, dx = bicgstab_envelope(orc, "staggered", bc, kw, restart_freq=rf if restart else 0)
    xg, ig = glx.minv_unpreconditioned("bicgstab", "staggered", bc, None, links=links, mass=mass, lx=lx, ly=ly,
                                       tol=1e-10, max_iters=2000, restart=restart, restart_freq=rf)
    assert ig.success and i0.success and ig.name == i0.name
    check_in_envelope(xg, ig, x0, its, np.maximum(dx, 1e-14), f"dispatcher BiCGStab restart={restart}")


def test_dispatcher_invalid_and_refused(glx, orc):
    links, bc, _ = lattice_inputs(orc, 16, 16, seed=6)
    x, info = glx.minv_unpreconditioned("invalid", "staggered", bc, None, links=links, mass=0.1, lx=16, ly=16)
    assert (info.iter, info.success, info.ops_count) == (0, False, 0) and not np.any(x)
    for solver in ("cr", "bicgstab_l", "gmres", "sor", "minres"):
        with pytest.raises(glx.GlxError):
            glx.minv_unpreconditioned(solver, "staggered", bc, None, links=links, mass=0.1, lx=16, ly=16)
