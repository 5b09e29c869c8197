"""Mixed-precision CG in s-step form: two CG iterations per launch (no
reference counterpart; the operator is D^dag D of operators.cpp:444-453).

The mixed solver takes this form on D^dag D (glx_set_sstep(2, rows), the
default) when lx % 128 == 0, ly % 4 == 0 and ly >= 8.  Each launch
builds the basis p, Ap, A^2 p, r, Ar in FP32 with its Gram matrix, and the two
CG iterations run on coordinates in FP64; reliable updates recompute the FP64
true residual as in the one-iteration form.  Checks per case:
  * success, and the FP64 true relative residual <= tol recomputed with the
    oracle's operator (and equal to the one the solver reports);
  * the solution within ENVELOPE times the error of the oracle's FP64 CG
    solve at the same tolerance, both against its solve at 1e-13;
  * the iteration count close to the one-iteration mixed solver's on the same
    inputs (both are CG in exact arithmetic) and to the oracle's;
  * max_iter stops at odd and even counts, frequent reliable updates
    (delta = 0.5), a nonzero initial guess, b = 0, graph replay and other
    rows per block.
"""
import numpy as np
import pytest

from helpers import ENVELOPE, cg_envelope, eo_envelope, lattice_inputs, relerr

pytestmark = pytest.mark.gpu


def _solve(glx, lat, mass, b, x0=None, sstep=2, rows=0, **kw):
    glx.set_sstep(sstep, rows)
    try:
        db = glx.DeviceArray(b)
        dx = glx.DeviceArray(np.zeros_like(b) if x0 is None else x0)
        info = glx.dev_cg_mixed(lat, "staggered_normal", mass, dx, db, **kw)
        return dx.download(), info
    finally:
        glx.set_sstep()


def _check_solution(orc, links, b, x, info, L, Ly, mass, tol):
    r = b - orc.apply("staggered_normal", x, links, mass, L, Ly)
    relres = np.linalg.norm(r) / np.linalg.norm(b)
    assert relres <= tol * 1.001, relres
    assert abs(np.sqrt(info.resSq) / np.linalg.norm(b) - relres) <= 1e-3 * relres
    return relres


@pytest.mark.parametrize("link_prec", [1, 2])
@pytest.mark.parametrize("L,Ly,mass,beta", [(128, 128, 0.05, 4.0), (256, 64, 0.1, 6.25), (128, 72, 0.05, 4.0)])
def test_sstep_matches_oracle(glx, orc, L, Ly, mass, beta, link_prec):
    links, bc, _ = lattice_inputs(orc, L, Ly, beta=beta, seed=71 + L + Ly)
    lat = glx.Lattice(links, L, Ly)
    x, info = _solve(glx, lat, mass, bc, max_iter=20000, eps=1e-10, delta=0.1, link_prec=link_prec)
    assert info.success and info.reliable_updates >= 1
    _check_solution(orc, links, bc, x, info, L, Ly, mass, 1e-10)
    xs, err_o, xo, io = cg_envelope(orc, "staggered_normal", bc, links, mass, L, Ly)
    print(f"{L}x{Ly} links {link_prec}: |x - x*|/|x*| {relerr(x, xs):.2e}, FP64 CG at the same tolerance "
          f"{err_o:.2e}")
    assert relerr(x, xs) <= ENVELOPE * err_o, (relerr(x, xs), err_o)
    assert info.iter <= 1.5 * io.iter + 20, (info.iter, io.iter)
    # the one-iteration form on the same inputs: the same CG in exact arithmetic
    x1, i1 = _solve(glx, lat, mass, bc, sstep=1, max_iter=20000, eps=1e-10, delta=0.1, link_prec=link_prec)
    assert i1.success
    assert abs(info.iter - i1.iter) <= max(6, 0.03 * i1.iter), (info.iter, i1.iter)
    assert relerr(x1, xs) <= ENVELOPE * err_o, (relerr(x1, xs), err_o)


@pytest.mark.parametrize("max_iter", [9, 10, 1])
def test_sstep_max_iter(glx, orc, max_iter):
    """An unconverged solve stops after exactly max_iter iterations (odd and
    even: the stop falls on either iteration of a launch) with success = false
    and the true residual of the returned x."""
    L = 128
    links, bc, _ = lattice_inputs(orc, L, L, beta=4.0, seed=5)
    lat = glx.Lattice(links, L, L)
    x, info = _solve(glx, lat, 0.02, bc, max_iter=max_iter, eps=1e-10)
    assert not info.success
    assert info.iter == max_iter, info.iter
    r = bc - orc.apply("staggered_normal", x, links, 0.02, L, L)
    assert abs(np.sqrt(info.resSq) - np.linalg.norm(r)) <= 1e-6 * np.linalg.norm(r)
    # the same number of CG iterations as the oracle's: the residual agrees
    xo, io = orc.solve("cg", "staggered_normal", bc, None, links=links, mass=0.02, lx=L, ly=L, max_iter=max_iter,
                       eps=1e-10)
    ro = np.linalg.norm(bc - orc.apply("staggered_normal", xo, links, 0.02, L, L))
    assert abs(np.linalg.norm(r) - ro) <= 1e-4 * ro, (np.linalg.norm(r), ro)


def test_sstep_frequent_reliable_updates(glx, orc):
    """delta = 0.5: a reliable update every few iterations, each followed by a
    restart launch from the true residual; the solve still converges."""
    L = 128
    links, bc, _ = lattice_inputs(orc, L, L, beta=4.0, seed=8)
    lat = glx.Lattice(links, L, L)
    x, info = _solve(glx, lat, 0.05, bc, max_iter=20000, eps=1e-10, delta=0.5)
    _, i01 = _solve(glx, lat, 0.05, bc, max_iter=20000, eps=1e-10, delta=0.1)
    assert info.success and info.reliable_updates > 2 * i01.reliable_updates, (info.reliable_updates,
                                                                              i01.reliable_updates)
    _check_solution(orc, links, bc, x, info, L, L, 0.05, 1e-10)


def test_sstep_initial_guess(glx, orc):
    """The solve starts from r = b - A x0."""
    L = 128
    links, bc, _ = lattice_inputs(orc, L, L, beta=4.0, seed=9)
    lat = glx.Lattice(links, L, L)
    xo, _ = orc.solve("cg", "staggered_normal", bc, None, links=links, mass=0.05, lx=L, ly=L, max_iter=20000,
                      eps=1e-4)
    x, info = _solve(glx, lat, 0.05, bc, x0=xo, max_iter=20000, eps=1e-10)
    _, i0 = _solve(glx, lat, 0.05, bc, max_iter=20000, eps=1e-10)
    assert info.success and info.iter < i0.iter, (info.iter, i0.iter)
    _check_solution(orc, links, bc, x, info, L, L, 0.05, 1e-10)
    # an exact guess: the residual is zero, nothing changes
    x2, info2 = _solve(glx, lat, 0.05, np.zeros_like(bc), max_iter=100, eps=1e-10)
    assert info2.success and np.all(x2 == 0) and info2.resSq == 0.0


@pytest.mark.parametrize("L,Ly,rows,graph", [(256, 96, 8, False), (256, 96, 32, False), (256, 96, 16, True),
                                             (128, 8, 16, False), (384, 12, 8, False)])
def test_sstep_launch_shapes(glx, orc, L, Ly, rows, graph):
    """Rows per block 8 / 16 / 32, graph replay, and the smallest lattices the
    kernel takes (one strip row shorter than its rows per block; a strip of
    8 rows whose 4-row halos wrap onto the other strip)."""
    links, bc, _ = lattice_inputs(orc, L, Ly, beta=6.25, seed=12)
    lat = glx.Lattice(links, L, Ly)
    glx.set_tuning(group=6, use_graph=graph)
    try:
        x, info = _solve(glx, lat, 0.1, bc, rows=rows, max_iter=20000, eps=1e-10)
    finally:
        glx.set_tuning()
    assert info.success
    _check_solution(orc, links, bc, x, info, L, Ly, 0.1, 1e-10)


def test_sstep_kernel_timing(glx, orc):
    """The timing entry of the s-step kernel runs and reports its bytes."""
    L = 256
    links, _, _ = lattice_inputs(orc, L, L, beta=4.0, seed=1)
    lat = glx.Lattice(links, L, L)
    ms, by = glx.time_cg_kernel(lat, 0.05, prec=3, reps=20)
    assert ms > 0 and by == 112
    ms, by = glx.time_cg_kernel(lat, 0.05, prec=4, reps=20)
    assert ms > 0 and by == 108


def test_sstep_autotune(glx, orc):
    """The autotuner installs one of its candidates for the lattice size
    (0 = one-iteration form) and solves of that size use it."""
    L = 256
    links, bc, _ = lattice_inputs(orc, L, L, beta=4.0, seed=21)
    lat = glx.Lattice(links, L, L)
    try:
        res = glx.autotune_sstep(lat, reps=10)
        assert res["rows"] in (0, 8, 16, 32, 64, 128) and res["ms_per_iteration"] > 0 and res["form"] in (0, 1)
        assert glx.get_sstep_rows(L, L) == res["rows"] and glx.get_sstep_form_for(L, L) == res["form"]
        x, info = _solve(glx, lat, 0.05, bc, max_iter=20000, eps=1e-10)
        assert info.success
        _check_solution(orc, links, bc, x, info, L, L, 0.05, 1e-10)
    finally:
        glx.set_sstep_rows(L, L, -1)
        glx.set_sstep_form_entry(L, L, -1)
    assert glx.get_sstep_rows(L, L) == 8 and glx.get_sstep_form_for(L, L) == 0


@pytest.mark.parametrize("lx,ly,link", [(256, 64, "mixed"), (128, 8, "mixed"), (384, 40, "mixed_f16"), (512, 72, "mixed")])
def test_sstep_eo_schur(glx, orc, lx, ly, link):
    """The even/odd solve of D x = b (prepare, CG on m^2 - D_eo D_oe over the
    compact even sites, reconstruct) with its mixed CG in s-step form (64
    compact columns per strip): the full residual bound of the pipeline, the
    solution of the reference-order FP64 pipeline, and an iteration count close
    to the one-iteration form's (where that form takes the lattice: its FP32
    kernel needs lx / 2 a multiple of 128)."""
    m = 0.05
    links, bc, _ = lattice_inputs(orc, lx, ly, beta=6.25, seed=41 + lx + ly)
    lat = glx.Lattice(links, lx, ly)

    def solve(method, sstep):
        glx.set_sstep(sstep)
        try:
            dx = glx.DeviceArray(np.zeros_like(bc))
            info, r2 = glx.dev_solve_staggered_eo(lat, m, dx, glx.DeviceArray(bc), method, max_iter=5000, eps=1e-10)
            return dx.download(), info
        finally:
            glx.set_sstep()

    x, info = solve(link, 2)
    assert info.success
    r = bc - orc.apply("staggered", x, links, m, lx, ly)
    assert np.linalg.norm(r) <= 1e-10 * (m + 2.0) / m * np.linalg.norm(bc) * 1.01
    xs, err_o = eo_envelope(orc, bc, links, m, lx, ly)
    print(f"{lx}x{ly} {link}: |x - x*|/|x*| {relerr(x, xs):.2e}, FP64 even/odd CG at the same tolerance "
          f"{err_o:.2e}")
    assert relerr(x, xs) <= ENVELOPE * err_o, (relerr(x, xs), err_o)
    if (lx // 2) % 128 == 0:
        _, i1 = solve(link, 1)
        assert abs(info.iter - i1.iter) <= max(6, 0.03 * i1.iter), (info.iter, i1.iter)
