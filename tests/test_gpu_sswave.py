"""The s-step launch in wave-strip form (glx_set_sstep_form(1, band_rows); no
reference counterpart, the operator is D^dag D of operators.cpp:444-453):
every wave marches a strip of 56 columns and band_rows rows on its own, with
the x-neighbours from the next and previous lanes (DPP) and every stage's rows
in registers.  The algorithm is the s-step CG of the default form; only the
FP32 rounding of the hops and the order of the FP64 sums differ.  Checks per
case, as for the default form:
  * success, the FP64 true relative residual <= tol recomputed with the
    oracle's operator (and equal to the one the solver reports);
  * the solution within ENVELOPE times the error of the oracle's FP64 CG
    solve at the same tolerance, both against its solve at 1e-13;
  * the iteration count within max(6, 3 %) of the default form's on the same
    inputs (both are the same CG in exact arithmetic);
  * strips that end inside a wave (lx not a multiple of 56), bands that end
    early (ly not a multiple of band_rows), lattices of 8 rows whose halo rows
    wrap onto the band itself, FP16 links, max_iter stops on either iteration
    of a launch, frequent reliable updates, an initial guess, b = 0, graph
    replay, and the 1024 x 1024 workload against its oracle solve.
"""
import numpy as np
import pytest

from helpers import ENVELOPE, cg_envelope, eo_envelope, lattice_inputs, relerr

pytestmark = pytest.mark.gpu


def _solve(glx, lat, mass, b, x0=None, form=1, band_rows=0, **kw):
    glx.set_sstep_form(form, band_rows)
    try:
        db = glx.DeviceArray(b)
        dx = glx.DeviceArray(np.zeros_like(b) if x0 is None else x0)
        info = glx.dev_cg_mixed(lat, "staggered_normal", mass, dx, db, **kw)
        return dx.download(), info
    finally:
        glx.set_sstep_form()


def _true_relres(orc, links, b, x, L, Ly, mass):
    r = b - orc.apply("staggered_normal", x, links, mass, L, Ly)
    return float(np.linalg.norm(r) / np.linalg.norm(b))


def _check(orc, links, b, x, info, L, Ly, mass, tol=1e-10):
    rr = _true_relres(orc, links, b, x, L, Ly, mass)
    assert info.success and rr <= tol * 1.001, rr
    assert abs(np.sqrt(info.resSq) / np.linalg.norm(b) - rr) <= 1e-3 * rr
    return rr


@pytest.mark.parametrize("link_prec", [1, 2])
@pytest.mark.parametrize("L,Ly,mass,beta,band_rows", [
    (128, 128, 0.05, 4.0, 0),     # three strips, the last of 16 columns
    (256, 64, 0.1, 6.25, 22),     # bands of 22, 22, 20 rows
    (128, 72, 0.05, 4.0, 10),     # ragged: 72 = 7 x 10 + 2
    (384, 40, 0.05, 4.0, 16),     # seven strips (the last 48 columns), bands 16, 16, 8
])
def test_wave_form_matches_oracle(glx, orc, L, Ly, mass, beta, band_rows, link_prec):
    links, bc, _ = lattice_inputs(orc, L, Ly, beta=beta, seed=171 + L + Ly)
    lat = glx.Lattice(links, L, Ly)
    x, info = _solve(glx, lat, mass, bc, band_rows=band_rows, max_iter=20000, eps=1e-10, delta=0.1,
                     link_prec=link_prec)
    rr = _check(orc, links, bc, x, info, L, Ly, mass)
    xs, err_o, xo, io = cg_envelope(orc, "staggered_normal", bc, links, mass, L, Ly)
    assert relerr(x, xs) <= ENVELOPE * err_o, (relerr(x, xs), err_o)
    x0, i0 = _solve(glx, lat, mass, bc, form=0, max_iter=20000, eps=1e-10, delta=0.1, link_prec=link_prec)
    print(f"{L}x{Ly} bands of {band_rows or 'auto'} rows, links {link_prec}: {info.iter} iterations "
          f"(default form {i0.iter}, oracle FP64 {io.iter}), true relres {rr:.3e}, |x - x_o|/|x_o| "
          f"{relerr(x, xo):.2e}; |x - x*|/|x*| {relerr(x, xs):.2e} (FP64 CG at the same tolerance {err_o:.2e})")
    assert i0.success and abs(info.iter - i0.iter) <= max(6, 0.03 * i0.iter), (info.iter, i0.iter)
    assert relerr(x0, xs) <= ENVELOPE * err_o, (relerr(x0, xs), err_o)


def test_wave_form_rows_of_eight(glx, orc):
    """An 8-row lattice: one band whose four halo rows above and below wrap
    onto its own rows."""
    L, Ly, mass = 256, 8, 0.1
    links, bc, _ = lattice_inputs(orc, L, Ly, beta=6.25, seed=12)
    lat = glx.Lattice(links, L, Ly)
    x, info = _solve(glx, lat, mass, bc, max_iter=20000, eps=1e-10)
    _check(orc, links, bc, x, info, L, Ly, mass)


@pytest.mark.parametrize("max_iter", [9, 10, 1])
def test_wave_form_max_iter(glx, orc, max_iter):
    """An unconverged solve stops after exactly max_iter iterations with the
    true residual of the returned x, which agrees with the oracle's after the
    same number of CG iterations."""
    L = 128
    links, bc, _ = lattice_inputs(orc, L, L, beta=4.0, seed=5)
    lat = glx.Lattice(links, L, L)
    x, info = _solve(glx, lat, 0.02, bc, max_iter=max_iter, eps=1e-10)
    assert not info.success and info.iter == max_iter, info.iter
    r = bc - orc.apply("staggered_normal", x, links, 0.02, L, L)
    assert abs(np.sqrt(info.resSq) - np.linalg.norm(r)) <= 1e-6 * np.linalg.norm(r)
    xo, _ = orc.solve("cg", "staggered_normal", bc, None, links=links, mass=0.02, lx=L, ly=L, max_iter=max_iter,
                      eps=1e-10)
    ro = np.linalg.norm(bc - orc.apply("staggered_normal", xo, links, 0.02, L, L))
    assert abs(np.linalg.norm(r) - ro) <= 1e-4 * ro, (np.linalg.norm(r), ro)


def test_wave_form_reliable_updates_and_guess(glx, orc):
    """delta = 0.5 (a reliable update and restart every few iterations), an
    initial guess, and b = 0."""
    L = 128
    links, bc, _ = lattice_inputs(orc, L, L, beta=4.0, seed=8)
    lat = glx.Lattice(links, L, L)
    x, info = _solve(glx, lat, 0.05, bc, max_iter=20000, eps=1e-10, delta=0.5)
    _, i01 = _solve(glx, lat, 0.05, bc, max_iter=20000, eps=1e-10, delta=0.1)
    assert info.reliable_updates > 2 * i01.reliable_updates
    _check(orc, links, bc, x, info, L, L, 0.05)
    xg, _ = orc.solve("cg", "staggered_normal", bc, None, links=links, mass=0.05, lx=L, ly=L, max_iter=20000,
                      eps=1e-4)
    x2, info2 = _solve(glx, lat, 0.05, bc, x0=xg, max_iter=20000, eps=1e-10)
    assert info2.iter < i01.iter
    _check(orc, links, bc, x2, info2, L, L, 0.05)
    x3, info3 = _solve(glx, lat, 0.05, np.zeros_like(bc), max_iter=100, eps=1e-10)
    assert info3.success and np.all(x3 == 0) and info3.resSq == 0.0


def test_wave_form_graph_replay(glx, orc):
    L, Ly = 256, 96
    links, bc, _ = lattice_inputs(orc, L, Ly, beta=6.25, seed=13)
    lat = glx.Lattice(links, L, Ly)
    glx.set_tuning(group=6, use_graph=True)
    try:
        x, info = _solve(glx, lat, 0.1, bc, band_rows=16, max_iter=20000, eps=1e-10)
    finally:
        glx.set_tuning()
    _check(orc, links, bc, x, info, L, Ly, 0.1)


def test_wave_form_workload_1024(glx, orc):
    """The 1024 x 1024 workload (m = 0.02, theta ~ N(0, 0.5^2)): the true
    residual meets 1e-10 and the solution is within 5e-11 of the oracle's
    FP64 CG solution (the default form's bound)."""
    L, mass = 1024, 0.02
    links, bc, _ = lattice_inputs(orc, L, L, beta=4.0, seed=3003)
    lat = glx.Lattice(links, L, L)
    x, info = _solve(glx, lat, mass, bc, max_iter=100000, eps=1e-10, delta=0.1)
    rr = _check(orc, links, bc, x, info, L, L, mass)
    xo, io = orc.solve("cg", "staggered_normal", bc, None, links=links, mass=mass, lx=L, ly=L, max_iter=100000,
                       eps=1e-10)
    err = relerr(x, xo)
    print(f"1024^2 wave form: {info.iter} iterations (oracle FP64 {io.iter}), {info.reliable_updates} reliable "
          f"updates, true relres {rr:.3e}, |x - x_o|/|x_o| = {err:.2e}")
    assert err <= 5e-11
    assert info.iter <= 1.2 * io.iter


def test_traffic_only_timing(glx, orc):
    """The wave-strip launch's loads and stores alone take less time than
    the launch itself, for both link precisions; lattices the s-step form
    does not take are refused."""
    L = 512
    links, _, _ = lattice_inputs(orc, L, L, beta=4.0, seed=2)
    lat = glx.Lattice(links, L, L, f32=True, f16=True)
    glx.set_sstep_form(1, 0)
    try:
        for lp in (glx.LINKS_F32, glx.LINKS_F16):
            t = glx.time_sscg_traffic(lat, lp, reps=20)
            ms, _ = glx.time_cg_kernel(lat, 0.05, prec=2 + lp, reps=20)
            assert 0 < t < ms, (lp, t, ms)
    finally:
        glx.set_sstep_form()
    small = glx.Lattice(links[:2 * 100 * 512], 100, 512)
    with pytest.raises(glx.GlxError):
        glx.time_sscg_traffic(small, glx.LINKS_F32, reps=1)


def test_wave_form_kernel_timing(glx, orc):
    """The timing entry runs the wave-strip launch when that form is set."""
    L = 256
    links, _, _ = lattice_inputs(orc, L, L, beta=4.0, seed=1)
    lat = glx.Lattice(links, L, L)
    glx.set_sstep_form(1, 0)
    try:
        ms, by = glx.time_cg_kernel(lat, 0.05, prec=3, reps=20)
    finally:
        glx.set_sstep_form()
    assert ms > 0 and by == 112
    assert glx.get_sstep_form() == -1


@pytest.mark.parametrize("lx,ly,link", [(256, 64, "mixed"), (128, 8, "mixed"), (384, 40, "mixed_f16"),
                                        (512, 72, "mixed"), (640, 44, "mixed")])
def test_wave_form_even_odd(glx, orc, lx, ly, link):
    """The even/odd solve of D x = b whose mixed CG runs the wave-strip
    launch on the compact Schur operator: the full residual bound of the
    pipeline, the solution within ENVELOPE of the FP64 even/odd CG's error,
    and the iteration count within max(6, 3 %) of the block form's."""
    m = 0.05
    links, bc, _ = lattice_inputs(orc, lx, ly, beta=6.25, seed=43 + lx + ly)
    lat = glx.Lattice(links, lx, ly)

    def solve(form):
        glx.set_sstep_form(form, 0)
        try:
            dx = glx.DeviceArray(np.zeros_like(bc))
            info, r2 = glx.dev_solve_staggered_eo(lat, m, dx, glx.DeviceArray(bc), link, max_iter=5000, eps=1e-10)
            return dx.download(), info
        finally:
            glx.set_sstep_form()

    x, info = solve(1)
    assert info.success
    r = bc - orc.apply("staggered", x, links, m, lx, ly)
    assert np.linalg.norm(r) <= 1e-10 * (m + 2.0) / m * np.linalg.norm(bc) * 1.01
    xs, err_o = eo_envelope(orc, bc, links, m, lx, ly)
    print(f"{lx}x{ly} {link} wave form: {info.iter} iterations, |x - x*|/|x*| {relerr(x, xs):.2e}, FP64 even/odd "
          f"CG at the same tolerance {err_o:.2e}")
    assert relerr(x, xs) <= ENVELOPE * err_o, (relerr(x, xs), err_o)
    _, i0 = solve(0)
    assert abs(info.iter - i0.iter) <= max(6, 0.03 * i0.iter), (info.iter, i0.iter)


def test_wave_form_even_odd_workload_1024(glx, orc):
    """The 1024 x 1024 even/odd solve (m = 0.02, theta ~ N(0, 0.5^2)) in the
    wave-strip form (its default at this size) against the oracle's FP64
    even/odd pipeline."""
    L, m = 1024, 0.02
    links, bc, _ = lattice_inputs(orc, L, L, beta=4.0, seed=3003)
    lat = glx.Lattice(links, L, L)
    assert glx.get_sstep_form_for(L, L) == 1
    dx = glx.DeviceArray(np.zeros_like(bc))
    info, r2 = glx.dev_solve_staggered_eo(lat, m, dx, glx.DeviceArray(bc), "mixed", max_iter=100000, eps=1e-10)
    x = dx.download()
    assert info.success
    r = bc - orc.apply("staggered", x, links, m, L, L)
    assert np.linalg.norm(r) <= 1e-10 * (m + 2.0) / m * np.linalg.norm(bc) * 1.01
    xs, err_o = eo_envelope(orc, bc, links, m, L, L)
    print(f"1024^2 even/odd wave form: {info.iter} iterations, |x - x*|/|x*| {relerr(x, xs):.2e} (FP64 even/odd CG "
          f"{err_o:.2e})")
    assert relerr(x, xs) <= ENVELOPE * err_o, (relerr(x, xs), err_o)
