"""Closed forms and invariants that pin the oracle where fixtures cannot (CPU).

Each check is chosen so that a plausible mistake fails it: a dropped hopping
term, a wrong sign or phase (eta), a transposed link index, a missing
conjugate, a wrong parity convention, an off-by-one in the lattice index.
"""
import numpy as np
import pytest

from helpers import lattice_inputs, relerr


# ------------------------------------------------------------ lattice index
@pytest.mark.parametrize("dims,nc", [((4, 3), 1), ((5, 7), 2), ((2, 2), 3), ((6, 1), 1)])
def test_lattice_index_bijection(orc, dims, nc):
    """lattice.h:52-139: i = (y*lx + x)*nc + c, parity of x + y, by brute force."""
    lx, ly = dims
    seen = set()
    for y in range(ly):
        for x in range(lx):
            for c in range(nc):
                i = orc.lattice_index(dims, nc, (x, y), c)
                assert i == (y * lx + x) * nc + c
                (cx, cy), cc = orc.lattice_coord(dims, nc, i)
                assert (cx, cy, cc) == (x, y, c)
                assert orc.lattice_is_even(dims, nc, i) == ((x + y) % 2 == 0)
                seen.add(i)
    assert seen == set(range(lx * ly * nc))


# -------------------------------------------------------------- mt19937 KAT
def test_mt19937_known_answer(orc):
    """The C++ standard ([rand.predef]) fixes the 10000th output of a
    default-constructed std::mt19937 (seed 5489) as 4123659995."""
    import ctypes as C
    lib = orc.orc_lib()
    g = orc.OrcMT()
    lib.orc_mt_seed(C.byref(g), 5489)
    for _ in range(9999):
        lib.orc_mt_next(C.byref(g))
    assert lib.orc_mt_next(C.byref(g)) == 4123659995


def test_normal_distribution_moments(orc):
    g = orc.Rng(3)
    v = g.gaussian_real(200000)
    assert abs(v.mean()) < 0.01 and abs(v.std() - 1) < 0.01
    u = g.gauss_gauge(200, 250, 6.25)
    th = np.angle(u)
    assert abs(th.std() - 0.4) < 0.005


# --------------------------------------------------------------- operators
def plane_wave(lx, ly, nx, ny):
    x = np.arange(lx)[None, :]
    y = np.arange(ly)[:, None]
    return np.exp(2j * np.pi * (nx * x / lx + ny * y / ly)).reshape(-1)


@pytest.mark.parametrize("nx,ny", [(0, 0), (1, 0), (2, 3), (5, 7)])
def test_free_laplace_spectrum(orc, nx, ny):
    lx, ly, m2 = 8, 12, 0.3
    v = plane_wave(lx, ly, nx, ny)
    lam = 4 + m2 - 2 * np.cos(2 * np.pi * nx / lx) - 2 * np.cos(2 * np.pi * ny / ly)
    bound = 1e-14 * (8 + m2) * np.linalg.norm(v)
    assert np.linalg.norm(orc.apply("laplace", v, None, m2, lx, ly) - lam * v) < bound
    assert np.linalg.norm(orc.apply("laplace_u1", v, orc.unit_gauge(lx, ly), m2, lx, ly) - lam * v) < bound
    assert np.linalg.norm(orc.apply("laplace_real", v.real.copy(), None, m2, lx, ly) - lam * v.real) < bound


@pytest.mark.parametrize("nx,ny", [(0, 0), (1, 2), (3, 5), (4, 6)])
def test_free_staggered_normal_spectrum(orc, nx, ny):
    """U = 1: D^dag D e^{ipx} = (m^2 + sin^2 p_x + sin^2 p_y) e^{ipx}."""
    lx, ly, m = 8, 12, 0.2
    v = plane_wave(lx, ly, nx, ny)
    px, py = 2 * np.pi * nx / lx, 2 * np.pi * ny / ly
    lam = m * m + np.sin(px) ** 2 + np.sin(py) ** 2
    # error bound relative to |A| |v| (|A| <= m^2 + 2): near p = pi the
    # eigenvalue is m^2 and forms by cancellation of O(1) terms
    r = orc.apply("staggered_normal", v, orc.unit_gauge(lx, ly), m, lx, ly) - lam * v
    assert np.linalg.norm(r) < 1e-14 * (m * m + 2) * np.linalg.norm(v)


def test_free_laplace_fft_solve(orc):
    """A CG solve of the free Laplacian equals the FFT solution."""
    lx, ly, m2 = 32, 16, 0.01
    b = orc.Rng(8).gaussian_real(lx * ly)
    x, info = orc.solve("cg", "laplace_real", b, None, lx=lx, ly=ly, mass=m2, eps=1e-12, max_iter=5000)
    kx = 2 * np.pi * np.fft.fftfreq(lx)[None, :]
    ky = 2 * np.pi * np.fft.fftfreq(ly)[:, None]
    lam = 4 + m2 - 2 * np.cos(kx) - 2 * np.cos(ky)
    xf = np.fft.ifft2(np.fft.fft2(b.reshape(ly, lx)) / lam).real.reshape(-1)
    assert info.success and relerr(x, xf) < 1e-9


def test_staggered_identities(orc):
    """staggered_pieces.cpp:263-323: D^dag = g5 D g5, D = D_eo + D_oe + m,
    D^dag D = m^2 - D_eo D_oe - D_oe D_eo; g5 D = gamma5 o D; and the
    even-only m^2 - D_eo D_oe is the even part of D^dag D on even input."""
    lx, ly, m = 12, 10, 0.3
    links, v, _ = lattice_inputs(orc, lx, ly, beta=2.0, seed=4)
    A = lambda op, w: orc.apply(op, w, links, m, lx, ly)
    assert relerr(A("staggered_dagger", v), A("gamma5", A("staggered", A("gamma5", v)))) < 1e-14
    assert relerr(A("staggered", v), A("deo", v) + A("doe", v) + m * v) < 1e-14
    lhs = A("staggered_normal", v)
    rhs = m * m * v - A("deo", A("doe", v)) - A("doe", A("deo", v))
    assert relerr(lhs, rhs) < 1e-14
    assert relerr(A("staggered_gamma5", v), A("gamma5", A("staggered", v))) < 1e-14
    x, y = np.meshgrid(np.arange(lx), np.arange(ly))
    even = ((x + y) % 2 == 0).reshape(-1)
    ve = np.where(even, v, 0)
    assert relerr(A("m2mdeodoe", ve), np.where(even, A("staggered_normal", ve), 0)) < 1e-14


def test_staggered_antihermitian_hopping(orc):
    """<u, D v> = <D^dag u, v>, and K = m - D is anti-Hermitian."""
    lx, ly, m = 10, 8, 0.25
    links, v, _ = lattice_inputs(orc, lx, ly, beta=1.5, seed=6)
    u = orc.Rng(60).gaussian_complex(lx * ly)
    Dv = orc.apply("staggered", v, links, m, lx, ly)
    Ddu = orc.apply("staggered_dagger", u, links, m, lx, ly)
    assert abs(np.vdot(u, Dv) - np.vdot(Ddu, v)) < 1e-12 * abs(np.vdot(u, Dv))
    Kv = m * v - Dv
    Ku = m * u - orc.apply("staggered", u, links, m, lx, ly)
    assert abs(np.vdot(u, Kv) + np.vdot(Ku, v)) < 1e-12 * np.linalg.norm(u) * np.linalg.norm(Kv)


def test_gauge_covariance(orc):
    """D[U^g] (g psi) = g (D[U] psi) with U_mu(x) -> g(x) U_mu(x) g(x+mu)^*;
    the plaquette is gauge invariant (u1_test.cpp:40-83)."""
    lx, ly, m = 8, 8, 0.1
    links, v, _ = lattice_inputs(orc, lx, ly, beta=3.0, seed=2)
    t = orc.Rng(12).rand_trans(lx, ly)
    lg = orc.apply_gauge_trans(links, t, lx, ly)
    for op in ("staggered", "staggered_dagger", "staggered_normal", "laplace_u1"):
        lhs = orc.apply(op, t * v, lg, m, lx, ly)
        rhs = t * orc.apply(op, v, links, m, lx, ly)
        assert relerr(lhs, rhs) < 1e-14, op
    assert abs(orc.plaquette(lg, lx, ly) - orc.plaquette(links, lx, ly)) < 1e-14


def test_stencil_form_equals_function_form(orc):
    """staggered_stencil.cpp:215-251 and staggered_pieces.cpp:361-508."""
    lx, ly, m = 12, 10, 0.3
    links, v, _ = lattice_inputs(orc, lx, ly, beta=2.0, seed=14)
    for f, s in (("staggered", "stencil"), ("staggered_dagger", "stencil_dagger"),
                 ("staggered_gamma5", "stencil_gamma5"), ("deo", "stencil_eo"), ("doe", "stencil_oe"),
                 ("m2mdeodoe", "stencil_m2mdeodoe")):
        assert relerr(orc.apply(s, v, links, m, lx, ly), orc.apply(f, v, links, m, lx, ly)) < 1e-14, s


# --------------------------------------------------------------------- BLAS
def test_blas_against_numpy(orc):
    g = orc.Rng(21)
    a, b = g.gaussian_complex(1000), g.gaussian_complex(1000)
    n = 1000
    bound = 2 * n * np.finfo(float).eps * np.sum(np.abs(a) * np.abs(b))
    assert abs(orc.blas("dot", a, b) - np.vdot(a, b)) < bound
    assert abs(orc.blas("norm2sq", a) - np.vdot(a, a).real) < bound
    assert abs(orc.blas("diffnorm2sq", a, b) - np.linalg.norm(a - b) ** 2) < 4 * bound
    ar, br = a.real.copy(), b.real.copy()
    assert abs(orc.blas("dot", ar, br) - ar @ br) < bound


# ------------------------------------------------------------------ solvers
def test_solvers_solve(orc):
    """Every solver's answer satisfies its own residual bound against an
    independent dense solve."""
    lx = ly = 8
    links, b, br = lattice_inputs(orc, lx, ly, beta=4.0, seed=30)
    m = 0.2
    n = lx * ly
    D = np.stack([orc.apply("staggered", e, links, m, lx, ly) for e in np.eye(n, dtype=complex)], axis=1)
    xs = np.linalg.solve(D, b)
    for solver in ("bicgstab", "gcr"):
        x, info = orc.solve(solver, "staggered", b, None, links=links, mass=m, lx=lx, ly=ly, eps=1e-12)
        assert info.success and relerr(x, xs) < 1e-9
    x, info = orc.solve("cg", "staggered_normal", b, None, links=links, mass=m, lx=lx, ly=ly, eps=1e-12)
    assert info.success and relerr(x, np.linalg.solve(D.conj().T @ D, b)) < 1e-9
    # resSq is the absolute squared true residual (generic_cg.cpp:217-218)
    r = b - orc.apply("staggered_normal", x, links, m, lx, ly)
    assert abs(info.resSq - np.vdot(r, r).real) <= 1e-6 * info.resSq


def test_multishift_equals_independent_solves(orc):
    """CG-M per shift equals independent CG on A + sigma (multishift.cpp:253-289)."""
    lx = ly = 16
    links, b, _ = lattice_inputs(orc, lx, ly, beta=6.25, seed=17)
    sh = np.array([0.0, 0.009, 0.004, 0.049, 0.999])
    X, info, res, _ = orc.cg_m("laplace_u1", b, sh, links=links, mass=0.001, lx=lx, ly=ly, eps=1e-12)
    for k, s in enumerate(sh):
        x, _ = orc.solve("cg", "laplace_u1", b, None, links=links, mass=0.001 + s, lx=lx, ly=ly, eps=1e-13)
        assert relerr(X[k], x) < 1e-8, (k, relerr(X[k], x))
        assert res[k] <= (1e-12 ** 2) * np.vdot(b, b).real * 1.01


def test_solver_failure_paths(orc):
    """success = false at max_iter; restart bookkeeping; ||b|| = 0 gives 0/0."""
    lx = ly = 8
    links, b, _ = lattice_inputs(orc, lx, ly, seed=1)
    _, info = orc.solve("cg", "staggered_normal", b, None, links=links, mass=0.1, lx=lx, ly=ly, max_iter=3)
    assert not info.success and info.iter == 3 and info.ops_count == 5
    _, info = orc.solve("bicgstab", "staggered", b, None, links=links, mass=0.1, lx=lx, ly=ly, max_iter=3)
    assert not info.success and info.iter == 3 and info.ops_count == 8
    _, info = orc.solve("gcr", "staggered", b, None, links=links, mass=0.1, lx=lx, ly=ly, max_iter=4,
                        restart_freq=2)
    assert info.name == "GCR(2)" and info.iter == 4 and not info.success
    _, info = orc.solve("cg", "staggered_normal", 0 * b, None, links=links, mass=0.1, lx=lx, ly=ly, max_iter=3)
    assert np.isnan(info.resSq) or info.resSq == 0.0


# ----------------------------------------------------------------- gradient
def test_force_matches_finite_differences(orc):
    """dS/dtheta and dS/dm of S = phi^dag (D^dag D)^-1 phi against central
    differences of S on a small lattice (dense solves)."""
    lx, ly, m = 6, 4, 0.3
    links, phi, _ = lattice_inputs(orc, lx, ly, beta=2.0, seed=40)
    n = lx * ly
    theta = np.angle(links)

    def S(th, mass):
        u = np.exp(1j * th)
        D = np.stack([orc.apply("staggered", e, u, mass, lx, ly) for e in np.eye(n, dtype=complex)], axis=1)
        return np.vdot(phi, np.linalg.solve(D.conj().T @ D, phi)).real, D

    s0, D = S(theta, m)
    chi = np.linalg.solve(D.conj().T @ D, phi)
    psi = D @ chi
    f, dm = orc.force(chi, psi, links, m, lx, ly)
    h = 1e-5
    for j in (0, 1, 7, 2 * n - 1, 13, 30):
        tp, tm = theta.copy(), theta.copy()
        tp[j] += h
        tm[j] -= h
        fd = (S(tp, m)[0] - S(tm, m)[0]) / (2 * h)
        assert abs(fd - f[j]) < 1e-6 * max(1.0, abs(fd)), (j, fd, f[j])
    fdm = (S(theta, m + h)[0] - S(theta, m - h)[0]) / (2 * h)
    assert abs(fdm - dm) < 1e-6 * abs(fdm)
