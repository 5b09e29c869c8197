"""GPU parity at the benchmark workloads' own sizes, against the oracle's
solves of the same seeded inputs (std::mt19937 streams drawn by the oracle,
as the reference's drivers draw them).

Two FP64 oracles: the reference's solvers as they are (serial left-to-right
sums, generic_vector.h:87-144), and the same solvers with correctly rounded
reductions (reductions="extended": long double, compensated, rounded once;
tests/test_oracle_extended_reductions.py pins it).  The GPU's reductions are
tree sums, so its FP64 solvers are held to the second at 1e-12, and must be
no farther from the first than the first is from the second: at these sizes
the serial sums alone move the reference's solution by 5e-12 (256^2) and
6e-12 (1024^2) from the correctly reduced one, more than the north star's
1e-12, while the GPU's tree sums stay within ~3e-13 of it.

  staggered 256^2, m = 0.05, theta ~ N(0, 0.4^2) (beta = 6.25):
      CG on D^dag D through the host-pointer C ABI (the drop-in): the
      oracle's iteration count, ops_count and success, x within 1e-12 of
      the correctly reduced oracle's;
      BiCGStab on D: inside the oracle's own envelope under 1-ulp
      perturbations of b (generic_bicgstab.cpp:205-341 is not stable under
      rounding: the spread is measured here, not assumed).
  staggered 1024^2, m = 0.02, theta ~ N(0, 0.5^2) (beta = 4):
      FP64 CG (reference order) with the oracle's iterations and x within
      1e-12 of the correctly reduced oracle's; single-pass FP64 CG; the
      mixed-precision CG in both forms (two iterations per launch and one)
      with FP32 and FP16 links, each with its true residual recomputed by
      the oracle's FP64 operator and its error against the oracle's FP64
      solution printed and bounded.
  multishift batch: one lattice for each L in {64, 128, 192, 256, 384, 512},
      m ~ U[0.01, 0.1], theta ~ N(0, 0.4^2), 4 right-hand sides each, the 12
      shifts sigma_k = 10^(-3 + 0.5 k), resid_freq_check 10, through one
      batched call; each of the 24 systems against its own oracle CG-M
      (generic_cg_m.cpp:312-600): its iterations and ops_count, every x_k
      within 1e-11 of the correctly reduced oracle's, every final |r_k|^2
      within 10% of the reference's beyond the two oracles' spread; a
      system whose limiting shift crosses the tolerance within rounding at
      a check may stop one check apart (at most two of the 24).

The oracle runs in threads (its C calls release the GIL), so the 1024^2 CG
(~20 s on one core) and the 24 multishift solves fit the per-test limit.
"""
import numpy as np
import pytest

from helpers import bicgstab_envelope, check_in_envelope, lattice_inputs, pmap, relerr

pytestmark = pytest.mark.gpu

# ------------------------------------------------------------ 256^2, FP64 --
L2, M2 = 256, 0.05


@pytest.fixture(scope="module")
def c2(orc):
    links, bc, _ = lattice_inputs(orc, L2, L2, beta=6.25, seed=2002)
    return links, bc


def oracle_pair(orc, solver, op, b, kw):
    """(serial, extended) oracle solves of the same system, in two threads."""
    return pmap(lambda red: orc.solve(solver, op, b, None, reductions=red, **kw), ["serial", "extended"])


def check_fp64_cg(what, xg, ig, serial, extended, tol=1e-12):
    (xs, i_s), (xe, i_e) = serial, extended
    d_se = relerr(xs, xe)
    d_ge, d_gs = relerr(xg, xe), relerr(xg, xs)
    print(f"{what}: iter {ig.iter} (oracle {i_s.iter}, extended {i_e.iter}); |x - x_ext|/|x_ext| = {d_ge:.2e}, "
          f"|x - x_ref|/|x_ref| = {d_gs:.2e}, reference vs extended {d_se:.2e}")
    assert d_ge <= tol, d_ge
    assert d_gs <= d_se + tol, (d_gs, d_se)


def test_c2_cg_normal_matches_oracle(glx, orc, c2):
    links, b = c2
    kw = dict(links=links, mass=M2, lx=L2, ly=L2, max_iter=100000, eps=1e-10)
    serial, extended = oracle_pair(orc, "cg", "staggered_normal", b, kw)
    io = serial[1]
    xg, ig = glx.solve("cg", "staggered_normal", b, None, **kw)
    assert (ig.iter, ig.ops_count, ig.success, ig.name) == (io.iter, io.ops_count, io.success, "CG")
    assert extended[1].iter == io.iter
    check_fp64_cg("C2 CG on D^dag D", xg, ig, serial, extended)
    assert abs(ig.resSq - io.resSq) <= 0.05 * io.resSq


def test_c2_bicgstab_inside_oracle_envelope(glx, orc, c2):
    links, b = c2
    kw = dict(links=links, mass=M2, lx=L2, ly=L2, max_iter=100000, eps=1e-10)
    x0, i0, its, dx = bicgstab_envelope(orc, "staggered", b, kw)
    xg, ig = glx.solve("bicgstab", "staggered", b, None, **kw)
    assert ig.success and i0.success and ig.name == "BiCGStab" and ig.ops_count == 2 * ig.iter + 2
    check_in_envelope(xg, ig, x0, its, dx, "C2 BiCGStab on D")
    # the reported residual is the true one of the returned x
    r = b - orc.apply("staggered", xg, links, M2, L2, L2)
    assert abs(np.sqrt(ig.resSq) - np.linalg.norm(r)) <= 1e-3 * np.linalg.norm(r)
    assert np.linalg.norm(r) <= 1e-10 * np.linalg.norm(b) * 1.0001


# ---------------------------------------------------- 1024^2, FP64 / mixed --
L3, M3 = 1024, 0.02


@pytest.fixture(scope="module")
def c3(glx, orc):
    links, bc, _ = lattice_inputs(orc, L3, L3, beta=4.0, seed=3003)
    kw = dict(links=links, mass=M3, lx=L3, ly=L3, max_iter=100000, eps=1e-10)
    serial, extended = oracle_pair(orc, "cg", "staggered_normal", bc, kw)
    lat = glx.Lattice(links, L3, L3, f32=True, f16=True)
    return links, bc, serial[0], serial[1], lat, serial, extended


def _true_relres(orc, links, b, x):
    r = b - orc.apply("staggered_normal", x, links, M3, L3, L3)
    return float(np.linalg.norm(r) / np.linalg.norm(b))


def test_c3_fp64_cg_matches_oracle(glx, orc, c3):
    """At cond(A) ~ 1e4 the iterated residual of any FP64 CG is fixed only to
    ~1 % by rounding, so the stop (relres < 1e-10) of two valid
    implementations may fall one iteration apart where the residual is within
    a few per cent of the tolerance; the solutions are then compared after
    the same number of iterations."""
    links, b, xo, io, _, serial, extended = c3
    kw = dict(links=links, mass=M3, lx=L3, ly=L3, max_iter=100000, eps=1e-10)
    xg, ig = glx.solve("cg", "staggered_normal", b, None, **kw)
    assert ig.success and ig.ops_count == ig.iter + 2
    if ig.iter != io.iter:
        assert abs(ig.iter - io.iter) == 1, (ig.iter, io.iter)
        _, io_h, hist = orc.solve("cg", "staggered_normal", b, None, hist=True, **dict(kw, max_iter=max(ig.iter,
                                                                                                   io.iter)))
        k = min(ig.iter, io.iter) - 1  # the oracle's relres where the two stops differ
        assert abs(hist[k] - 1e-10) <= 0.03e-10, hist[k]
        serial, extended = pmap(lambda red: orc.solve("cg", "staggered_normal", b, None, reductions=red,
                                                      **dict(kw, max_iter=ig.iter)), ["serial", "extended"])
        print(f"C3: stops one iteration apart (oracle relres {hist[k]:.4e} at iteration {k + 1}); compared after "
              f"{ig.iter} iterations")
    check_fp64_cg("C3 FP64 CG", xg, ig, serial, extended)


def test_c3_fp64_single_pass_cg(glx, orc, c3):
    links, b, xo, io, lat, _, extended = c3
    dx = glx.DeviceArray(np.zeros_like(b))
    info = glx.dev_cg_fused(lat, "staggered_normal", M3, dx, glx.DeviceArray(b), max_iter=100000, eps=1e-10)
    x = dx.download()
    rr = _true_relres(orc, links, b, x)
    print(f"C3 single-pass FP64 CG: iter {info.iter} (oracle {io.iter}), true relres {rr:.3e}, "
          f"|x - x_o|/|x_o| = {relerr(x, xo):.2e}, |x - x_ext|/|x_ext| = {relerr(x, extended[0]):.2e}")
    assert info.success and rr <= 1e-10 * 1.001
    assert abs(info.iter - io.iter) <= 2
    assert relerr(x, xo) <= 1e-11


@pytest.mark.parametrize("sstep", [2, 1])
@pytest.mark.parametrize("link_prec", [1, 2])
def test_c3_mixed_against_oracle(glx, orc, c3, sstep, link_prec):
    """Mixed precision (FP32 bulk, FP32 or FP16 links, FP64 reliable updates
    at delta 0.1): the oracle's FP64 operator recomputes the true residual of
    the returned x (<= 1e-10, equal to the one the solver reports), and x is
    within 5e-11 of the oracle's FP64 solution (measured 5e-12: both
    solutions carry residuals of 1e-10 whose error components differ)."""
    links, b, xo, io, lat, _, _ = c3
    glx.set_sstep(sstep)
    try:
        dx = glx.DeviceArray(np.zeros_like(b))
        info = glx.dev_cg_mixed(lat, "staggered_normal", M3, dx, glx.DeviceArray(b), max_iter=100000, eps=1e-10,
                                delta=0.1, link_prec=link_prec)
    finally:
        glx.set_sstep()
    x = dx.download()
    rr = _true_relres(orc, links, b, x)
    err = relerr(x, xo)
    print(f"C3 mixed (s={sstep}, links {'FP32' if link_prec == 1 else 'FP16'}): iter {info.iter} "
          f"(oracle FP64 {io.iter}), reliable updates {info.reliable_updates}, true relres {rr:.3e}, "
          f"|x - x_o|/|x_o| = {err:.2e}")
    assert info.success and rr <= 1e-10 * 1.001
    assert abs(np.sqrt(info.resSq) / np.linalg.norm(b) - rr) <= 1e-3 * rr
    assert err <= 5e-11
    assert info.iter <= 1.2 * io.iter


# ------------------------------------------------- multishift batch sizes --
SHIFTS = 10.0 ** (-3.0 + 0.5 * np.arange(12))
LS = [64, 128, 192, 256, 384, 512]


def test_c4_batch_every_size_matches_oracle(glx, orc):
    rng = np.random.default_rng(4004)
    systems, ref = [], []
    for n, L in enumerate(LS):
        links, b0, _ = lattice_inputs(orc, L, L, beta=6.25, seed=4100 + n)
        lat = glx.Lattice(links, L, L, f32=False, f16=False)
        mass = float(rng.uniform(0.01, 0.1))
        for r in range(4):
            b = b0 if r == 0 else orc.Rng(4200 + 10 * n + r).gaussian_complex(L * L)
            xs = [glx.DeviceArray(np.zeros(L * L, np.complex128)) for _ in SHIFTS]
            systems.append((lat, mass, glx.DeviceArray(b), xs))
            ref.append((links, b, mass, L))
    infos, st = glx.dev_cg_m_batched(systems, SHIFTS, resid_freq_check=10, max_iter=100000, eps=1e-10,
                                     stats=True)
    jobs = [(a, red) for a in ref for red in ("serial", "extended")]
    runs = pmap(lambda j: orc.cg_m("staggered_normal", j[0][1], SHIFTS, links=j[0][0], mass=j[0][2], lx=j[0][3],
                                   ly=j[0][3], resid_freq_check=10, max_iter=100000, eps=1e-10, reductions=j[1]),
                jobs)
    orc_runs, ext_runs = runs[0::2], runs[1::2]
    tol = 1e-10
    n_exact = 0
    for (lat, mass, db, xs), (links, b, m, L), info, (X, io, ro, so), (Xe, ie, re, _) in zip(systems, ref, infos,
                                                                                             orc_runs, ext_runs):
        assert np.array_equal(so, SHIFTS)
        assert info.resSq == 0.0 and info.name == "CG-M" and info.success == io.success
        bn2 = float(np.vdot(b, b).real)
        xg = [xs[k].download() for k in range(SHIFTS.size)]
        errs = [relerr(xg[k], X[k]) for k in range(SHIFTS.size)]
        errs_e = [relerr(xg[k], Xe[k]) for k in range(SHIFTS.size)]
        d_se = [relerr(X[k], Xe[k]) for k in range(SHIFTS.size)]
        kap = [(m * m + s + 4.0) / (m * m + s) for s in SHIFTS]
        print(f"L={L} m={m:.4f}: iter {info.iter} (oracle {io.iter}, extended {ie.iter}), max_k |x_k - x_ref,k|"
              f"/|x_ref,k| = {max(errs):.2e}, vs extended {max(errs_e):.2e} (reference vs extended {max(d_se):.2e})")
        if info.iter == io.iter:
            n_exact += 1
            assert info.ops_count == io.ops_count
            for k in range(SHIFTS.size):
                if ie.iter == io.iter:
                    # within 1e-11 of the correctly reduced oracle, and no
                    # farther from the reference than it is
                    assert errs_e[k] < 1e-11, (L, k, errs_e[k])
                    assert errs[k] < d_se[k] + 1e-11, (L, k, errs[k], d_se[k])
                    # the final |r_k|^2 of a converged shift moves with the
                    # rounding of the sums (the oracle's serial and correctly
                    # rounded sums give values up to 1.8x apart on these
                    # systems' shifts at L = 64 and 128): within 10% of the
                    # reference's beyond that spread
                    spread = abs(re[k] - ro[k])
                    assert abs(info.resSqmrhs[k] - ro[k]) <= spread + 0.1 * ro[k] + 1e-24 * bn2, \
                        (k, info.resSqmrhs[k], ro[k], re[k])
                else:
                    # the correctly reduced oracle stopped at another check
                    # (its limiting shift crossed the tolerance within
                    # rounding): the reference's own stop, bounded as below
                    assert abs(ie.iter - io.iter) == 10, (L, ie.iter, io.iter)
                    assert errs[k] < 2 * tol * kap[k], (L, k)
                    assert np.sqrt(info.resSqmrhs[k] / bn2) <= tol * 1.01
        else:
            # CG-M tests convergence every resid_freq_check iterations: a
            # limiting shift whose residual reaches the tolerance within
            # rounding at a check may stop one check apart (both solutions
            # then meet the tolerance)
            assert abs(info.iter - io.iter) == 10, (L, info.iter, io.iter)
            assert info.ops_count - io.ops_count == info.iter - io.iter
            first = info.resSqmrhs if info.iter < io.iter else ro
            assert np.sqrt(max(first) / bn2) >= 0.95 * tol, (L, np.sqrt(max(first) / bn2))
            for k in range(SHIFTS.size):
                assert errs[k] < 2 * tol * kap[k], (L, k)
                assert np.sqrt(info.resSqmrhs[k] / bn2) <= tol * 1.01 and np.sqrt(ro[k] / bn2) <= tol * 1.01
    assert n_exact >= len(systems) - 2, n_exact
    # the reported work: every system's iterations; x_s updates (one per
    # active shift and window) at most 12 per iteration, p_s updates at most
    # as many; D^dag D 64 B per site in the first iteration, 96 B (p formed
    # from r and the last p) in every later one
    assert st["iterations"] == sum(i.iter for i in infos)
    assert st["shift_updates_x"] <= 12 * st["iterations"] and st["shift_updates_p"] <= st["shift_updates_x"]
    assert st["bytes_stencil"] == sum(lat.lx * lat.ly * (64.0 + 96.0 * (i.iter - 1))
                                      for (lat, _, _, _), i in zip(systems, infos))
    assert st["bytes_r_pass"] == sum(48.0 * lat.lx * lat.ly * i.iter for (lat, _, _, _), i in zip(systems, infos))
