"""Batched multishift CG (glx_dev_cg_m_batched): many independent CG-M
solves (generic_cg_m.cpp:312-600 per system) on staggered D^dag D lattices of
different sizes and masses, advanced together -- one launch per phase over
all unconverged systems, iteration groups replayed from a hipGraph.

Each system must reproduce the oracle's own CG-M run: the same iteration
count, ops_count and success, per-shift solutions to 1e-11 and per-shift
true residuals, whatever the other systems in the batch do; shifts are
passed out of order (as tests/multishift/multishift.cpp:224-232 does) to
exercise the in-solver permutation.  CG-M tests convergence only every
resid_freq_check iterations, so when the limiting shift's residual reaches
the tolerance within rounding at a check (the side that stops first shows a
true relative residual within 5% of it), the device and the oracle may stop
one check interval apart; then both solutions satisfy the tolerance and are
compared to the accuracy it allows, 2 tol kappa_k with
kappa_k = (m^2 + sigma_k + 4) / (m^2 + sigma_k).  Everything else must match
exactly.
"""
import numpy as np
import pytest

from helpers import lattice_inputs, relerr

pytestmark = pytest.mark.gpu

SHIFTS = 10.0 ** (-3.0 + 0.5 * np.array([0, 2, 1, 3, 5, 4, 6, 8, 7, 9, 11, 10]))


def make_batch(glx, orc, shapes, rhs_per_lattice, seed=5):
    rng = np.random.default_rng(seed)
    systems, ref = [], []
    for n, (lx, ly) in enumerate(shapes):
        links, b0, _ = lattice_inputs(orc, lx, ly, beta=6.25, seed=100 + n)
        lat = glx.Lattice(links, lx, ly, f32=False, f16=False)
        mass = float(rng.uniform(0.01, 0.1))
        for r in range(rhs_per_lattice):
            b = b0 if r == 0 else orc.Rng(200 + 10 * n + r).gaussian_complex(lx * ly)
            xs = [glx.DeviceArray(np.zeros(lx * ly, np.complex128)) for _ in SHIFTS]
            db = glx.DeviceArray(b)
            systems.append((lat, mass, db, xs))
            ref.append((links, b, mass, lx, ly))
    return systems, ref


@pytest.mark.parametrize("use_graph", [True, False])
def test_batched_multishift_matches_oracle(glx, orc, use_graph):
    shapes = [(64, 64), (128, 64), (64, 192), (192, 128), (128, 8), (256, 64)]
    systems, ref = make_batch(glx, orc, shapes, rhs_per_lattice=2)
    infos = glx.dev_cg_m_batched(systems, SHIFTS, resid_freq_check=10, max_iter=4000, eps=1e-10,
                                 use_graph=use_graph)
    assert len(infos) == len(systems)
    for (lat, mass, db, xs), (links, b, m, lx, ly), info in zip(systems, ref, infos):
        X, io, ro, so = orc.cg_m("staggered_normal", b, SHIFTS, links=links, mass=m, lx=lx, ly=ly,
                                 resid_freq_check=10, max_iter=4000, eps=1e-10)
        assert info.resSq == 0.0 and info.name == "CG-M" and info.success == io.success
        bn2 = float(np.vdot(b, b).real)
        tol = 1e-10
        if info.iter == io.iter:
            assert info.ops_count == io.ops_count
            for k in range(SHIFTS.size):
                assert relerr(xs[k].download(), X[k]) < 1e-11, (lx, ly, k)
                assert abs(info.resSqmrhs[k] - ro[k]) <= 0.1 * ro[k] + 1e-24 * bn2, (k, info.resSqmrhs[k], ro[k])
        else:
            # one check interval apart at a marginal threshold crossing
            assert abs(info.iter - io.iter) == 10, (lx, ly, info.iter, io.iter)
            assert info.ops_count - io.ops_count == info.iter - io.iter
            first = info.resSqmrhs if info.iter < io.iter else ro
            assert np.sqrt(max(first) / bn2) >= 0.95 * tol, (lx, ly, np.sqrt(max(first) / bn2))
            for k in range(SHIFTS.size):
                kap = (m * m + SHIFTS[k] + 4.0) / (m * m + SHIFTS[k])
                assert relerr(xs[k].download(), X[k]) < 2 * tol * kap, (lx, ly, k)
                assert np.sqrt(info.resSqmrhs[k] / bn2) <= tol * 1.01 and np.sqrt(ro[k] / bn2) <= tol * 1.01


def test_batched_equals_single_solves(glx, orc):
    """The batch gives each system what a solve of its own gives: the same
    iterations and ops_count, and solutions equal to the rounding of their
    different reduction orders (block counts) and of the batch's windowed
    shift updates, which move x_k by ~1e-12 here (the bar of the oracle
    comparisons, 1e-11)."""
    shapes = [(64, 64), (128, 128)]
    systems, ref = make_batch(glx, orc, shapes, rhs_per_lattice=1, seed=9)
    infos = glx.dev_cg_m_batched(systems, SHIFTS, max_iter=4000)
    for (lat, mass, db, xs), info in zip(systems, infos):
        ys = [glx.DeviceArray(np.zeros(lat.lx * lat.ly, np.complex128)) for _ in SHIFTS]
        one = glx.dev_cg_m(lat, "staggered_normal", mass, ys, db, SHIFTS, max_iter=4000)
        assert (info.iter, info.ops_count) == (one.iter, one.ops_count)
        for k in range(SHIFTS.size):
            assert relerr(xs[k].download(), ys[k].download()) < 1e-11


def test_batched_rejects_unsupported_lattice(glx, orc):
    links, b, _ = lattice_inputs(orc, 24, 10, beta=6.25, seed=3)
    lat = glx.Lattice(links, 24, 10, f32=False, f16=False)
    xs = [glx.DeviceArray(np.zeros(240, np.complex128)) for _ in SHIFTS]
    with pytest.raises(glx.GlxError, match="lx % 64"):
        glx.dev_cg_m_batched([(lat, 0.1, glx.DeviceArray(b), xs)], SHIFTS)
