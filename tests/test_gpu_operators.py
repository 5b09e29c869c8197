"""GPU parity of every lattice operator against the oracle.

The device kernels compute the reference's operators (operators.cpp:28-616,
coarse_stencil.cpp:12-722, operators_stencil.cpp:14-236) with FMA contraction
and a different but exact-in-real-arithmetic term grouping (e.g. D^dag D as
m^2 - K^2), so agreement is to rounding: relative 1e-14 in the 2-norm.
Lattice shapes include non-multiples of the 64 x 8 tile and lattices smaller
than one tile (wrap-around halos).
"""
import numpy as np
import pytest

from helpers import lattice_inputs, relerr

pytestmark = pytest.mark.gpu

SHAPES = [(16, 16), (64, 8), (24, 10), (4, 4), (130, 18), (8, 6)]
OPS = ["laplace", "laplace_u1", "staggered", "staggered_dagger", "staggered_normal", "staggered_gamma5",
       "gamma5", "deo", "doe", "m2mdeodoe", "stencil", "stencil_dagger", "stencil_gamma5", "stencil_eo",
       "stencil_oe", "stencil_m2mdeodoe"]
TOL = 1e-14


@pytest.mark.parametrize("lx,ly", SHAPES)
@pytest.mark.parametrize("op", OPS)
def test_operator_matches_oracle(glx, orc, op, lx, ly):
    links, bc, _ = lattice_inputs(orc, lx, ly, beta=2.0, seed=7 + lx)
    mass = 0.37
    want = orc.apply(op, bc, links, mass, lx, ly)
    got = glx.apply(op, bc, links, mass, lx, ly)
    assert relerr(got, want) < TOL, (op, lx, ly, relerr(got, want))


@pytest.mark.parametrize("lx,ly", [(16, 7), (24, 9), (130, 11)])
def test_m2mdeodoe_odd_ly(glx, orc, lx, ly):
    """m^2 - D_eo D_oe where ly is odd: across the periodic wrap an even
    site's neighbour is even, so the intermediate D_oe p must be zero on
    the even sites (it is the oracle's D_oe, operators.cpp:499-539)."""
    links, bc, _ = lattice_inputs(orc, lx, ly, beta=2.0, seed=11 + ly)
    want = orc.apply("m2mdeodoe", bc, links, 0.37, lx, ly)
    got = glx.apply("m2mdeodoe", bc, links, 0.37, lx, ly)
    assert relerr(got, want) < TOL, (lx, ly, relerr(got, want))


@pytest.mark.parametrize("lx,ly,nc", [(16, 16, 1), (24, 10, 1), (8, 6, 2), (5, 7, 3)])
def test_real_laplace_matches_oracle(glx, orc, lx, ly, nc):
    g = orc.Rng(3)
    b = g.gaussian_real(lx * ly * nc)
    want = orc.apply("laplace_real", b, None, 0.01, lx, ly, nc)
    got = glx.apply("laplace_real", b, None, 0.01, lx, ly, nc)
    assert relerr(got, want) < TOL


def test_complex_laplace_colours(glx, orc):
    g = orc.Rng(5)
    b = g.gaussian_complex(6 * 4 * 2)
    want = orc.apply("laplace", b, None, 0.3, 6, 4, 2)
    got = glx.apply("laplace", b, None, 0.3, 6, 4, 2)
    assert relerr(got, want) < TOL


@pytest.mark.parametrize("stencil", [False, True])
def test_eoprec_prepare_reconstruct(glx, orc, stencil):
    lx, ly, m = 16, 12, 0.2
    links, bc, _ = lattice_inputs(orc, lx, ly, beta=3.0, seed=11)
    xe = orc.Rng(12).gaussian_complex(lx * ly)
    assert relerr(glx.eoprec_prepare(bc, links, m, lx, ly, stencil),
                  orc.eoprec_prepare(bc, links, m, lx, ly, stencil)) < TOL
    assert relerr(glx.eoprec_reconstruct(xe, bc, links, m, lx, ly, stencil),
                  orc.eoprec_reconstruct(xe, bc, links, m, lx, ly, stencil)) < TOL


@pytest.mark.parametrize("op", ["stencil", "stencil_dagger", "stencil_gamma5"])
def test_stencil_coefficients(glx, orc, op):
    lx, ly = 12, 10
    links, _, _ = lattice_inputs(orc, lx, ly, beta=2.0, seed=9)
    a = glx.stencil_coeffs(op, links, 0.25, lx, ly)
    b = orc.stencil_coeffs(op, links, 0.25, lx, ly)
    for x, y in zip(a, b):
        # coefficients are products of exact constants and one link: exact
        assert np.array_equal(x, y)


def test_large_lattice_normal_operator(glx, orc):
    """The fused D^dag D at the 1024^2 workload size against the oracle."""
    lx = ly = 1024
    links, bc, _ = lattice_inputs(orc, lx, ly, beta=4.0, seed=21)
    got = glx.apply("staggered_normal", bc, links, 0.02, lx, ly)
    want = orc.apply("staggered_normal", bc, links, 0.02, lx, ly)
    assert relerr(got, want) < TOL
