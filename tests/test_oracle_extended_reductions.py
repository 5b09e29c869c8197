"""The oracle's extended-reduction solvers (orc_*_cx): the reference's
complex solvers with every dot product and norm accumulated in long double
and rounded once.  Pinned here, on the CPU, against math.fsum (exactly
rounded sums of the exactly computed products):
  * each reduction equals the correctly rounded value to within one unit in
    the last place, where the serial double sum of the reference may not;
  * CG with these reductions equals a numpy CG in the reference's order whose
    reductions are math.fsum sums (same iterations, x to 1e-14);
  * on the 256^2 staggered normal system, the reference's serial sums move
    its solution by ~5e-12 from this one (the reason the GPU's FP64 CG, whose
    reductions are tree sums, is held to this variant at 1e-12).
"""
import math

import numpy as np
import pytest

from helpers import lattice_inputs, relerr


def fsum_dot(a, b):
    ar, ai, br, bi = a.real, a.imag, b.real, b.imag
    return complex(math.fsum(np.concatenate([ar * br, ai * bi])), math.fsum(np.concatenate([ar * bi, -ai * br])))


def _vec(rng, n):
    v = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    # 26 significant bits: products are exact doubles, so fsum is exact
    re = np.ldexp(np.round(np.ldexp(v.real, 25)), -25)
    im = np.ldexp(np.round(np.ldexp(v.imag, 25)), -25)
    return re + 1j * im


@pytest.mark.parametrize("n", [1, 17, 4096, 200000])
def test_extended_reductions_are_correctly_rounded(orc, n):
    rng = np.random.default_rng(n)
    a, b = _vec(rng, n), _vec(rng, n)
    want = fsum_dot(a, b)
    got = orc.blas("dot", a, b, impl="orcx")
    for g, w in ((got.real, want.real), (got.imag, want.imag)):
        assert abs(g - w) <= np.spacing(abs(w)), (g, w)
    nw = math.fsum(np.concatenate([a.real * a.real, a.imag * a.imag]))
    assert abs(orc.blas("norm2sq", a, impl="orcx") - nw) <= np.spacing(nw)
    d = a - b
    dw = math.fsum(np.concatenate([d.real * d.real, d.imag * d.imag]))
    assert abs(orc.blas("diffnorm2sq", a, b, impl="orcx") - dw) <= np.spacing(dw) * 2


def test_serial_sum_is_not_correctly_rounded(orc):
    """The reference's serial sum misses the correctly rounded value on a long
    vector (which the extended variant hits): the two reductions differ."""
    rng = np.random.default_rng(3)
    a, b = _vec(rng, 200000), _vec(rng, 200000)
    want = fsum_dot(a, b)
    ser = orc.blas("dot", a, b)
    assert ser != want and orc.blas("dot", a, b, impl="orcx") == pytest.approx(want, rel=1e-16)


def _np_cg(A, b, maxit, eps, dot):
    """CG in the reference's order (generic_cg.cpp:278-382) with `dot`."""
    x = np.zeros_like(b)
    r = b - A(x)
    p = r.copy()
    Ap = A(p)
    bn = math.sqrt(dot(b, b).real)
    rsq = dot(r, r).real
    for k in range(maxit):
        alpha = rsq / dot(p, Ap)
        x = x + alpha * p
        r = r - alpha * Ap
        rn = dot(r, r).real
        if math.sqrt(rn) < eps * bn or k == maxit - 1:
            break
        beta = rn / rsq
        rsq = rn
        p = r + beta * p
        Ap = A(p)
    return x, k + 1


def test_extended_cg_tracks_fsum_cg(orc):
    """At 256^2 the extended CG and a numpy CG in the reference's order with
    math.fsum reductions take the same iterations and agree to the level of
    their other scalar roundings (numpy's complex division is not the
    oracle's), an order of magnitude closer than the serial-sum reference."""
    L, m = 256, 0.05
    links, b, _ = lattice_inputs(orc, L, L, beta=6.25, seed=2002)
    A = lambda v: orc.apply("staggered_normal", v, links, m, L, L)
    xn, itn = _np_cg(A, b, 100000, 1e-10, fsum_dot)
    kw = dict(links=links, mass=m, lx=L, ly=L, max_iter=100000, eps=1e-10)
    xe, ie = orc.solve("cg", "staggered_normal", b, None, reductions="extended", **kw)
    xs, i_s = orc.solve("cg", "staggered_normal", b, None, **kw)
    assert ie.iter == itn == i_s.iter
    assert relerr(xe, xn) < 5e-13
    assert relerr(xs, xn) > 5 * relerr(xe, xn)


def test_serial_reference_moves_from_extended(orc):
    """The size of the effect at the 256^2 workload: the reference's own CG
    (serial sums) and the extended variant take the same iterations, and
    their solutions differ by more than 1e-12."""
    L, m = 256, 0.05
    links, b, _ = lattice_inputs(orc, L, L, beta=6.25, seed=2002)
    kw = dict(links=links, mass=m, lx=L, ly=L, max_iter=100000, eps=1e-10)
    xs, i_s = orc.solve("cg", "staggered_normal", b, None, **kw)
    xe, i_e = orc.solve("cg", "staggered_normal", b, None, reductions="extended", **kw)
    assert i_s.iter == i_e.iter
    d = relerr(xs, xe)
    assert 1e-12 < d < 5e-11, d
