"""The operators.h functions off the solver path on the GPU (general stencils
built on the device from the links) against the oracle's function forms, and
solves that pass them as the operator.  The stencil form adds the same terms
with coefficients precomputed, so results agree to rounding (1e-13 on O(1)
values)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
OPS = ["staggered_free", "staggered_gamma5_free", "staggered_2linklaplace", "symmshift_x", "symmshift_y",
       "index_operator"]


@pytest.mark.parametrize("lx,ly", [(8, 6), (5, 7), (2, 2), (1, 4), (64, 48)])
@pytest.mark.parametrize("op", OPS)
def test_operator(glx, orc, op, lx, ly):
    u = orc.Rng(3).gauss_gauge(lx, ly, 2.0)
    v = orc.Rng(4).gaussian_complex(lx * ly)
    want = orc.apply(op, v, u, 0.3, lx, ly, wilson=0.7)
    got = glx.apply(op, v, u, 0.3, lx, ly, wilson=0.7)
    assert np.max(np.abs(got - want)) < 1e-13


def test_solves(glx, orc):
    L, m = 32, 0.2
    u = orc.Rng(11).gauss_gauge(L, L, 4.0)
    b = orc.Rng(12).gaussian_complex(L * L)
    x, info = glx.solve("bicgstab", "staggered_2linklaplace", b, links=u, mass=m, lx=L, ly=L, eps=1e-10, wilson=0.3)
    assert info.success
    r = b - orc.apply("staggered_2linklaplace", x, u, m, L, L, wilson=0.3)
    assert np.linalg.norm(r) <= 1.01e-10 * np.linalg.norm(b)
    # the index operator is Hermitian and indefinite: unrestarted GCR on a
    # 16^2 lattice (a Krylov space of at most 256 directions)
    L = 16
    u = orc.Rng(13).gauss_gauge(L, L, 4.0)
    b = orc.Rng(14).gaussian_complex(L * L)
    x, info = glx.solve("gcr", "index_operator", b, links=u, mass=m, lx=L, ly=L, eps=1e-10, max_iter=2000)
    assert info.success
    r = b - orc.apply("index_operator", x, u, m, L, L)
    assert np.linalg.norm(r) <= 1.01e-10 * np.linalg.norm(b)
