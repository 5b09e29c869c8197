"""The oracle's restatement of the operators.h functions off the solver path
(operators.cpp:127-846): the free staggered operator and its gamma5 form,
staggered + two-link Laplacian, the symmetric shifts and the index operator.

Pinned by reference-written fixtures and the reference build (bit for bit),
and by structure: the free operators are the U(1) ones on unit links, the
two-link term vanishes at wilson_coeff 0, the symmetric shifts and the index
operator are Hermitian, and at m = 0 the index operator is i D."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
G = np.load(os.path.join(HERE, "golden", "golden.npz"))
OPS = ["staggered_free", "staggered_gamma5_free", "staggered_2linklaplace", "symmshift_x", "symmshift_y",
       "index_operator"]


def same_bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("op", OPS)
def test_golden(orc, op):
    out = orc.apply(op, G["op_in"], G["op_links"], 0.37, 8, 6, wilson=0.7)
    assert same_bits(out, G[f"op_out_{op}"])


def test_reference_build(orc):
    if not orc.ref_available():
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    for lx, ly in ((8, 6), (5, 7), (2, 2), (1, 4), (16, 16)):
        u = orc.Rng(3).gauss_gauge(lx, ly, 2.0)
        v = orc.Rng(4).gaussian_complex(lx * ly)
        for op in OPS:
            a = orc.apply(op, v, u, 0.3, lx, ly, wilson=0.7)
            b = orc.apply(op, v, u, 0.3, lx, ly, impl="ref", wilson=0.7)
            assert same_bits(a, b), (op, lx, ly)


def test_structure(orc):
    lx, ly, m = 8, 6, 0.3
    u = orc.Rng(5).gauss_gauge(lx, ly, 2.0)
    one = np.ones(2 * lx * ly, np.complex128)
    rng = np.random.default_rng(0)
    v = rng.standard_normal(lx * ly) + 1j * rng.standard_normal(lx * ly)
    w = rng.standard_normal(lx * ly) + 1j * rng.standard_normal(lx * ly)
    assert np.allclose(orc.apply("staggered_free", v, None, m, lx, ly), orc.apply("staggered", v, one, m, lx, ly),
                       rtol=0, atol=1e-15)
    assert np.allclose(orc.apply("staggered_gamma5_free", v, None, m, lx, ly),
                       orc.apply("staggered_gamma5", v, one, m, lx, ly), rtol=0, atol=1e-15)
    assert np.allclose(orc.apply("staggered_2linklaplace", v, u, m, lx, ly, wilson=0.0),
                       orc.apply("staggered", v, u, m, lx, ly), rtol=0, atol=1e-15)
    for op in ("symmshift_x", "symmshift_y", "index_operator"):
        lhs = np.vdot(w, orc.apply(op, v, u, m, lx, ly))
        rhs = np.vdot(orc.apply(op, w, u, m, lx, ly), v)
        assert abs(lhs - rhs) < 1e-12, op
    assert np.allclose(orc.apply("index_operator", v, u, 0.0, lx, ly), 1j * orc.apply("staggered", v, u, 0.0, lx, ly),
                       rtol=0, atol=1e-15)
    # the two-link term alone is the U-transported second difference
    d = orc.apply("staggered_2linklaplace", v, u, m, lx, ly, wilson=1.0) - orc.apply("staggered", v, u, m, lx, ly)
    U = u.reshape(ly, lx, 2)
    V = v.reshape(ly, lx)
    want = 4 * V - U[:, :, 0] * np.roll(U[:, :, 0], -1, 1) * np.roll(V, -2, 1) \
        - np.conj(np.roll(U[:, :, 0], 1, 1)) * np.conj(np.roll(U[:, :, 0], 2, 1)) * np.roll(V, 2, 1) \
        - U[:, :, 1] * np.roll(U[:, :, 1], -1, 0) * np.roll(V, -2, 0) \
        - np.conj(np.roll(U[:, :, 1], 1, 0)) * np.conj(np.roll(U[:, :, 1], 2, 0)) * np.roll(V, 2, 0)
    assert np.allclose(d.reshape(ly, lx), want, rtol=0, atol=1e-13)
