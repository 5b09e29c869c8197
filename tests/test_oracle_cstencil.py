"""The oracle's general stencil_2d (any Nc, two-link terms, single-direction
sdir, pieces all / eo / oe / tb / bt, the even/odd helpers and the generator;
coarse_stencil.cpp:12-1627, operators_stencil.cpp:179-236).

Pinned by
- fixtures the reference build wrote (tests/golden/make_golden.py): bit for bit;
- the reference build itself where oracle/_ref exists: every (Nc, two-link,
  sdir, piece) combination, bit for bit;
- structure that fixes the layout and the term selection: a colour-diagonal
  Nc = 2 stencil is two Nc = 1 stencils; the single-direction pieces add up
  to the full stencil; tb + bt + the same-half blocks give the full stencil;
  eo + oe + clover give the full one-link stencil; generating a stencil from a
  stencil reproduces it; Nc = 1 agrees with the one-link oracle (itself pinned
  against the staggered operator).
"""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
G = np.load(os.path.join(HERE, "golden", "golden.npz"))
CST_CASES = [("all", 0), ("all", 1), ("all", 3), ("all", 7), ("eo", 0), ("oe", 4), ("eo", 9), ("tb", 0), ("bt", 0),
             ("tb", 12), ("bt", 1)]


def same_bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def golden_stencil(orc):
    sh = G["cst_shifts"]
    return orc.CStencil(5, 4, 3, G["cst_clover"], G["cst_hopping"], G["cst_two"], complex(sh[0], sh[1]),
                        complex(sh[2], sh[3]), complex(sh[4], sh[5]))


@pytest.mark.parametrize("piece,sdir", CST_CASES)
def test_golden_pieces(orc, piece, sdir):
    st = golden_stencil(orc)
    st.sdir = sdir
    assert same_bits(orc.cstencil_apply(st, G["cst_in"], piece, lhs=G["cst_in2"]), G[f"cst_out_{piece}_{sdir}"])


def test_golden_eo_helpers_and_generator(orc):
    st = golden_stencil(orc)
    v, w = G["cst_in"], G["cst_in2"]
    assert same_bits(orc.cstencil_m2mdeodoe(st, v), G["cst_m2mdeodoe"])
    assert same_bits(orc.cstencil_eoprec_prepare(st, v), G["cst_prepare"])
    assert same_bits(orc.cstencil_eoprec_reconstruct(st, v, w), G["cst_reconstruct"])
    gen = orc.cstencil_generate(6, 5, 1, True, op="staggered_normal", links=G["cst_gen_links"], mass=0.2)
    assert same_bits(gen.clover, G["cst_gen_clover"])
    assert same_bits(gen.hopping, G["cst_gen_hopping"])
    assert same_bits(gen.two_link, G["cst_gen_two"])


def test_reference_build_all_combinations(orc):
    if not orc.ref_available():
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    rng = np.random.default_rng(0)
    for nc in (1, 2, 4):
        for has_two in (False, True):
            for sdir in range(14):
                for piece in orc.PIECES:
                    st = orc.CStencil.random(rng, 5, 6, nc, has_two, shifts=(0.3 + 0.1j, -0.2j, 0.7), sdir=sdir)
                    v = rng.standard_normal(st.n) + 1j * rng.standard_normal(st.n)
                    init = rng.standard_normal(st.n) + 1j * rng.standard_normal(st.n)
                    a = orc.cstencil_apply(st, v, piece, lhs=init)
                    b = orc.cstencil_apply(st, v, piece, impl="ref", lhs=init)
                    assert same_bits(a, b), (nc, has_two, sdir, piece)


def _colour_blocks(st, orc, keep):
    """A copy of st whose (row, column) entries outside keep(r, c) are zero."""
    nc = st.nc
    r, c = np.meshgrid(np.arange(nc), np.arange(nc), indexing="ij")
    mask = keep(r, c).astype(float).reshape(-1)

    def cut(a):
        return None if a is None else (a.reshape(-1, nc * nc) * mask).reshape(-1)

    return orc.CStencil(st.lx, st.ly, nc, cut(st.clover), cut(st.hopping), cut(st.two_link), sdir=st.sdir)


def test_colour_diagonal_is_independent_copies(orc):
    rng = np.random.default_rng(1)
    lx, ly = 6, 5
    a = orc.CStencil.random(rng, lx, ly, 1, True)
    b = orc.CStencil.random(rng, lx, ly, 1, True)
    V = lx * ly

    def merge(x, y):  # nc = 2 block-diagonal coefficients from two nc = 1 stencils
        t = x.size // V
        out = np.zeros((t, V, 2, 2), np.complex128)
        out[:, :, 0, 0] = x.reshape(t, V)
        out[:, :, 1, 1] = y.reshape(t, V)
        return out.reshape(-1)

    st = orc.CStencil(lx, ly, 2, merge(a.clover, b.clover), merge(a.hopping, b.hopping),
                      merge(a.two_link, b.two_link))
    v = rng.standard_normal(2 * V) + 1j * rng.standard_normal(2 * V)
    out = orc.cstencil_apply(st, v).reshape(V, 2)
    assert np.allclose(out[:, 0], orc.cstencil_apply(a, v.reshape(V, 2)[:, 0].copy()), rtol=0, atol=1e-13)
    assert np.allclose(out[:, 1], orc.cstencil_apply(b, v.reshape(V, 2)[:, 1].copy()), rtol=0, atol=1e-13)


@pytest.mark.parametrize("nc", [1, 3, 4])
def test_pieces_add_up(orc, nc):
    rng = np.random.default_rng(2 + nc)
    st = orc.CStencil.random(rng, 7, 6, nc, True)
    v = rng.standard_normal(st.n) + 1j * rng.standard_normal(st.n)
    full = orc.cstencil_apply(st, v)
    parts = np.zeros_like(full)
    for sdir in range(1, 14):
        st.sdir = sdir
        parts += orc.cstencil_apply(st, v)
    st.sdir = 0
    assert np.allclose(parts, full, rtol=0, atol=1e-12)
    if nc > 1:
        h = nc // 2
        same_half = _colour_blocks(st, orc, lambda r, c: (r < h) == (c < h))
        tb_bt = orc.cstencil_apply(st, v, "tb") + orc.cstencil_apply(st, v, "bt")
        assert np.allclose(tb_bt + orc.cstencil_apply(same_half, v), full, rtol=0, atol=1e-12)
        # tb only writes the top rows, bt only the bottom rows
        assert not orc.cstencil_apply(st, v, "tb").reshape(-1, nc)[:, h:].any()
        assert not orc.cstencil_apply(st, v, "bt").reshape(-1, nc)[:, :h].any()
    one = orc.CStencil(st.lx, st.ly, nc, st.clover, st.hopping)
    clover_only = orc.CStencil(st.lx, st.ly, nc, st.clover, sdir=1)
    eo_oe = orc.cstencil_apply(one, v, "eo") + orc.cstencil_apply(one, v, "oe")
    assert np.allclose(eo_oe + orc.cstencil_apply(clover_only, v), orc.cstencil_apply(one, v), rtol=0, atol=1e-12)


@pytest.mark.parametrize("nc,has_two", [(1, False), (2, True), (3, False)])
def test_generator_reproduces_stencil(orc, nc, has_two):
    rng = np.random.default_rng(10 + nc)
    src = orc.CStencil.random(rng, 6, 5, nc, has_two, shifts=(0.25, 0.1j, -0.3))
    gen = orc.cstencil_generate(6, 5, nc, has_two, src=src)
    # shifts land on the clover diagonal
    diag = np.zeros((6 * 5, nc, nc), np.complex128)
    idx = np.arange(nc)
    y, x = np.divmod(np.arange(30), 6)
    even = ((x + y) % 2 == 0)[:, None]
    top = (idx < nc // 2)[None, :]
    diag[:, idx, idx] = src.shift + np.where(even, 1, -1) * src.eo_shift + np.where(top, 1, -1) * src.dof_shift
    assert np.allclose(gen.clover, src.clover + diag.reshape(-1), rtol=0, atol=1e-14)
    assert np.allclose(gen.hopping, src.hopping, rtol=0, atol=1e-14)
    if has_two:
        assert np.allclose(gen.two_link, src.two_link, rtol=0, atol=1e-14)


def test_nc1_matches_one_link_oracle(orc):
    links = orc.Rng(4).gauss_gauge(8, 6, 2.0)
    v = orc.Rng(6).gaussian_complex(48)
    cl, hop, sh = orc.stencil_coeffs("stencil", links, 0.3, 8, 6)
    st = orc.CStencil(8, 6, 1, cl, hop, None, complex(sh[0], sh[1]), complex(sh[2], sh[3]), complex(sh[4], sh[5]))
    assert same_bits(orc.cstencil_apply(st, v), orc.apply("stencil", v, links, 0.3, 8, 6))
    assert same_bits(orc.cstencil_apply(st, v, "eo"), orc.apply("stencil_eo", v, links, 0.3, 8, 6))
    assert same_bits(orc.cstencil_m2mdeodoe(st, v), orc.apply("stencil_m2mdeodoe", v, links, 0.3, 8, 6))
