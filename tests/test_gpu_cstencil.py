"""General stencil_2d on the GPU (any Nc, two-link terms, sdir, pieces all /
eo / oe / tb / bt, the even/odd helpers, the generator and solves) against
the oracle (oracle/orc_ops.c, bit-exact against the reference build).

Tolerance: the device sums the same products in the same order but may fuse
multiply-adds, so entries agree to rounding: |device - oracle| <= 1e-14 per
accumulated product (13 Nc products at most, inputs O(1))."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rand_stencil(orc, glx, rng, lx, ly, nc, has_two, shifts=(0.3 + 0.1j, -0.2j, 0.7), sdir=0):
    o = orc.CStencil.random(rng, lx, ly, nc, has_two, shifts=shifts, sdir=sdir)
    g = glx.StencilNc(lx, ly, nc, o.clover, o.hopping, o.two_link, o.shift, o.eo_shift, o.dof_shift, sdir)
    return o, g


def tol(nc):
    return 1e-14 * 13 * nc * 10


CASES = [(nc, two, sdir) for nc in (1, 2, 3, 8) for two in (False, True) for sdir in (0, 1, 3, 7, 12)]


@pytest.mark.parametrize("nc,has_two,sdir", CASES)
def test_callback_pieces(glx, orc, nc, has_two, sdir):
    print(f"case nc={nc} has_two={has_two} sdir={sdir}", flush=True)
    rng = np.random.default_rng(100 * nc + 10 * has_two + sdir)
    o, g = rand_stencil(orc, glx, rng, 6, 5, nc, has_two, sdir=sdir)
    v = rng.standard_normal(o.n) + 1j * rng.standard_normal(o.n)
    init = rng.standard_normal(o.n) + 1j * rng.standard_normal(o.n)
    for piece in ("all", "eo", "oe", "tb", "bt"):
        want = orc.cstencil_apply(o, v, piece, lhs=init)
        got = glx.cstencil_apply(g, v, piece, lhs=init)
        assert np.max(np.abs(got - want)) <= tol(nc), piece


@pytest.mark.parametrize("nc,has_two", [(1, False), (2, True), (4, True), (16, False)])
def test_device_handle(glx, orc, nc, has_two):
    rng = np.random.default_rng(7 + nc)
    o, g = rand_stencil(orc, glx, rng, 32, 24, nc, has_two)
    v = rng.standard_normal(o.n) + 1j * rng.standard_normal(o.n)
    h = glx.CStencilDev(g)
    din, dout = glx.DeviceArray(v), glx.DeviceArray(np.zeros_like(v))
    for piece in ("all", "eo", "oe", "tb", "bt", "m2mdeodoe"):
        for sdir in ((0, 2, 9) if piece != "m2mdeodoe" else (0,)):
            o.sdir = sdir
            want = orc.cstencil_m2mdeodoe(o, v) if piece == "m2mdeodoe" else orc.cstencil_apply(o, v, piece)
            dout.upload(np.zeros_like(v))
            h.apply(dout, din, piece, sdir)
            if piece in ("all", "tb", "bt") and sdir >= 6 and not has_two:
                want = np.zeros_like(v)  # left as it was (zeros)
            scale = 13 * nc if piece == "m2mdeodoe" else 1
            assert np.max(np.abs(dout.download() - want)) <= tol(nc) * scale, (piece, sdir)
    h.close()


@pytest.mark.parametrize("nc", [1, 3])
def test_eo_helpers(glx, orc, nc):
    rng = np.random.default_rng(30 + nc)
    o, g = rand_stencil(orc, glx, rng, 8, 6, nc, False, shifts=(0.4, 0, 0))
    v = rng.standard_normal(o.n) + 1j * rng.standard_normal(o.n)
    w = rng.standard_normal(o.n) + 1j * rng.standard_normal(o.n)
    assert np.max(np.abs(glx.cstencil_apply(g, v, "m2mdeodoe") - orc.cstencil_m2mdeodoe(o, v))) <= tol(nc) * 13 * nc
    assert np.max(np.abs(glx.cstencil_eoprec_prepare(g, v) - orc.cstencil_eoprec_prepare(o, v))) <= tol(nc)
    assert np.max(np.abs(glx.cstencil_eoprec_reconstruct(g, v, w) - orc.cstencil_eoprec_reconstruct(o, v, w))) \
        <= tol(nc) / 0.4


def test_generator_from_device_operator(glx, orc):
    """generate_stencil_2d of the staggered normal operator (two-link terms),
    the operator running on the GPU, equals the oracle's generator."""
    lx, ly, m = 6, 6, 0.2
    links = orc.Rng(5).gauss_gauge(lx, ly, 2.0)
    want = orc.cstencil_generate(lx, ly, 1, True, op="staggered_normal", links=links, mass=m)
    op = glx.StaggeredU1Op(links.ctypes.data, m, lx, ly, 1, 0.0)
    import ctypes as C
    got = glx.cstencil_generate(lx, ly, 1, True, glx.op_callback("staggered_normal"),
                                C.cast(C.pointer(op), C.c_void_p))
    for a, b in ((got.clover, want.clover), (got.hopping, want.hopping), (got.two_link, want.two_link)):
        assert np.max(np.abs(a - b)) < 1e-14
    # and the generated stencil applies the operator
    v = orc.Rng(6).gaussian_complex(lx * ly)
    assert np.max(np.abs(glx.cstencil_apply(got, v) - orc.apply("staggered_normal", v, links, m, lx, ly))) < 1e-13


def test_solves(glx, orc):
    rng = np.random.default_rng(11)
    lx, ly, nc = 16, 12, 4
    o, g = rand_stencil(orc, glx, rng, lx, ly, nc, True, shifts=(0, 0, 0))
    # diagonally dominant: clover = 20 I + small, hopping and two-link small
    o.clover = (0.1 * o.clover.reshape(-1, nc, nc) + 20 * np.eye(nc)).reshape(-1)
    o.hopping *= 0.3
    o.two_link *= 0.1
    g = glx.StencilNc(lx, ly, nc, o.clover, o.hopping, o.two_link)
    b = rng.standard_normal(o.n) + 1j * rng.standard_normal(o.n)
    for solver in ("bicgstab", "gcr"):
        x, info = glx.cstencil_solve(solver, g, b, eps=1e-10)
        assert info.success, solver
        r = b - orc.cstencil_apply(o, x)
        assert np.linalg.norm(r) <= 1.01e-10 * np.linalg.norm(b), solver
    # device-resident CG on m^2 - D_eo D_oe of the staggered stencil (nc = 1)
    L, m = 32, 0.1
    links = orc.Rng(8).gauss_gauge(L, L, 4.0)
    cl, hop, sh = orc.stencil_coeffs("stencil", links, m, L, L)
    st1 = orc.CStencil(L, L, 1, cl, hop, None, complex(sh[0], sh[1]))
    h = glx.CStencilDev(glx.StencilNc(L, L, 1, cl, hop, None, complex(sh[0], sh[1])))
    be = orc.cstencil_eoprec_prepare(st1, orc.Rng(9).gaussian_complex(L * L))
    x, db = glx.DeviceArray(np.zeros_like(be)), glx.DeviceArray(be)
    info = h.solve("cg", x, db, "m2mdeodoe", max_iter=5000, eps=1e-10)
    assert info.success
    r = be - orc.cstencil_m2mdeodoe(st1, x.download())
    assert np.linalg.norm(r) <= 1.01e-10 * np.linalg.norm(be)
    h.close()
