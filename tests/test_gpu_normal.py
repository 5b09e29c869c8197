"""Row-marching D^dag D kernel (operators.cpp:444-453 is the operator).

Every launch shape (rows per block, including ragged last strips, the
narrowest lattice the kernel takes and strips whose halo rows wrap onto
their own rows) against the oracle in FP64 and against the tiled kernel in
FP32; solver parity with the kernel in the loop; the autotuner's choice and
its JSON cache.
"""
import json

import numpy as np
import pytest

from helpers import ENVELOPE, cg_envelope, lattice_inputs, relerr

pytestmark = pytest.mark.gpu

SHAPES = [(64, 8), (64, 12), (128, 72), (64, 40), (192, 100), (256, 64)]
ROWS = [8, 12, 16, 32, 64]


@pytest.mark.parametrize("lx,ly", SHAPES)
@pytest.mark.parametrize("rows", ROWS)
def test_normal_fp64_matches_oracle(glx, orc, rows_reset, lx, ly, rows):
    links, bc, _ = lattice_inputs(orc, lx, ly, beta=2.0, seed=3 + lx + ly)
    mass = 0.37
    glx.set_normal_rows(lx, ly, 16, rows)
    assert glx.get_normal_rows(lx, ly, 16) == rows
    want = orc.apply("staggered_normal", bc, links, mass, lx, ly)
    got = glx.apply("staggered_normal", bc, links, mass, lx, ly)
    assert relerr(got, want) < 1e-14, relerr(got, want)


@pytest.mark.parametrize("lx,ly", SHAPES)
@pytest.mark.parametrize("rows", [8, 16, 64])
@pytest.mark.parametrize("link_prec", [1, 2])
def test_normal_fp32_matches_tiled(glx, orc, rows_reset, lx, ly, rows, link_prec):
    links, bc, _ = lattice_inputs(orc, lx, ly, beta=2.0, seed=5 + lx)
    lat = glx.Lattice(links, lx, ly)
    b32 = bc.astype(np.complex64)
    din = glx.DeviceArray(b32)
    dout = glx.DeviceArray(np.zeros_like(b32))
    glx.set_normal_rows(lx, ly, 8, 0)
    glx.dev_apply(lat, "staggered_normal", 0.2, dout, din, fp32=True, link_prec=link_prec)
    tiled = dout.download()
    glx.set_normal_rows(lx, ly, 8, rows)
    glx.dev_apply(lat, "staggered_normal", 0.2, dout, din, fp32=True, link_prec=link_prec)
    march = dout.download()
    # same arithmetic up to FMA contraction choices
    assert relerr(march, tiled) < 1e-6
    want = orc.apply("staggered_normal", b32.astype(np.complex128), links, 0.2, lx, ly)
    assert relerr(march, want) < (1e-6 if link_prec == 1 else 2e-3)


@pytest.mark.parametrize("rows", [0, 8, 16])
def test_cg_with_normal_kernel(glx, orc, rows_reset, rows):
    """The reference-order CG keeps exact iteration and ops counts with the
    row-marching kernel in the loop (its reductions are FP64 and the stencil
    agrees with the oracle to rounding)."""
    lx, ly, mass = 128, 72, 0.1
    links, bc, _ = lattice_inputs(orc, lx, ly, beta=4.0, seed=77)
    glx.set_normal_rows(lx, ly, 16, rows)
    lat = glx.Lattice(links, lx, ly)
    db = glx.DeviceArray(bc)
    dx = glx.DeviceArray(np.zeros_like(bc))
    info = glx.dev_solve("cg", lat, "staggered_normal", mass, dx, db, max_iter=5000, eps=1e-10)
    xo, io = orc.solve("cg", "staggered_normal", bc, None, links=links, mass=mass, lx=lx, ly=ly, max_iter=5000,
                       eps=1e-10)
    assert (info.iter, info.ops_count, info.success) == (io.iter, io.ops_count, io.success)
    assert relerr(dx.download(), xo) < 1e-11


@pytest.mark.parametrize("rows", [8, 16, 32])
def test_mixed_and_fused_with_normal_kernel(glx, orc, rows_reset, rows):
    lx = ly = 128
    mass = 0.05
    links, bc, _ = lattice_inputs(orc, lx, ly, beta=4.0, seed=78)
    glx.set_normal_rows(lx, ly, 16, rows)
    glx.set_normal_rows(lx, ly, 8, rows)
    lat = glx.Lattice(links, lx, ly)
    db = glx.DeviceArray(bc)
    xs, err_o, xo, io = cg_envelope(orc, "staggered_normal", bc, links, mass, lx, ly)
    for fn in (glx.dev_cg_fused, glx.dev_cg_mixed):
        dx = glx.DeviceArray(np.zeros_like(bc))
        info = fn(lat, "staggered_normal", mass, dx, db, max_iter=20000, eps=1e-10)
        x = dx.download()
        r = bc - orc.apply("staggered_normal", x, links, mass, lx, ly)
        assert info.success and np.linalg.norm(r) / np.linalg.norm(bc) <= 1.05e-10
        assert relerr(x, xs) <= ENVELOPE * err_o, (relerr(x, xs), err_o)
        assert info.iter <= 1.5 * io.iter + 20


def test_autotune_and_cache(glx, orc, rows_reset, tmp_path):
    lx = ly = 128
    links, _, _ = lattice_inputs(orc, lx, ly, beta=4.0, seed=79)
    lat = glx.Lattice(links, lx, ly)
    cache = str(tmp_path / "tune.json")
    res = glx.autotune_normal(lat, reps=5, cache=cache)
    assert not res["cached"]
    assert res["rows64"] in (0, 8, 12, 16, 24, 32, 48, 64, 96, 128) and res["ms64"] > 0
    assert res["rows32"] in (0, 8, 12, 16, 24, 32, 48, 64, 96, 128) and res["ms32"] > 0
    assert glx.get_normal_rows(lx, ly, 16) == res["rows64"]
    assert glx.get_normal_rows(lx, ly, 8) == res["rows32"]
    with open(cache) as f:
        table = json.load(f)
    assert table["128x128"]["rows64"] == res["rows64"]
    glx.set_normal_rows(lx, ly, 16, 8)
    again = glx.autotune_normal(lat, cache=cache)
    assert again["cached"] and glx.get_normal_rows(lx, ly, 16) == res["rows64"]
    rows_reset.add((lx, ly, 8))  # the autotuner's own entry, removed with the rest
