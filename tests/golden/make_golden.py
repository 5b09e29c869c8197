#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/ from the reference itself.

Runs only where the reference's sources are available: `make -C oracle`
compiles them (unchanged, from /root/reference) into oracle/_ref/libref.so,
and this script drives that build through oracle/orc.py (impl="ref") on small
seeded inputs.  The fixtures hold inputs and outputs only (numpy arrays and
JSON); tests/test_oracle_golden.py checks that the oracle restatement
reproduces every one of them bit for bit.

Usage: python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import orc  # noqa: E402

OPS = ["laplace", "laplace_u1", "staggered", "staggered_dagger", "staggered_normal", "staggered_gamma5", "gamma5",
       "deo", "doe", "m2mdeodoe", "stencil", "stencil_dagger", "stencil_gamma5", "stencil_eo", "stencil_oe",
       "stencil_m2mdeodoe"]

# operators.h functions off the solver path (wilson_coeff 0.7 for the two-link term)
EXTRA_OPS = ["staggered_free", "staggered_gamma5_free", "staggered_2linklaplace", "symmshift_x", "symmshift_y",
             "index_operator"]

# (piece, sdir) pairs of the general-stencil fixtures
CST_CASES = [("all", 0), ("all", 1), ("all", 3), ("all", 7), ("eo", 0), ("oe", 4), ("eo", 9), ("tb", 0), ("bt", 0),
             ("tb", 12), ("bt", 1)]


def main():
    if not orc.ref_available():
        orc.build()
    assert orc.ref_available(), "oracle/_ref/libref.so is needed (make -C oracle with the reference present)"
    arrays = {}
    meta = {"solves": {}, "cg_m": {}}

    # random streams of std::mt19937(1337): reals, complex, gauge, transform
    g = orc.Rng(1337, "ref")
    arrays["rng_gauss_real"] = g.gaussian_real(101)
    arrays["rng_gauss_complex"] = g.gaussian_complex(64)
    arrays["rng_gauss_gauge_8x6_b625"] = g.gauss_gauge(8, 6, 6.25)
    arrays["rng_rand_gauge_4x4"] = g.rand_gauge(4, 4)
    arrays["rng_rand_trans_8x6"] = g.rand_trans(8, 6)
    del g

    # operator applications on an 8x6 lattice (and Laplacians with colours)
    lx, ly, m = 8, 6, 0.37
    g = orc.Rng(7, "ref")
    links = g.gauss_gauge(lx, ly, 2.0)
    v = g.gaussian_complex(lx * ly)
    vr = g.gaussian_real(lx * ly * 2)
    vc2 = g.gaussian_complex(lx * ly * 2)
    del g
    arrays["op_links"] = links
    arrays["op_in"] = v
    arrays["op_in_real_nc2"] = vr
    arrays["op_in_cplx_nc2"] = vc2
    for op in OPS:
        arrays[f"op_out_{op}"] = orc.apply(op, v, links, m, lx, ly, impl="ref")
    for op in EXTRA_OPS:
        arrays[f"op_out_{op}"] = orc.apply(op, v, links, m, lx, ly, impl="ref", wilson=0.7)
    arrays["op_out_laplace_real_nc2"] = orc.apply("laplace_real", vr, None, m, lx, ly, 2, impl="ref")
    arrays["op_out_laplace_nc2"] = orc.apply("laplace", vc2, None, m, lx, ly, 2, impl="ref")
    for st in (False, True):
        arrays[f"eo_prepare_{int(st)}"] = orc.eoprec_prepare(v, links, m, lx, ly, st, impl="ref")
        arrays[f"eo_reconstruct_{int(st)}"] = orc.eoprec_reconstruct(v, v[::-1].copy(), links, m, lx, ly, st,
                                                                      impl="ref")
    for op in ("stencil", "stencil_dagger", "stencil_gamma5"):
        cl, hop, sh = orc.stencil_coeffs(op, links, m, lx, ly, impl="ref")
        arrays[f"coef_{op}_clover"], arrays[f"coef_{op}_hopping"], arrays[f"coef_{op}_shifts"] = cl, hop, sh
    arrays["plaquette"] = np.array([orc.plaquette(links, lx, ly, impl="ref")])
    t = orc.Rng(9, "ref").rand_trans(lx, ly)
    arrays["gtrans_t"] = t
    arrays["gtrans_out"] = orc.apply_gauge_trans(links, t, lx, ly, impl="ref")
    for w in ("dot", "norm2sq", "diffnorm2sq"):
        arrays[f"blas_{w}_c"] = np.array([orc.blas(w, v, v[::-1].copy(), impl="ref")])
        arrays[f"blas_{w}_r"] = np.array([orc.blas(w, vr, vr[::-1].copy(), impl="ref")])

    # general stencil_2d: nc = 3 with two-link terms on 5x4, coefficients and
    # vectors drawn from std::mt19937(13)
    g = orc.Rng(13, "ref")
    clx, cly, cnc = 5, 4, 3
    blk = clx * cly * cnc * cnc
    cst = orc.CStencil(clx, cly, cnc, g.gaussian_complex(blk), g.gaussian_complex(4 * blk),
                       g.gaussian_complex(8 * blk), 0.3 + 0.1j, -0.2j, 0.7)
    cv = g.gaussian_complex(clx * cly * cnc)
    cw = g.gaussian_complex(clx * cly * cnc)
    del g
    arrays["cst_clover"], arrays["cst_hopping"], arrays["cst_two"] = cst.clover, cst.hopping, cst.two_link
    arrays["cst_shifts"], arrays["cst_in"], arrays["cst_in2"] = cst.shifts6(), cv, cw
    for piece, sdir in CST_CASES:
        cst.sdir = sdir
        arrays[f"cst_out_{piece}_{sdir}"] = orc.cstencil_apply(cst, cv, piece, impl="ref", lhs=cw)
    cst.sdir = 0
    arrays["cst_m2mdeodoe"] = orc.cstencil_m2mdeodoe(cst, cv, impl="ref")
    arrays["cst_prepare"] = orc.cstencil_eoprec_prepare(cst, cv, impl="ref")
    arrays["cst_reconstruct"] = orc.cstencil_eoprec_reconstruct(cst, cv, cw, impl="ref")
    gen = orc.cstencil_generate(6, 5, 1, True, op="staggered_normal", links=orc.Rng(5, "ref").gauss_gauge(6, 5, 2.0),
                                mass=0.2, impl="ref")
    arrays["cst_gen_links"] = orc.Rng(5, "ref").gauss_gauge(6, 5, 2.0)
    arrays["cst_gen_clover"], arrays["cst_gen_hopping"], arrays["cst_gen_two"] = gen.clover, gen.hopping, gen.two_link

    # solves (16x16 unless stated), with DETAIL verbosity text
    g = orc.Rng(11, "ref")
    L = 16
    sl = g.gauss_gauge(L, L, 6.25)
    sb = g.gaussian_complex(L * L)
    sbr = g.gaussian_real(L * L)
    sx0 = 0.1 * g.gaussian_complex(L * L)
    del g
    arrays["solve_links"], arrays["solve_b"], arrays["solve_br"], arrays["solve_x0"] = sl, sb, sbr, sx0
    cases = [
        ("cg", "laplace_real", 0.01, 0, False), ("cg", "staggered_normal", 0.05, 0, True),
        ("bicgstab", "staggered", 0.05, 0, True), ("bicgstab", "laplace_real", 0.01, 0, False),
        ("gcr", "staggered", 0.05, 0, False), ("gcr", "laplace_real", 0.01, 0, True),
        ("cg", "staggered_normal", 0.05, 7, False), ("bicgstab", "staggered", 0.05, 5, False),
        ("gcr", "staggered", 0.05, 9, False), ("cg", "stencil_m2mdeodoe", 0.05, 0, False),
        ("cg", "staggered_normal", 0.1, 3, False),
    ]
    for solver, op, mass, rf, x0 in cases:
        b = sbr if op == "laplace_real" else sb
        if op == "stencil_m2mdeodoe":
            b = orc.eoprec_prepare(sb, sl, mass, L, L, True, impl="ref")
        guess = None
        if x0:
            guess = sx0.real.copy() if op == "laplace_real" else sx0
        max_iter = 3 if rf == 3 else 500
        x, info = orc.solve(solver, op, b, guess, links=sl, mass=mass, lx=L, ly=L, max_iter=max_iter, eps=1e-10,
                            restart_freq=rf, verb_level=3, prefix="[G]: ", fmt="%.6e", impl="ref")
        key = f"{solver}_{op}_m{mass}_rf{rf}_x0{int(x0)}"
        arrays[f"solve_x_{key}"] = x
        if op == "stencil_m2mdeodoe":
            arrays[f"solve_rhs_{key}"] = b
        meta["solves"][key] = dict(solver=solver, op=op, mass=mass, restart_freq=rf, x0=x0, max_iter=max_iter,
                                   resSq=info.resSq, iter=info.iter, success=info.success,
                                   ops_count=info.ops_count, name=info.name, text=info.text)
    sh = np.array([0.001, 0.01, 0.005, 0.05, 0.1, 0.5, 1.0])
    sh = sh - sh[0]
    for op, mass, wf in (("laplace_real", 0.001, True), ("laplace_u1", 0.001, False),
                         ("staggered_normal", 0.05, True)):
        b = sbr if op == "laplace_real" else sb
        X, info, res, sh_after = orc.cg_m(op, b, sh, links=sl, mass=mass, lx=L, ly=L, resid_freq_check=10,
                                          max_iter=2000, eps=1e-10, worst_first=wf, verb_level=3, prefix="[M]: ",
                                          fmt="%.6e", impl="ref")
        key = f"cgm_{op}_wf{int(wf)}"
        arrays[f"{key}_X"] = X
        arrays[f"{key}_res"] = res
        meta["cg_m"][key] = dict(op=op, mass=mass, worst_first=wf, shifts=sh.tolist(),
                                 shifts_after=sh_after.tolist(), iter=info.iter, success=info.success,
                                 ops_count=info.ops_count, resSq=info.resSq, text=info.text)

    np.savez_compressed(os.path.join(HERE, "golden.npz"), **arrays)
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    print("wrote", len(arrays), "arrays and", len(meta["solves"]) + len(meta["cg_m"]), "solve records")


if __name__ == "__main__":
    main()
