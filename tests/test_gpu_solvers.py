"""GPU parity of the Krylov solvers against the oracle.

CG, GCR and CG-M: identical iteration count, ops_count, success flag and
name; solution within XTOL of the oracle's (1e-13 for the Laplacians, 1e-11
for the staggered normal operator, whose condition number ~1e3 amplifies the
different rounding of the GPU's tree reductions and FMA contraction over the
iterations; measured differences are 3e-16 .. 1e-12).

BiCGStab is not trajectory-stable under rounding (its iteration count moves
by a few per cent between any two FP64 implementations), so its bar is: the
same success flag, an iteration count within 10 % of the oracle's, and a
solution whose distance to the exact solution is within 3x of the oracle's
own distance.

The squared true residual resSq is a 1e-20-level quantity formed from the
solution; it is compared through the solution and by being below eps^2 |b|^2
whenever the solve reports success.  Verbosity lines must agree in every
field except the last digits of residuals.
"""
import numpy as np
import pytest

from helpers import (bicgstab_envelope, check_in_envelope, lattice_inputs, oracle_text_envelope, perturbed, pmap,
                     relerr, text_envelope)

pytestmark = pytest.mark.gpu

def xtol(op):
    return 1e-13 if op.startswith("laplace") else 1e-11


def check_info(got, want, name=None, bnorm2=None, eps=1e-10):
    assert got.iter == want.iter, (got, want)
    assert got.ops_count == want.ops_count, (got, want)
    assert got.success == want.success, (got, want)
    if name:
        assert got.name == name
    if got.success and bnorm2 is not None:
        assert got.resSq <= (eps * eps) * bnorm2 * 1.0001


# The verbosity text's RelRes values may differ from the oracle's by this
# multiple of the oracle's own spread on the same line (text_envelope: 1-ulp
# perturbations of b and the long double reductions), or by the printed
# precision (%g: six significant digits).
TEXT_ENVELOPE = 3.0
TEXT_QUANT = 1e-5


def check_text(got, want, env, floor=1e-13):
    """Verbosity lines agree in every field but RelRes exactly; the RelRes
    values of line i agree to max(TEXT_QUANT, TEXT_ENVELOPE * env[i]), except
    at the FP64 rounding floor, below `floor`, where only their smallness is
    compared.  Returns the largest deviation seen as a multiple of its
    line's allowance."""
    gl, wl = got.splitlines(), want.splitlines()
    assert len(gl) == len(wl) == len(env), (len(gl), len(wl), len(env), gl[-3:], wl[-3:])
    worst = 0.0
    for a, b, e in zip(gl, wl, env):
        ha, hb = a.split("RelRes")[0], b.split("RelRes")[0]
        assert ha == hb, (a, b)
        va = [float(t) for t in a.split("RelRes")[1].split()]
        vb = [float(t) for t in b.split("RelRes")[1].split()]
        assert len(va) == len(vb)
        tol = max(TEXT_QUANT, TEXT_ENVELOPE * e)
        for x, y in zip(va, vb):
            if max(abs(x), abs(y)) < floor:
                continue
            dev = abs(x - y) / abs(y)
            worst = max(worst, dev / tol)
            assert dev <= tol, (a, b, dev, e)
    return worst


def oracle_sensitivity(orc, solver, op, b, kw, trials=3, seed=0):
    """How far the oracle's own result moves when b is perturbed by 1e-16
    (relative, per entry): the rounding-level sensitivity of this solve."""
    x0, _ = orc.solve(solver, op, b, None, **kw)
    rng = np.random.default_rng(seed)
    worst = 0.0
    for _ in range(trials):
        if np.iscomplexobj(b):
            bp = b * (1 + 1e-16 * rng.standard_normal(b.size))
        else:
            bp = b * (1 + 1e-16 * rng.standard_normal(b.size))
        x1, _ = orc.solve(solver, op, bp, None, **kw)
        worst = max(worst, relerr(x1, x0))
    return worst


CASES = [
    # solver, op, lx, ly, mass, beta, with_x0
    ("cg", "laplace_real", 64, 64, 0.01, 6.25, False),
    ("cg", "staggered_normal", 32, 32, 0.05, 6.25, True),
    ("cg", "staggered_normal", 24, 10, 0.1, 2.0, False),
    ("cg", "laplace_u1", 16, 16, 0.05, 3.0, True),
    ("gcr", "staggered", 16, 16, 0.1, 6.25, False),
    ("gcr", "laplace_real", 32, 32, 0.01, 6.25, True),
    ("gcr", "stencil", 16, 16, 0.1, 6.25, False),
    ("cg", "laplace", 20, 12, 0.2, 6.25, True),
]

BICG_CASES = [
    ("staggered", 32, 32, 0.05, 6.25, True),
    ("laplace_real", 32, 32, 0.01, 6.25, False),
    ("stencil", 16, 16, 0.1, 6.25, False),
    ("staggered_gamma5", 16, 16, 0.2, 6.25, False),
]


@pytest.mark.parametrize("solver,op,lx,ly,mass,beta,with_x0", CASES)
def test_solver_matches_oracle(glx, orc, solver, op, lx, ly, mass, beta, with_x0):
    links, bc, br = lattice_inputs(orc, lx, ly, beta=beta, seed=100 + lx)
    b = br if op == "laplace_real" else bc
    x0 = 0.1 * b[::-1].copy() if with_x0 else None
    kw = dict(links=links, mass=mass, lx=lx, ly=ly, max_iter=2000, eps=1e-10, verb_level=3, prefix="[T]: ")
    xo, io = orc.solve(solver, op, b, x0, **kw)
    xg, ig = glx.solve(solver, op, b, x0, **kw)
    check_info(ig, io, io.name, bnorm2=float(np.vdot(b, b).real))
    assert relerr(xg, xo) < xtol(op), relerr(xg, xo)
    base, env = oracle_text_envelope(orc, solver, op, b, x0, kw)
    assert base == io.text
    print(f"{solver} {op}: text within {check_text(ig.text, io.text, env):.2f} of its allowance")


@pytest.mark.parametrize("op,lx,ly,mass,beta,with_x0", BICG_CASES)
def test_bicgstab_matches_oracle(glx, orc, op, lx, ly, mass, beta, with_x0):
    links, bc, br = lattice_inputs(orc, lx, ly, beta=beta, seed=100 + lx)
    b = br if op == "laplace_real" else bc
    x0 = 0.1 * b[::-1].copy() if with_x0 else None
    kw = dict(links=links, mass=mass, lx=lx, ly=ly, max_iter=4000, eps=1e-10)
    xo, io = orc.solve("bicgstab", op, b, x0, **kw)
    xg, ig = glx.solve("bicgstab", op, b, x0, **kw)
    # reference solution far below the solver tolerance
    xs, _ = orc.solve("gcr", op, b, None, **dict(kw, eps=1e-14))
    assert ig.success == io.success and ig.name == "BiCGStab"
    assert abs(ig.iter - io.iter) <= max(2, 0.1 * io.iter), (ig.iter, io.iter)
    assert ig.ops_count == 2 * ig.iter + 2
    assert relerr(xg, xs) <= 3 * max(relerr(xo, xs), 1e-12), (relerr(xg, xs), relerr(xo, xs))
    if ig.success:
        assert ig.resSq <= 1e-20 * float(np.vdot(b, b).real) * 1.0001


def test_eoprec_solve_matches_direct(glx, orc):
    """prepare -> CG on m^2 - D_eo D_oe -> reconstruct equals the direct
    solution of D x = b (staggered_pieces.cpp:330-346), for both the function
    and the stencil form."""
    lx = ly = 32
    m = 0.1
    links, bc, _ = lattice_inputs(orc, lx, ly, beta=6.25, seed=8)
    xs, _ = orc.solve("gcr", "staggered", bc, None, links=links, mass=m, lx=lx, ly=ly, max_iter=4000, eps=1e-14)
    for stencil in (False, True):
        be = glx.eoprec_prepare(bc, links, m, lx, ly, stencil)
        op = "stencil_m2mdeodoe" if stencil else "m2mdeodoe"
        kw = dict(links=links, mass=m, lx=lx, ly=ly, max_iter=2000, eps=1e-12)
        xe_o, io = orc.solve("cg", op, be, None, **kw)
        xe, ig = glx.solve("cg", op, be, None, **kw)
        check_info(ig, io, "CG")
        assert relerr(xe, xe_o) < 1e-11
        x = glx.eoprec_reconstruct(xe, bc, links, m, lx, ly, stencil)
        assert relerr(x, xs) < 1e-9


@pytest.mark.parametrize("solver", ["cg", "bicgstab", "gcr"])
def test_restarted_matches_oracle(glx, orc, solver):
    """CG(9) and GCR(9): the oracle's iterations, ops and text (to its
    envelope); BiCGStab(9), whose trajectory is not stable under rounding:
    iterations and solution inside the oracle's own spread under 1-ulp
    perturbations of b."""
    lx = ly = 16
    op = "staggered_normal" if solver == "cg" else "staggered"
    links, bc, _ = lattice_inputs(orc, lx, ly, beta=6.25, seed=5)
    kw = dict(links=links, mass=0.05, lx=lx, ly=ly, max_iter=600, eps=1e-10, restart_freq=9, verb_level=3,
              prefix="[R]: ")
    if solver == "bicgstab":
        # at m = 0.05 the oracle's own BiCGStab(9) takes 503 .. 764 iterations
        # over 1-ulp perturbations of b; m = 0.2 keeps the spread (143 .. 186)
        # well inside max_iter
        kw0 = dict(links=links, mass=0.2, lx=lx, ly=ly, max_iter=2000, eps=1e-10)
        xg, ig = glx.solve(solver, op, bc, None, **dict(kw0, restart_freq=9))
        x0, i0, its, dx = bicgstab_envelope(orc, op, bc, kw0, restart_freq=9)
        assert ig.name == i0.name == "BiCGStab(9)" and ig.success and i0.success
        check_in_envelope(xg, ig, x0, its, dx, "BiCGStab(9) 16^2")
        r = bc - orc.apply(op, xg, links, 0.2, lx, ly)
        assert np.linalg.norm(r) <= 1e-10 * np.linalg.norm(bc) * 1.0001
        return
    xg, ig = glx.solve(solver, op, bc, None, **kw)
    xo, io = orc.solve(solver, op, bc, None, **kw)
    check_info(ig, io, io.name)
    assert relerr(xg, xo) < 1e-9
    _, env = oracle_text_envelope(orc, solver, op, bc, None, kw)
    print(f"{solver}(9): text within {check_text(ig.text, io.text, env):.2f} of its allowance")


@pytest.mark.parametrize("verb", [0, 1, 2, 3])
def test_verbosity_levels(glx, orc, verb):
    lx = ly = 16
    links, bc, _ = lattice_inputs(orc, lx, ly, seed=2)
    kw = dict(links=links, mass=0.1, lx=lx, ly=ly, max_iter=500, eps=1e-10, restart_freq=7, verb_level=verb,
              prefix="[V]: ", fmt="%.6e")
    _, io = orc.solve("cg", "staggered_normal", bc, None, **kw)
    _, ig = glx.solve("cg", "staggered_normal", bc, None, **kw)
    _, env = oracle_text_envelope(orc, "cg", "staggered_normal", bc, None, kw)
    check_text(ig.text, io.text, env)


def test_max_iter_quirks(glx, orc):
    """success = false when stopping at k == max_iter-1 (generic_cg.cpp:208-214)."""
    lx = ly = 16
    links, bc, _ = lattice_inputs(orc, lx, ly, seed=3)
    for mi in (1, 2, 5):
        kw = dict(links=links, mass=0.1, lx=lx, ly=ly, max_iter=mi, eps=1e-10)
        for solver, op in (("cg", "staggered_normal"), ("bicgstab", "staggered"), ("gcr", "staggered")):
            xo, io = orc.solve(solver, op, bc, None, **kw)
            xg, ig = glx.solve(solver, op, bc, None, **kw)
            check_info(ig, io)
            # CG and GCR are rounding-stable over a few iterations; this
            # BiCGStab start is near breakdown (|x| jumps to ~8 after one
            # step) and the oracle itself moves by ~5e-7 after 5 iterations
            # when b moves by 1e-16, so it is held to that sensitivity.
            bound = 1e-13
            if solver == "bicgstab":
                bound = max(bound, 10 * oracle_sensitivity(orc, solver, op, bc, kw))
            assert relerr(xg, xo) < bound, (solver, mi, relerr(xg, xo), bound)


def test_bicgstab_restart(glx, orc):
    """BiCGStab(n): the restart driver's bookkeeping (name, ops = sum of the
    inner solves' ops, success re-evaluated from the true residual)."""
    lx = ly = 16
    links, bc, _ = lattice_inputs(orc, lx, ly, beta=6.25, seed=5)
    kw = dict(links=links, mass=0.2, lx=lx, ly=ly, max_iter=2000, eps=1e-10, restart_freq=25, verb_level=2,
              prefix="[B]: ")
    xg, ig = glx.solve("bicgstab", "staggered", bc, None, **kw)
    xs, _ = orc.solve("gcr", "staggered", bc, None, **dict(kw, eps=1e-14, restart_freq=0, verb_level=0))
    assert ig.name == "BiCGStab(25)"
    lines = ig.text.splitlines()
    assert lines[-1].startswith("[B]: BiCGStab(25) Success")
    restarts = [l for l in lines[:-1] if l.startswith("[B]: BiCGStab(25) Iter")]
    assert len(restarts) >= 1
    last_it, last_ops = restarts[-1].split()[3], restarts[-1].split()[5]
    assert int(last_it) == ig.iter and int(last_ops) == ig.ops_count
    assert ig.success and relerr(xg, xs) < 1e-8


@pytest.mark.parametrize("op,mass,wf", [("laplace_real", 0.001, True), ("laplace_real", 0.001, False),
                                         ("laplace_u1", 0.001, True), ("staggered_normal", 0.05, False)])
def test_multishift_matches_oracle(glx, orc, op, mass, wf):
    lx = ly = 16
    links, bc, br = lattice_inputs(orc, lx, ly, beta=6.25, seed=17)
    b = br if op == "laplace_real" else bc
    sh = np.array([0.001, 0.01, 0.005, 0.05, 0.1, 0.5, 1.0])
    sh = sh - sh[0]
    kw = dict(links=links, mass=mass, lx=lx, ly=ly, resid_freq_check=10, max_iter=2000, eps=1e-10,
              worst_first=wf, verb_level=3, prefix="[M]: ", fmt="%.6e")
    Xo, io, ro, so = orc.cg_m(op, b, sh, **kw)
    Xg, ig, sg = glx.minv_vector_cg_m(op, b, sh, **kw)
    check_info(ig, io, "CG-M")
    assert ig.resSq == 0.0
    assert np.array_equal(sg, sh) and np.array_equal(so, sh)
    bn2 = float(np.vdot(b, b).real)
    for k in range(sh.size):
        assert relerr(Xg[k], Xo[k]) < xtol(op), (k, relerr(Xg[k], Xo[k]))
        # per-shift squared residuals: both below tolerance or both above
        assert (ig.resSqmrhs[k] <= 1e-20 * bn2) == (ro[k] <= 1e-20 * bn2) or abs(ig.resSqmrhs[k] - ro[k]) < 0.1 * ro[k]
    jobs = list(range(3)) + ([-1] if np.iscomplexobj(b) else [])
    others = pmap(lambda k: orc.cg_m(op, b if k < 0 else perturbed(b, k), sh,
                                     **dict(kw, reductions="extended" if k < 0 else "serial"))[1].text, jobs)
    print(f"CG-M {op}: text within {check_text(ig.text, io.text, text_envelope(io.text, others)):.2f} of its "
          f"allowance")
