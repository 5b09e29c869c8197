"""Shared input generation for the tests (seeded std::mt19937 streams drawn
by the oracle, bit-identical to the reference's own generators) and the
envelopes the GPU results are held to: what the oracle's own results do
under 1-ulp perturbations of the input and other valid reduction orders."""
import concurrent.futures as cf
import os

import numpy as np

WORKERS = min(8, os.cpu_count() or 4)


def pmap(fn, items):
    """map in threads (the oracle's C calls release the GIL)."""
    with cf.ThreadPoolExecutor(WORKERS) as ex:
        return list(ex.map(fn, items))


def lattice_inputs(orc, lx, ly, beta=6.25, seed=1337):
    """(links, complex rhs, real rhs) from one seeded stream."""
    g = orc.Rng(seed)
    links = g.gauss_gauge(lx, ly, beta)
    bc = g.gaussian_complex(lx * ly)
    br = g.gaussian_real(lx * ly)
    return links, bc, br


def relerr(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


# A solver that stops on the same tolerance as an FP64 CG solve is held to
# this multiple of that solve's own error.  Two solves that stop at the same
# residual norm can differ severalfold in error (the residual's content in
# the low modes decides it: the CPU restatement of the s-step algorithm stops
# at 8.8e-11 with 7.6 times the error of the oracle's CG at 8.2e-11); the GPU
# solvers follow the FP64 CG trajectory closely and measure 0.8 - 1.1 times
# it.  The a-priori bound of either error, cond(A) * tol, is 10^2 - 10^4
# times larger on these systems.
ENVELOPE = 4.0


def cg_envelope(orc, op, b, links, mass, lx, ly, eps=1e-10, max_iter=100000):
    """The accuracy an FP64 CG solve stopped at `eps` reaches on this system:
    (x_star, err, x_o, info_o) with x_star the oracle's CG solution at eps =
    1e-13 (the system's solution to ~1e-13), x_o and info_o the oracle's CG
    solve at `eps`, and err = |x_o - x_star| / |x_star|."""
    kw = dict(links=links, mass=mass, lx=lx, ly=ly, max_iter=max_iter)
    xs, _ = orc.solve("cg", op, b, None, eps=1e-13, **kw)
    xo, io = orc.solve("cg", op, b, None, eps=eps, **kw)
    return xs, relerr(xo, xs), xo, io


def eo_envelope(orc, b, links, mass, lx, ly, eps=1e-10, max_iter=100000):
    """cg_envelope of the even/odd solve of D x = b: the Schur system on the
    even sites by CG, then the odd sites reconstructed (the oracle's
    eoprec pipeline); returns (x_star, err)."""
    be = orc.eoprec_prepare(b, links, mass, lx, ly)
    kw = dict(links=links, mass=mass, lx=lx, ly=ly, max_iter=max_iter)
    xs = orc.eoprec_reconstruct(orc.solve("cg", "m2mdeodoe", be, None, eps=1e-13, **kw)[0], b, links, mass, lx, ly)
    xo = orc.eoprec_reconstruct(orc.solve("cg", "m2mdeodoe", be, None, eps=eps, **kw)[0], b, links, mass, lx, ly)
    return xs, relerr(xo, xs)


def perturbed(b, k, frac=0.01):
    """b with the last bit of the real part of a seeded 1 % of its entries
    flipped up (k odd) or down (k even)."""
    rng = np.random.default_rng(1000 + k)
    idx = rng.choice(b.size, size=max(1, int(b.size * frac)), replace=False)
    re = b.real.copy()
    re[idx] = np.nextafter(re[idx], np.inf if k % 2 else -np.inf)
    return re + 1j * b.imag if np.iscomplexobj(b) else re


def bicgstab_envelope(orc, op, b, kw, n=8, restart_freq=0):
    """The oracle's own spread: BiCGStab on b and on n 1-ulp perturbations
    of it.  Returns (x of b, info of b, iteration counts, relative distances
    of the perturbed solutions from the unperturbed one)."""
    kw = dict(kw, restart_freq=restart_freq)
    runs = pmap(lambda k: orc.solve("bicgstab", op, b if k < 0 else perturbed(b, k), None, **kw),
                list(range(-1, n)))
    x0, i0 = runs[0]
    its = np.array([i.iter for _, i in runs])
    dx = np.array([relerr(x, x0) for x, _ in runs[1:]])
    return x0, i0, its, dx


def check_in_envelope(xg, ig, x0, its, dx, what):
    """GPU iterations within the oracle runs' mean +- 3 sigma, its distance
    from the oracle's unperturbed x within twice the largest distance a
    1-ulp perturbation causes."""
    mu, sd = float(its.mean()), float(max(its.std(), 1.0))
    dist = relerr(xg, x0)
    print(f"{what}: GPU iter {ig.iter}, oracle iters {sorted(its.tolist())} (mean {mu:.1f}, sd {sd:.1f}); "
          f"|x_gpu - x_o|/|x_o| = {dist:.2e}, oracle spread {dx.min():.2e} .. {dx.max():.2e}")
    assert abs(ig.iter - mu) <= 3.0 * sd, (ig.iter, mu, sd)
    assert dist <= 2.0 * dx.max(), (dist, dx.max())


def relres_fields(text):
    """The numbers after "RelRes" on every verbosity line (empty where none)."""
    return [[float(t) for t in line.split("RelRes")[1].split()] if "RelRes" in line else []
            for line in text.splitlines()]


def text_envelope(base, others):
    """Per line of the verbosity text `base`: the largest relative change of
    its RelRes values over the texts `others` (the same solve on perturbed
    inputs or with other reductions), as a running maximum over the lines
    (the spread grows along a solve, and a line that some run did not reach
    takes the spread of the lines before it)."""
    vb = relres_fields(base)
    env = [0.0] * len(vb)
    for t in others:
        vo = relres_fields(t)
        for i in range(min(len(vb), len(vo))):
            if len(vo[i]) != len(vb[i]):
                continue
            for x, y in zip(vo[i], vb[i]):
                if y != 0.0:
                    env[i] = max(env[i], abs(x - y) / abs(y))
    run = 0.0
    for i, e in enumerate(env):
        run = max(run, e)
        env[i] = run
    return env


def oracle_text_envelope(orc, solver, op, b, x0, kw, n=3):
    """text_envelope of the oracle's verbosity text of this solve over n
    1-ulp perturbations of b and, for complex operators, its long double
    reductions."""
    jobs = [("p", k) for k in range(n)] + ([("e", 0)] if np.iscomplexobj(b) else [])

    def run(job):
        kind, k = job
        if kind == "e":
            return orc.solve(solver, op, b, x0, reductions="extended", **kw)[1].text
        return orc.solve(solver, op, perturbed(b, k), x0, **kw)[1].text

    base = orc.solve(solver, op, b, x0, **kw)[1].text
    return base, text_envelope(base, pmap(run, jobs))
