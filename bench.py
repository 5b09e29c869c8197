#!/usr/bin/env python3
"""Benchmark of the MI355X solver path on its headline workload.

Workload (one GPU): CG on the 2D U(1) staggered normal operator D^dag D,
L = 1024 x 1024, m = 0.02, link angles theta ~ N(0, 0.5^2), complex Gaussian
rhs, to true relative residual 1e-10 -- plain FP64 CG against mixed-precision
CG (FP32 Krylov vectors and links, FP64 reliable updates at delta = 0.1).
Inputs are synthetic (numpy, fixed seed) and resident in HBM before timing.
The launch shape of the D^dag D kernel is autotuned for the lattice first
(outside the timed region).

A "step" is one complete mixed-precision solve from x = 0.  `value` is its
time to solution in ms (lower is better).  The JSON line also carries the
FP64 solves' times (the reference-order CG and the single-pass CG), the
speed-ups, iteration counts, the solutions' relative errors against the FP64
GPU solve and against the reference CPU solve of the same system, the
roofline of the mixed solve's dominant kernel (the s-step launch) measured
live with HIP events, the plain D and D^dag D stencils at this size, 2x this
size and 4096^2 (beyond the 256 MB last-level cache) as fractions of the HBM
peak and of the streaming ceiling measured in the same run, the multishift
batch with its per-phase roofline, and the reference CPU path timed on the
host in the same run (the complete headline solve on one core; the batch on
all host cores).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--L 1024]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "generic-linalg_amd"))

HBM_PEAK_GBS = 8000.0


def make_inputs(L: int, sigma: float, seed: int):
    rng = np.random.default_rng(seed)
    theta = rng.normal(0.0, sigma, size=2 * L * L)
    links = np.exp(1j * theta)
    b = rng.normal(size=L * L) + 1j * rng.normal(size=L * L)
    return links.astype(np.complex128), b.astype(np.complex128)


def cpu_baseline(links, b, L, mass, gpu_x, grad=None, batch=None, small=None):
    """Reference CPU path (oracle/_ref, the reference's own sources built in
    this tree; the oracle restatement where that build is absent), timed on
    the host in the same run:
      * the headline solve itself, complete (FP64 CG to 1e-10 on one core,
        about 20 s): its time is `value`, and its solution gives the GPU
        solutions' relative errors against the FP64 CPU reference;
      * one core each: the 512^2 gradient (reference CG solve + D chi, the
        oracle's force) and the small workloads;
      * the whole multishift batch, its 192 independent CG-M solves spread
        over the host's cores (one solve per thread, longest first; the
        reference solvers are reentrant), each compared with its GPU
        solution."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc  # noqa: E402  (CPU baseline leg only)

    impl = "ref" if orc.ref_available() else "orc"
    t0 = time.perf_counter()
    xc, info = orc.solve("cg", "staggered_normal", b, None, links=links, mass=mass, lx=L, ly=L, max_iter=100000,
                         eps=1e-10, impl=impl)
    dt = time.perf_counter() - t0
    nx = float(np.linalg.norm(xc))
    out = {
        "value": round(dt * 1e3, 1),
        "unit": "ms (FP64 CG time to 1e-10)",
        "cores": 1,
        "kind": "reference" if impl == "ref" else "port",
        "sample": f"the complete {L}x{L} FP64 CG solve ({info.iter} iterations, success {info.success}) on one host "
                  f"core, {dt * 1e3 / (info.iter + 3):.1f} ms per iteration",
        "cpu_model": cpu_model(),
        "iters": info.iter,
        "relerr_vs_cpu": {k: float(np.linalg.norm(v - xc) / nx) for k, v in gpu_x.items()},
    }
    if grad is not None:
        Lg, mg = grad["L"], grad["mass"]
        rng = np.random.default_rng(21)
        theta = rng.normal(0.0, 0.4, 2 * Lg * Lg)
        phi = rng.normal(size=Lg * Lg) + 1j * rng.normal(size=Lg * Lg)
        lk = np.exp(1j * theta)
        t0 = time.perf_counter()
        chi, gi = orc.solve("cg", "staggered_normal", phi, None, links=lk, mass=mg, lx=Lg, ly=Lg, max_iter=100000,
                            eps=1e-10, impl=impl)
        psi = orc.apply("staggered", chi, lk, mg, Lg, Lg, impl=impl)
        orc.force(chi, psi, lk, mg, Lg, Lg)
        out["gradient_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
        out["gradient_sample"] = (f"full {Lg}x{Lg} gradient on one core: {impl} CG ({gi.iter} iterations) + D chi, "
                                  f"the oracle's force")
    if small is not None:
        gpu, inputs = small
        L1, b1 = inputs["C1"]
        L2, lk2, b2 = inputs["C2"]
        for name, solver, op, mass_, bb, lk, Ls in (("C1_laplace_cg", "cg", "laplace_real", 0.01, b1, None, L1),
                                                    ("C2_normal_cg", "cg", "staggered_normal", 0.05, b2, lk2, L2),
                                                    ("C2_bicgstab_D", "bicgstab", "staggered", 0.05, b2, lk2, L2)):
            t0 = time.perf_counter()
            xs, ci = orc.solve(solver, op, bb, None, links=lk, mass=mass_, lx=Ls, ly=Ls, max_iter=100000,
                               eps=1e-10, impl=impl)
            dtc = time.perf_counter() - t0
            xg = gpu[name]["x"]
            out[name] = {"ms": round(dtc * 1e3, 2), "iters": ci.iter,
                         "x_relerr_gpu_vs_cpu": float(np.linalg.norm(xg - xs) / np.linalg.norm(xs))}
    if batch is not None:
        out.update(cpu_batch(orc, impl, batch))
    return out


def host_threads():
    """Worker threads for the host leg: the host's cores, at most 16 (one
    GPU's share of the box)."""
    return max(1, min(16, os.cpu_count() or 1))


def cpu_batch(orc, impl, batch):
    """The multishift batch on the host: every system's CG-M (reference code)
    on its own thread, largest lattices first, all host threads busy; each
    result is compared with the GPU solution of the same system (per-shift
    relative error, iteration count)."""
    import concurrent.futures as cf

    systems, shifts = batch["host"], batch["shifts"]
    order = sorted(range(len(systems)), key=lambda k: -systems[k]["L"])
    nthr = host_threads()

    def one(k):
        s = systems[k]
        X, io, _, _ = orc.cg_m("staggered_normal", s["b"], shifts, links=s["links"], mass=s["mass"], lx=s["L"],
                               ly=s["L"], resid_freq_check=10, max_iter=100000, eps=1e-10, impl=impl)
        err = 0.0
        for q in range(len(shifts)):
            xg = s["x"][q].download()
            err = max(err, float(np.linalg.norm(xg - X[q]) / np.linalg.norm(X[q])))
        return k, io.iter, err

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(nthr) as ex:
        res = list(ex.map(one, order))
    dt = time.perf_counter() - t0
    same = sum(1 for k, it, _ in res if it == systems[k]["gpu_iter"])
    near = sum(1 for k, it, _ in res if abs(it - systems[k]["gpu_iter"]) <= 10)
    return {"batch_s": round(dt, 2), "batch_cores": nthr,
            "batch_sample": f"all {len(systems)} CG-M solves of the batch ({impl}), one per thread on {nthr} host "
                            f"threads, longest first",
            "batch_relerr_gpu_vs_cpu_max": max(e for _, _, e in res),
            "batch_iters_equal": same, "batch_iters_within_one_check": near}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip() + f" ({os.cpu_count()} logical CPUs visible)"
    except OSError:
        pass
    return None


def stencil_beyond_mall(glx, L=4096):
    """FP32 D^dag D (FP32 links) at 4096^2: 128 MiB per vector and 256 MiB of
    links, beyond the 256 MB last-level cache, so the figure is HBM-bound.
    The wave-march kernel at its own rows against the row-marching kernel
    at each candidate rows per block (the autotuner tunes the CG iteration
    kernel, which stages more data); the faster is reported."""
    V = L * L
    rng = np.random.default_rng(5)
    links = np.exp(1j * rng.normal(0.0, 0.5, size=2 * V)).astype(np.complex128)
    lat = glx.Lattice(links, L, L, f32=True, f16=False)
    x = (rng.normal(size=V) + 1j * rng.normal(size=V)).astype(np.complex64)
    d_in = glx.DeviceArray(x)
    d_out = glx.DeviceArray(np.zeros_like(x))
    best = None
    for rows in ("wave", 32, 64, 96, 128):
        glx.set_ddwave_rows(-1 if rows == "wave" else 0)
        if rows != "wave":
            glx.set_normal_rows(L, L, 8, rows)
        ms = glx.time_apply(lat, "staggered_normal", 0.02, d_out, d_in, reps=50, fp32=True,
                            link_prec=glx.LINKS_F32)
        if best is None or ms < best[1]:
            best = (rows, ms)
    rows, ms = best
    if rows != "wave":
        glx.set_normal_rows(L, L, 8, rows)
    gbs = 32 * V / (ms * 1e-3) / 1e9
    del d_in, d_out
    # the same operator in FP64 (64 B per site: 1 GiB per launch)
    x64 = x.astype(np.complex128)
    d_in, d_out = glx.DeviceArray(x64), glx.DeviceArray(np.zeros_like(x64))
    best64 = None
    for r64 in ("wave", 32, 64, 128):
        glx.set_ddwave_rows(-1 if r64 == "wave" else 0)
        if r64 != "wave":
            glx.set_normal_rows(L, L, 16, r64)
        m64 = glx.time_apply(lat, "staggered_normal", 0.02, d_out, d_in, reps=30)
        if best64 is None or m64 < best64[1]:
            best64 = (r64, m64)
    if best64[0] != "wave":
        glx.set_normal_rows(L, L, 16, best64[0])
    glx.set_ddwave_rows(-1)
    gbs64 = 64 * V / (best64[1] * 1e-3) / 1e9
    del lat, d_in, d_out
    glx.release_cache()
    return {"L": L, "rows": rows, "fp32_ms": ms, "fp32_GBs": round(gbs, 1),
            "frac": round(gbs / HBM_PEAK_GBS, 4), "fp64_rows": best64[0], "fp64_ms": best64[1],
            "fp64_GBs": round(gbs64, 1), "fp64_frac": round(gbs64 / HBM_PEAK_GBS, 4)}


def small_configs(glx, reps=5):
    """The reference's small workloads on the GPU (FP64, reference order of
    operations): C1 real Laplace 64^2 (diagonal 4 + 0.01) with CG; C2
    staggered 256^2 (m = 0.05, theta ~ N(0, 0.4^2)) with CG on D^dag D and
    BiCGStab on D.  Host-timed solves on device-resident vectors (best of
    `reps`); C1 is latency-bound, so no roofline fraction is given for it."""
    out = {}
    rng = np.random.default_rng(31)
    L1 = 64
    lat1 = glx.Lattice(np.ones(2 * L1 * L1, np.complex128), L1, L1, f32=False, f16=False)
    b1 = rng.standard_normal(L1 * L1)
    rng2 = np.random.default_rng(32)
    L2 = 256
    links2 = np.exp(1j * rng2.normal(0.0, 0.4, 2 * L2 * L2))
    b2 = rng2.standard_normal(L2 * L2) + 1j * rng2.standard_normal(L2 * L2)
    lat2 = glx.Lattice(links2, L2, L2, f32=False, f16=False)
    cases = (("C1_laplace_cg", "cg", lat1, "laplace_real", 0.01, b1),
             ("C2_normal_cg", "cg", lat2, "staggered_normal", 0.05, b2),
             ("C2_bicgstab_D", "bicgstab", lat2, "staggered", 0.05, b2))
    for name, solver, lat, op, mass, b in cases:
        bd = glx.DeviceArray(b)
        x = glx.DeviceArray(np.zeros_like(b))
        best, info = 1e9, None
        for _ in range(reps):
            x.upload(np.zeros_like(b))
            glx.synchronize()
            t0 = time.perf_counter()
            info = glx.dev_solve(solver, lat, op, mass, x, bd, max_iter=100000, eps=1e-10)
            glx.synchronize()
            best = min(best, time.perf_counter() - t0)
        out[name] = {"ms": round(best * 1e3, 3), "iters": info.iter, "success": info.success,
                     "L": lat.lx, "mass": mass, "x": x.download()}
    return out, {"C1": (L1, b1), "C2": (L2, links2, b2)}


def cstencil_traffic(L: int, nc: int):
    """HBM bytes per launch of k_cstencil for one bench shape, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary; None if absent."""
    path = os.path.join(ROOT, "profiles", "k_cstencil_pmc.txt")
    if not os.path.exists(path):
        return None
    hit = False
    with open(path) as f:
        for line in f:
            if line.startswith(f"L={L} nc={nc} "):
                hit = True
            elif hit and "traffic_bytes_per_launch:" in line:
                return float(line.split(":", 1)[1])
    return None


def general_stencil(glx):
    """General stencil_2d (coarse_stencil.cpp:12-392) on the device: the
    per-launch bytes are the coefficients (n_terms nc^2 16 B per site, read
    once) plus the input and output vectors (2 nc 16 B per site).  Two shapes
    of multigrid coarse operators: nc = 8 with two-link terms on 128^2
    (218 MB of coefficients, beyond the last-level cache) and nc = 16
    one-link on 64^2."""
    out = []
    rng = np.random.default_rng(9)
    for L, nc, two in ((128, 8, True), (64, 16, False)):
        V = L * L
        blk = V * nc * nc
        g = lambda k: (rng.standard_normal(k * blk) + 1j * rng.standard_normal(k * blk))
        st = glx.StencilNc(L, L, nc, g(1), g(4), g(8) if two else None, 0.1)
        h = glx.CStencilDev(st)
        del st
        v = (rng.standard_normal(V * nc) + 1j * rng.standard_normal(V * nc))
        d_in, d_out = glx.DeviceArray(v), glx.DeviceArray(np.zeros_like(v))
        ms = h.time(d_out, d_in, "all", reps=50)
        nterms = 13 if two else 5
        by = V * (nterms * nc * nc * 16 + 2 * nc * 16)
        gbs = by / (ms * 1e-3) / 1e9
        out.append({"L": L, "nc": nc, "two_link": two, "ms": ms, "bytes": by, "GBs": round(gbs, 1),
                    "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": cstencil_traffic(L, nc)})
        h.close()
        del d_in, d_out
    glx.release_cache()
    return out


def gauge_generation(glx, L=512, reps=10):
    """Gaussian U(1) field generated in HBM (Philox4x32-10, Box-Muller) with
    its FP32/FP16 copies, and the device mean plaquette, for 512^2."""
    lat = glx.Lattice.gaussian(L, L, 6.25, seed=1)
    glx.synchronize()
    t0 = time.perf_counter()
    for r in range(reps):
        lat.regenerate(6.25, seed=r + 2)
    glx.synchronize()
    gen_ms = (time.perf_counter() - t0) * 1e3 / reps
    t0 = time.perf_counter()
    for r in range(reps):
        p = lat.plaquette()
    plaq_ms = (time.perf_counter() - t0) * 1e3 / reps
    lat.close()
    return {"L": L, "beta": 6.25, "generate_ms": round(gen_ms, 4), "plaquette_ms": round(plaq_ms, 4),
            "plaquette": [p.real, p.imag], "note": "host-timed, each call synchronises"}


def gradient_timing(glx, L=512, mass=0.05, reps=3):
    """Gradient of S = phi^dag (D^dag D)^-1 phi on the 512^2 workload
    (theta ~ N(0, 0.4^2), phi complex Gaussian): adjoint solve to 1e-10 +
    D chi + per-link force, FP64 and FP32-bulk, gradient left in HBM."""
    rng = np.random.default_rng(21)
    theta = rng.normal(0.0, 0.4, 2 * L * L)
    phi = rng.normal(size=L * L) + 1j * rng.normal(size=L * L)
    lat = glx.Lattice(np.exp(1j * theta), L, L)
    glx.autotune_normal(lat, reps=20)
    dphi = glx.DeviceArray(phi)
    g = glx.DeviceArray(nbytes=2 * L * L * 8, dtype=np.float64, shape=(2 * L * L,))
    out = {"L": L, "mass": mass, "tol": 1e-10}
    ref = None
    for name, prec in (("fp64", 0), ("fp32_bulk", glx.LINKS_F32)):
        glx.staggered_grad(lat, mass, dphi, precision=prec, grad=g, download=False)  # warm-up
        ts = []
        for _ in range(reps):
            glx.synchronize()
            t0 = time.perf_counter()
            r = glx.staggered_grad(lat, mass, dphi, precision=prec, grad=g, download=False)
            glx.synchronize()
            ts.append(time.perf_counter() - t0)
        gv = g.download()
        if ref is None:
            ref = gv
        out[name] = {"ms": round(float(np.median(ts)) * 1e3, 3), "iters": r.info.iter, "dS_dm": r.dS_dm,
                     "relerr_vs_fp64": float(np.linalg.norm(gv - ref) / np.linalg.norm(ref))}
    del lat
    glx.release_cache()
    return out


def batch_multishift(glx, n_lat=48, rhs=4, seq=True, keep_host=True):
    """48 staggered lattices, L uniform on {64, ..., 512}, m ~ U[0.01, 0.1],
    theta ~ N(0, 0.4^2), 4 complex Gaussian right-hand sides each, 12 shifts
    sigma_k = 10^(-3 + 0.5 k), tol 1e-10, resid_freq_check 10: the 192
    multishift solves in one batched call (D^dag D in strips of 32 rows: the
    batch's own choice),
    with the algorithmic bytes of each phase of an iteration counted over the
    shifts active in it; then the same call again with every phase's launches
    between HIP events (no graph), for the batch's roofline; and the same
    solves one call each for comparison.  Returns (figures, host copies of
    the inputs with the GPU solutions for the CPU leg)."""
    rng = np.random.default_rng(4)
    Ls = rng.choice([64, 128, 192, 256, 384, 512], size=n_lat)
    masses = rng.uniform(0.01, 0.1, size=n_lat)
    shifts = 10.0 ** (-3.0 + 0.5 * np.arange(12))
    systems, host = [], []
    for i, L in enumerate(Ls):
        L = int(L)
        links = np.exp(1j * rng.normal(0.0, 0.4, 2 * L * L))
        lat = glx.Lattice(links, L, L, f32=False, f16=False)
        for _ in range(rhs):
            b = rng.normal(size=L * L) + 1j * rng.normal(size=L * L)
            xs = [glx.DeviceArray(nbytes=L * L * 16, shape=(L * L,)) for _ in shifts]
            systems.append((lat, float(masses[i]), glx.DeviceArray(b), xs))
            if keep_host:
                host.append({"L": L, "mass": float(masses[i]), "links": links, "b": b, "x": xs})
    glx.dev_cg_m_batched(systems[:8], shifts, max_iter=20)  # warm-up (code paths, allocations)
    glx.synchronize()
    t0 = time.perf_counter()
    infos, st = glx.dev_cg_m_batched(systems, shifts, resid_freq_check=10, max_iter=100000, eps=1e-10, stats=True)
    glx.synchronize()
    t_batch = time.perf_counter() - t0
    site_iters = float(sum(lat.lx * lat.ly * inf.iter for (lat, _, _, _), inf in zip(systems, infos)))
    out = {"systems": len(systems), "lattices": n_lat, "rhs": rhs, "shifts": 12,
           "sites_total": int(sum(int(L) * int(L) for L in Ls) * rhs),
           "batch_s": round(t_batch, 3), "all_success": all(i.success for i in infos),
           "iters_max": max(i.iter for i in infos), "iters_sum": sum(i.iter for i in infos),
           "site_iters": site_iters,
           "rows_per_block": 32}
    for h, inf in zip(host, infos):
        h["gpu_iter"] = inf.iter
    # the same solve with each phase between HIP events: per-phase roofline
    _, tp = glx.dev_cg_m_batched(systems, shifts, resid_freq_check=10, max_iter=100000, eps=1e-10,
                                 time_phases=True)
    total = st["bytes_stencil"] + st["bytes_r_pass"] + st["bytes_shift_pass"]
    # two launches per iteration: D^dag D (forming p in its prologue) and the
    # residual pass, which also applies each window's x_s, p_s updates
    phases = {}
    for name, by, ms in (("stencil", st["bytes_stencil"], tp["ms_stencil"]),
                         ("r_pass", st["bytes_r_pass"] + st["bytes_shift_pass"], tp["ms_r_pass"])):
        gbs = by / (ms * 1e-3) / 1e9 if ms > 0 else None
        phases[name] = {"GB": round(by / 1e9, 2), "ms": round(ms, 2), "GBs": None if gbs is None else round(gbs, 1),
                        "frac": None if gbs is None else round(gbs / HBM_PEAK_GBS, 4)}
    phases["r_pass"]["of_which_shift_updates_GB"] = round(st["bytes_shift_pass"] / 1e9, 2)
    dom = max(phases, key=lambda k: phases[k]["ms"])
    kernels = {"stencil": "k_normal_batch<InCgmP, EpiDotInOut> (D^dag D with p = r + alpha p in its prologue and "
                          "<p, Ap> in its epilogue, one launch over all unconverged systems)",
               "r_pass": "k_point_batch<PtCgmWindow> + k_point_batch<PtCgmRing> (at each window's end the x_s, "
                         "p_s updates of every active shift; then r += beta Ap into the residual ring and |r|^2)"}
    out["roofline"] = {
        "bound": "hbm", "kernel": kernels[dom], "phase": dom,
        "achieved": phases[dom]["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": phases[dom]["frac"],
        "bytes_model": "per site and iteration: D^dag D 96 B (64 B in the first); residual pass 48 B; at each "
                       "window's end (every 10 iterations and at every convergence check) 16 B per window term "
                       "+ 48 B per x_s update + 16 B per p_s update, over the shifts active in the window",
        "shift_updates_x": st["shift_updates_x"], "shift_updates_p": st["shift_updates_p"],
        "phases": phases,
        "whole_batch_GBs": round(total / t_batch / 1e9, 1),
        "whole_batch_frac": round(total / t_batch / 1e9 / HBM_PEAK_GBS, 4)}
    if seq:
        t0 = time.perf_counter()
        for lat, m, db, xs in systems:
            glx.dev_cg_m(lat, "staggered_normal", m, xs, db, shifts, resid_freq_check=10, max_iter=100000,
                         eps=1e-10)
        glx.synchronize()
        out["one_call_each_s"] = round(time.perf_counter() - t0, 3)
    del systems
    return out, ({"host": host, "shifts": shifts} if keep_host else None)


def stream_ceiling(glx, footprints=((64 << 20), (1 << 30)), reps=20):
    """What the device delivers to plain streaming kernels (read only, write
    only, copy, and three bytes read per byte written -- the stencils' ratio)
    over a footprint the size of the 1024^2 working set (64 MiB, inside the
    256 MB last-level cache) and one beyond it (1 GiB), HIP-event timed."""
    out = {}
    for fp in footprints:
        key = f"{fp >> 20}MiB"
        out[key] = {}
        for mode in ("read", "write", "copy", "read3_write1"):
            ms, by = glx.time_stream(mode, fp, reps)
            gbs = by / (ms * 1e-3) / 1e9
            out[key][mode] = {"GBs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    glx.release_cache()
    return out


def eo_solves(glx, lat, b, mass, reps=3):
    """D x = b on the same lattice by the even/odd pipeline on compacted
    vectors (prepare, CG on m^2 - D_eo D_oe to 1e-10, reconstruct): FP64
    single-pass and mixed precision, with the full residual |D x - b| / |b|."""
    out = {}
    db = glx.DeviceArray(b)
    dx = glx.DeviceArray(np.zeros_like(b))
    bn = float(np.linalg.norm(b))
    for method in ("cg", "cg_fused", "mixed"):
        glx.dev_solve_staggered_eo(lat, mass, dx, db, method, eps=1e-10)  # warm-up
        ts = []
        for _ in range(reps):
            glx.synchronize()
            t0 = time.perf_counter()
            info, r2 = glx.dev_solve_staggered_eo(lat, mass, dx, db, method, eps=1e-10)
            glx.synchronize()
            ts.append(time.perf_counter() - t0)
        out[method] = {"ms": round(float(np.median(ts)) * 1e3, 3), "iters": info.iter,
                       "full_relres": float(np.sqrt(r2) / bn)}
    return out


def stencil_timing(glx, lat, b, mass, L, reps=200):
    """The operators alone at L^2 and 2 L^2: D (the wave-march kernel on
    lattices of whole 64-column strips, else the tiled kernel) and the fused
    D^dag D (the wave-march kernel; DdD_rowmarch: the row-marching kernel),
    FP64 and FP32 (and FP16 links), HIP events around `reps` back-to-back
    launches.  A lone row-marching D^dag D application takes its own best
    rows per block (the table's entry is tuned for the CG iteration kernel,
    whose prologue and epilogue change the balance); the table is restored
    afterwards.  Algorithmic bytes per application: V (2 s_v + 2 s_u [+ 4])."""
    out = {"op": "D and fused D^dag D, algorithmic bytes V*(2 s_v + 2 s_u [+ s_scale])"}

    def measure(lt, vb, n, key):
        V = n * n
        d_in, d_out = glx.DeviceArray(vb), glx.DeviceArray(np.zeros_like(vb))
        v32 = vb.astype(np.complex64)
        i32, o32 = glx.DeviceArray(v32), glx.DeviceArray(np.zeros_like(v32))
        res = {"L": n}
        for op, tag in (("staggered", "D"), ("staggered_normal", "DdD"), ("staggered_normal", "DdD_rowmarch")):
            saved = (glx.get_normal_rows(n, n, 16), glx.get_normal_rows(n, n, 8))
            # DdD: the wave-march kernel at its own rows; DdD_rowmarch: the
            # row-marching (or tiled, 0) kernel at its best rows per block
            glx.set_ddwave_rows(0 if tag == "DdD_rowmarch" else -1)
            cand = (0, 8, 16, 32, 64) if tag == "DdD_rowmarch" else (None,)
            best = {}
            for rows in cand:
                if rows is not None:
                    glx.set_normal_rows(n, n, 16, rows)
                    glx.set_normal_rows(n, n, 8, rows)
                for prec, fp32, lp, by in (("fp64", False, glx.LINKS_F32, 64), ("fp32", True, glx.LINKS_F32, 32),
                                          ("fp16link", True, glx.LINKS_F16, 28)):
                    di, do = (i32, o32) if fp32 else (d_in, d_out)
                    ms = glx.time_apply(lt, op, mass, do, di, reps=reps, fp32=fp32, link_prec=lp)
                    if prec not in best or ms < best[prec][0]:
                        best[prec] = (ms, rows)
            glx.set_normal_rows(n, n, 16, saved[0])
            glx.set_normal_rows(n, n, 8, saved[1])
            glx.set_ddwave_rows(-1)
            for prec, (ms, rows) in best.items():
                if tag == "D":  # the wave-march kernels' rows per wave (0: another kernel)
                    rows = glx.get_dwave_rows(n, n, 16 if prec == "fp64" else 8)
                elif tag == "DdD":
                    rows = glx.get_ddwave_rows(n, n, 16 if prec == "fp64" else 8)
                by = {"fp64": 64, "fp32": 32, "fp16link": 28}[prec] * V
                gbs = by / (ms * 1e-3) / 1e9
                res[f"{tag}_{prec}"] = {"us": round(ms * 1e3, 2), "GBs": round(gbs, 1),
                                        "frac": round(gbs / HBM_PEAK_GBS, 4), "rows": rows}
        out[key] = res

    measure(lat, b, L, "at_L")
    L2 = 2 * L
    links2, b2 = make_inputs(L2, 0.5, seed=1338)
    lat2 = glx.Lattice(links2, L2, L2, f32=True, f16=True)
    measure(lat2, b2, L2, "at_2L")
    del lat2
    return out


def traffic_from_profile(kernel: str, links: str):
    """HBM bytes per launch of the dominant kernel as measured by rocprofv3
    FETCH_SIZE / WRITE_SIZE (scaled by a copy kernel of known size), read from
    the committed summary for this workload; None if there is none."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_fp32_{links}_L1024_pmc.txt")
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                vals["t"] = float(line.split(":", 1)[1])
    return vals.get("t")


def kernel_only(glx, args):
    """One CG iteration kernel alone on the headline lattice, `reps` launches
    between HIP events on the kernel's stream (no solve, no autotuning): the
    program a counter run profiles.  'batch': the batched multishift call
    alone (with its phase timing), for profiles of its kernels."""
    if args.kernel_only == "batch":
        out, _ = batch_multishift(glx, seq=False, keep_host=False)
        print(json.dumps(out))
        return 0
    if args.kernel_only == "stencil":
        links, b = make_inputs(args.L, args.sigma, seed=1337)
        lat = glx.Lattice(links, args.L, args.L, f32=True, f16=True)
        out = stencil_timing(glx, lat, b, args.mass, args.L, reps=args.reps)
        c_small = stream_ceiling(glx)["64MiB"]["read3_write1"]["GBs"]
        for v in out["at_L"].values():
            if isinstance(v, dict) and "GBs" in v:
                v["frac_of_ceiling"] = round(v["GBs"] / c_small, 4)
        print(json.dumps(out))
        return 0
    L = args.L
    links, _ = make_inputs(L, args.sigma, seed=1337)
    lat = glx.Lattice(links, L, L, f32=True, f16=True)
    lp = glx.LINKS_F16 if args.links == "f16" else glx.LINKS_F32
    if args.sstep_rows > 0:
        glx.set_sstep_rows(L, L, args.sstep_rows)
    prec = {"sscg": 2 + lp, "fast32": lp, "fast64": 0}[args.kernel_only]
    ms, bps = glx.time_cg_kernel(lat, args.mass, prec=prec, reps=args.reps)
    gbs = bps * L * L / (ms * 1e-3) / 1e9
    out = {"kernel": args.kernel_only, "L": L, "links": args.links, "reps": args.reps,
           "sstep_rows": glx.get_sstep_rows(L, L), "sstep_form": glx.get_sstep_form_for(L, L),
           "band_rows": args.band_rows, "us_per_launch": round(ms * 1e3, 3),
           "bytes_per_site": bps, "GBs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    if args.kernel_only == "sscg" and out["sstep_form"] == 1:
        out["traffic_only_us_per_launch"] = round(glx.time_sscg_traffic(lat, lp, reps=args.reps) * 1e3, 3)
    print(json.dumps(out))
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--L", type=int, default=1024)
    ap.add_argument("--mass", type=float, default=0.02)
    ap.add_argument("--sigma", type=float, default=0.5)
    ap.add_argument("--links", choices=["f32", "f16"], default="f32")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-big", action="store_true", help="skip the 4096^2 stencil measurement")
    ap.add_argument("--no-grad", action="store_true", help="skip the 512^2 gradient timing")
    ap.add_argument("--no-eo", action="store_true", help="skip the even/odd D x = b solves")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched multishift workload")
    ap.add_argument("--no-cstencil", action="store_true", help="skip the general stencil_2d timing")
    ap.add_argument("--sstep-rows", type=int, default=0,
                    help="rows per block of the mixed solver's s-step kernel (0: autotune; a profiler changes the "
                         "timings the autotuner compares, so profiled runs pin the shape of the unprofiled one)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="first run 3 device-to-device copies of this many MiB (a plain copy kernel of known size, "
                         "for scaling rocprofv3 byte counters)")
    ap.add_argument("--kernel-only", choices=["sscg", "fast32", "fast64", "batch", "stencil"], default=None,
                    help="only time one CG iteration kernel at --L (HIP events, --reps launches) and print its "
                         "line: the s-step launch, the FP32 one-iteration kernel or the FP64 single-pass kernel "
                         "(for counter runs of that kernel alone); 'batch': only the batched multishift call and "
                         "its phase timing, without the one-call-each comparison; 'stencil': only the operators alone "
                         "(D and D^dag D at L^2 and 2 L^2) and the streaming ceiling")
    ap.add_argument("--reps", type=int, default=100)
    ap.add_argument("--sstep-form", type=int, choices=[0, 1], default=None,
                    help="s-step launch form on D^dag D: 0 blocks of four waves through LDS rings, 1 wave strips in "
                         "registers (default: the library's)")
    ap.add_argument("--band-rows", type=int, default=0, help="rows per wave of the wave-strip form (0: by size)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU: the solver path has no multi-GPU code; "
                                   "run with --gpus 1"}))
        return 2

    import glx

    if args.copy_mib > 0:
        n = args.copy_mib << 16  # complex128 elements (16 bytes)
        src = glx.DeviceArray(np.ones(n, np.complex128))
        dst = glx.DeviceArray(np.zeros(n, np.complex128))
        for _ in range(3):
            glx.copy_device(dst, src)
        glx.synchronize()
        del src, dst

    if args.sstep_form is not None:
        glx.set_sstep_form(args.sstep_form, args.band_rows)

    if args.kernel_only:
        return kernel_only(glx, args)

    L, mass = args.L, args.mass
    V = L * L
    links, b = make_inputs(L, args.sigma, seed=1337)
    lat = glx.Lattice(links, L, L, f32=True, f16=True)
    glx.set_tuning(group=16, use_graph=False)
    tuned = glx.autotune_normal(lat, reps=40)
    if args.sstep_rows > 0:
        glx.set_sstep_rows(L, L, args.sstep_rows)
        tuned_ss = {"rows": args.sstep_rows}
    else:
        tuned_ss = glx.autotune_sstep(lat, reps=40)
    d_b = glx.DeviceArray(b)
    d_x = glx.DeviceArray(np.zeros_like(b))
    d_x64 = glx.DeviceArray(np.zeros_like(b))
    zero = np.zeros_like(b)
    link_prec = glx.LINKS_F16 if args.links == "f16" else glx.LINKS_F32

    def solve_fp64():
        d_x64.upload(zero)
        glx.synchronize()
        t0 = time.perf_counter()
        info = glx.dev_solve("cg", lat, "staggered_normal", mass, d_x64, d_b, max_iter=100000, eps=1e-10)
        glx.synchronize()
        return time.perf_counter() - t0, info

    def solve_fused():
        d_x64.upload(zero)
        glx.synchronize()
        t0 = time.perf_counter()
        info = glx.dev_cg_fused(lat, "staggered_normal", mass, d_x64, d_b, max_iter=100000, eps=1e-10)
        glx.synchronize()
        return time.perf_counter() - t0, info

    def solve_mixed(sstep=2):
        glx.set_sstep(sstep)
        d_x.upload(zero)
        glx.synchronize()
        t0 = time.perf_counter()
        info = glx.dev_cg_mixed(lat, "staggered_normal", mass, d_x, d_b, max_iter=100000, eps=1e-10, delta=0.1,
                                link_prec=link_prec)
        glx.synchronize()
        glx.set_sstep()
        return time.perf_counter() - t0, info

    for _ in range(args.warmup):
        solve_fp64()
        solve_fused()
        solve_mixed(1)
        solve_mixed()
    tfu = []
    for _ in range(max(1, args.steps)):
        dt, infofu = solve_fused()
        tfu.append(dt)
    xfu = d_x64.download()
    t64 = []
    for _ in range(max(1, args.steps)):
        dt, info64 = solve_fp64()
        t64.append(dt)
    tm1 = []
    for _ in range(max(1, args.steps)):
        dt, infom1 = solve_mixed(1)
        tm1.append(dt)
    tmx = []
    for _ in range(args.steps):
        dt, infomx = solve_mixed()
        tmx.append(dt)
    ms64 = float(np.median(t64)) * 1e3
    msfu = float(np.median(tfu)) * 1e3
    msmx = float(np.median(tmx)) * 1e3
    msm1 = float(np.median(tm1)) * 1e3
    best64 = min(ms64, msfu)
    x64 = d_x64.download()
    xmx = d_x.download()
    err_vs_fp64 = float(np.linalg.norm(xmx - x64) / np.linalg.norm(x64))

    # live roofline of the mixed solve's dominant kernel: the s-step launch
    # (two CG iterations: p, r, x from the previous basis, then A p, A^2 p,
    # A r and twelve FP64 partial sums), HIP events around back-to-back
    # launches on its own stream; the one-iteration kernel beside it
    it_ms, it_bps = glx.time_cg_kernel(lat, mass, prec=2 + link_prec, reps=200)
    ss_form = glx.get_sstep_form_for(L, L)  # the form the timed launch (and the solves) took
    it1_ms, it1_bps = glx.time_cg_kernel(lat, mass, prec=link_prec, reps=200)
    it64_ms, it64_bps = glx.time_cg_kernel(lat, mass, prec=0, reps=200)
    it_gbs = it_bps * V / (it_ms * 1e-3) / 1e9
    it64_gbs = it64_bps * V / (it64_ms * 1e-3) / 1e9
    # the wave-strip launch's loads and stores alone (no arithmetic): the
    # floor its access pattern sets; frac_of_floor = that floor / the launch
    pattern = None
    if ss_form == 1:
        pat_ms = glx.time_sscg_traffic(lat, link_prec, reps=200)
        pat_gbs = it_bps * V / (pat_ms * 1e-3) / 1e9
        pattern = {"ms_per_launch": pat_ms, "GBs": round(pat_gbs, 1), "frac": round(pat_gbs / HBM_PEAK_GBS, 4),
                   "frac_of_floor": round(pat_ms / it_ms, 4)}

    stencil = stencil_timing(glx, lat, b, mass, L)
    big = None
    if not args.no_big:
        big = stencil_beyond_mall(glx)

    ceiling = stream_ceiling(glx)
    c_small = ceiling["64MiB"]["read3_write1"]["GBs"]
    c_big = ceiling["1024MiB"]["read3_write1"]["GBs"]
    for key in ("at_L", "at_2L"):
        for v in stencil[key].values():
            if isinstance(v, dict) and "GBs" in v:
                v["frac_of_ceiling"] = round(v["GBs"] / c_small, 4)
    if big is not None:
        big["frac_of_ceiling"] = round(big["fp32_GBs"] / c_big, 4)
        big["fp64_frac_of_ceiling"] = round(big["fp64_GBs"] / c_big, 4)

    eo = None if args.no_eo else eo_solves(glx, lat, b, mass)
    grad = None if args.no_grad else gradient_timing(glx)
    batch, batch_host = (None, None) if args.no_batch else batch_multishift(glx, keep_host=not args.no_cpu)
    cst = None if args.no_cstencil else general_stencil(glx)
    small = small_configs(glx)
    gauge = None if args.no_cstencil else gauge_generation(glx)

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(links, b, L, mass, {"fp64": x64, "fp64_fused": xfu, "mixed": xmx}, grad=grad,
                               batch=batch_host, small=small)
        except Exception as e:  # the baseline is reported, never required
            cpu = {"value": None, "unit": "ms", "cores": 1, "kind": "port", "sample": f"failed: {e}"}
    del batch_host
    glx.release_cache()
    rel_cpu = (cpu or {}).get("relerr_vs_cpu", {})

    out = {
        "metric": "CG time-to-1e-10 & stencil HBM GB/s, 2D U(1) staggered: FP64 vs mixed",
        "value": round(msmx, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(msmx, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"fp32 bulk ({args.links} links) + fp64 reliable updates",
        "data": "synthetic (numpy seed 1337): theta ~ N(0, 0.5^2), complex Gaussian rhs",
        "config": {
            "workload": f"CG on staggered D^dag D, L={L}x{L}, m={mass}, theta~N(0,{args.sigma}^2), "
                        f"true relres 1e-10; FP64 CG vs mixed (delta 0.1)",
            "L": L, "mass": mass, "tol": 1e-10, "delta": 0.1,
        },
        "fp64_ms": round(ms64, 3),
        "fp64_fused_ms": round(msfu, 3),
        "speedup_vs_fp64": round(ms64 / msmx, 3),
        "speedup_basis": "plain FP64 GPU CG in the reference's order of operations (glx_dev_solve)",
        "speedup_vs_fp64_fused": round(msfu / msmx, 3),
        "mixed_form": (("s-step, two CG iterations per launch, "
                        + ("wave strips (form 1)" if glx.get_sstep_form_for(L, L) == 1
                           else f"{tuned_ss['rows']} rows per block (form 0)")
                        + f" ({'set' if args.sstep_rows > 0 else 'autotuned'})")
                       if tuned_ss["rows"] > 0 else "one CG iteration per launch (autotuned)"),
        "mixed_single_pass_ms": round(msm1, 3),
        "mixed_single_pass_iters": infom1.iter,
        "fp64_iters": info64.iter,
        "fp64_fused_iters": infofu.iter,
        "fp64_fused_relerr_vs_fp64": float(np.linalg.norm(xfu - x64) / np.linalg.norm(x64)),
        "mixed_iters": infomx.iter,
        "mixed_reliable_updates": infomx.reliable_updates,
        "mixed_true_relres": float(np.sqrt(infomx.resSq) / np.linalg.norm(b)),
        "fp64_true_relres": float(np.sqrt(info64.resSq) / np.linalg.norm(b)),
        "mixed_relerr_vs_fp64": err_vs_fp64,
        "mixed_relerr_vs_cpu": rel_cpu.get("mixed"),
        "fp64_relerr_vs_cpu": rel_cpu.get("fp64"),
        "fp64_fused_relerr_vs_cpu": rel_cpu.get("fp64_fused"),
        "normal_rows": {"fp64": tuned["rows64"], "fp32": tuned["rows32"]},
        "stencil": dict(stencil, beyond_mall=big),
        "stream_ceiling": ceiling,
        "cg_kernel_fp64": {"ms_per_launch": it64_ms, "bytes_per_site": it64_bps, "GBs": round(it64_gbs, 1)},
        "cg_kernel_fp32_single_pass": {"ms_per_launch": it1_ms, "bytes_per_site": it1_bps,
                                       "GBs": round(it1_bps * V / (it1_ms * 1e-3) / 1e9, 1)},
        "roofline": {
            "bound": "hbm",
            "kernel": (f"k_sscg_w<Links{args.links.upper()}> (s-step mixed CG launch = two CG iterations, wave-strip "
                       f"form: p, r, x from the previous basis + A p, A^2 p, A r + 12 partial sums)"
                       if ss_form == 1 else
                       f"k_sscg<2, Links{args.links.upper()}> (s-step mixed CG launch = two CG iterations: "
                       f"p, r, x from the previous basis + A p, A^2 p, A r + 12 partial sums)"),
            "achieved": round(it_gbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(it_gbs / HBM_PEAK_GBS, 4),
            "traffic": traffic_from_profile("k_sscg_w" if ss_form == 1 else "k_sscg", args.links),
            "bytes_per_launch": it_bps * V,
            "bytes_per_site": it_bps,
            "ms_per_launch": it_ms,
            "pattern_floor": pattern,
        },
        "eo_solve_D": eo,
        "gradient": grad,
        "batch_multishift": batch,
        "general_stencil": cst,
        "small_configs": {k: {kk: vv for kk, vv in v.items() if kk != "x"} for k, v in small[0].items()},
        "gauge_generation": gauge,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
