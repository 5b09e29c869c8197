"""ctypes front end for the oracle (TEST INFRASTRUCTURE ONLY).

Two libraries sit behind the same Python calls:

* ``liborc.so`` -- the plain-C restatement of the reference's solver path
  (orc_*.c); it travels with the tree and is the checker on any machine.
* ``_ref/libref.so`` -- the reference compiled from its own sources
  (oracle/Makefile); present only where the reference was available at build
  time.  It pins the restatement.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg may
import this module.  Vectors are numpy arrays: float64 for real, complex128 for
complex (interleaved re/im, the layout of std::complex<double>); links are a
complex128 array of 2*V entries, AoS ``u[2*(y*lx+x)+mu]``.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

# Operator ids shared with ref_shim.cpp.
OPS = {
    "laplace_real": 0,
    "laplace": 1,
    "laplace_u1": 2,
    "staggered": 3,
    "staggered_dagger": 4,
    "staggered_normal": 5,
    "staggered_gamma5": 6,
    "gamma5": 7,
    "deo": 8,
    "doe": 9,
    "m2mdeodoe": 10,
    "stencil": 11,
    "stencil_dagger": 12,
    "stencil_gamma5": 13,
    "stencil_eo": 14,
    "stencil_oe": 15,
    "stencil_m2mdeodoe": 16,
    "staggered_free": 30,
    "staggered_gamma5_free": 31,
    "staggered_2linklaplace": 32,
    "symmshift_x": 33,
    "symmshift_y": 34,
    "index_operator": 35,
}
SOLVERS = {"cg": 0, "bicgstab": 1, "gcr": 2}

_dp = C.POINTER(C.c_double)


def _p(a):
    return a.ctypes.data_as(_dp) if a is not None else None


class OrcInfo(C.Structure):
    _fields_ = [("resSq", C.c_double), ("iter", C.c_int), ("success", C.c_int),
                ("ops_count", C.c_int), ("name", C.c_char * 32)]


class OrcVerb(C.Structure):
    _fields_ = [("level", C.c_int), ("prefix", C.c_char_p), ("fmt", C.c_char_p),
                ("buf", C.c_void_p), ("cap", C.c_size_t), ("len", C.c_size_t)]


class OrcStagOp(C.Structure):
    _fields_ = [("lattice", C.c_void_p), ("mass", C.c_double), ("x_fine", C.c_int),
                ("y_fine", C.c_int), ("Nc", C.c_int), ("wilson_coeff", C.c_double)]


class OrcStencil(C.Structure):
    _fields_ = [("lx", C.c_int), ("ly", C.c_int), ("clover", C.c_void_p), ("hopping", C.c_void_p),
                ("shift", C.c_double * 2), ("eo_shift", C.c_double * 2), ("dof_shift", C.c_double * 2)]


class OrcCStencil(C.Structure):
    _fields_ = [("lx", C.c_int), ("ly", C.c_int), ("nc", C.c_int), ("has_two", C.c_int), ("sdir", C.c_int),
                ("clover", C.c_void_p), ("hopping", C.c_void_p), ("two_link", C.c_void_p),
                ("shift", C.c_double * 2), ("eo_shift", C.c_double * 2), ("dof_shift", C.c_double * 2)]


class C2(C.Structure):
    """double _Complex returned by value (two SSE registers)."""
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


class OrcMT(C.Structure):
    _fields_ = [("mt", C.c_uint32 * 624), ("idx", C.c_int)]


@dataclass
class SolveInfo:
    resSq: float
    iter: int
    success: bool
    ops_count: int
    name: str = ""
    text: str = ""


def build(quiet: bool = True) -> None:
    """Compile liborc.so (and _ref/libref.so where the reference exists)."""
    import subprocess
    out = subprocess.run(["make", "-C", HERE], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)


_orc = None
_ref = None


def orc_lib():
    global _orc
    if _orc is None:
        path = os.path.join(HERE, "liborc.so")
        if not os.path.exists(path):
            build()
        lib = C.CDLL(path)
        vp, i, d = C.c_void_p, C.c_int, C.c_double
        for nm in ("orc_cg_r", "orc_cg_c", "orc_bicgstab_r", "orc_bicgstab_c", "orc_gcr_r", "orc_gcr_c",
                   "orc_cg_cx", "orc_bicgstab_cx", "orc_gcr_cx"):
            f = getattr(lib, nm)
            f.restype = OrcInfo
            f.argtypes = [vp, vp, i, i, d, vp, vp, C.POINTER(OrcVerb), vp]
        for nm in ("orc_cg_restart_r", "orc_cg_restart_c", "orc_bicgstab_restart_r", "orc_bicgstab_restart_c",
                   "orc_gcr_restart_r", "orc_gcr_restart_c", "orc_cg_restart_cx", "orc_bicgstab_restart_cx",
                   "orc_gcr_restart_cx"):
            f = getattr(lib, nm)
            f.restype = OrcInfo
            f.argtypes = [vp, vp, i, i, d, i, vp, vp, C.POINTER(OrcVerb)]
        for nm in ("orc_cg_m_r", "orc_cg_m_c", "orc_cg_m_cx"):
            f = getattr(lib, nm)
            f.restype = OrcInfo
            f.argtypes = [vp, vp, i, i, i, i, d, vp, vp, vp, i, C.POINTER(OrcVerb), vp]
        lib.orc_dot_r.restype = d
        lib.orc_dot_r.argtypes = [vp, vp, i]
        lib.orc_norm2sq_r.restype = d
        lib.orc_norm2sq_r.argtypes = [vp, i]
        lib.orc_norm2sq_c.restype = d
        lib.orc_norm2sq_c.argtypes = [vp, i]
        lib.orc_diffnorm2sq_r.restype = d
        lib.orc_diffnorm2sq_r.argtypes = [vp, vp, i]
        lib.orc_diffnorm2sq_c.restype = d
        lib.orc_diffnorm2sq_c.argtypes = [vp, vp, i]
        lib.orc_dot_c.restype = C2
        lib.orc_dot_c.argtypes = [vp, vp, i]
        lib.orc_dot_cx.restype = C2
        lib.orc_dot_cx.argtypes = [vp, vp, i]
        lib.orc_norm2sq_cx.restype = d
        lib.orc_norm2sq_cx.argtypes = [vp, i]
        lib.orc_diffnorm2sq_cx.restype = d
        lib.orc_diffnorm2sq_cx.argtypes = [vp, vp, i]
        lib.orc_mt_seed.argtypes = [C.POINTER(OrcMT), C.c_uint32]
        lib.orc_mt_next.restype = C.c_uint32
        lib.orc_mt_next.argtypes = [C.POINTER(OrcMT)]
        lib.orc_canonical.restype = d
        lib.orc_canonical.argtypes = [C.POINTER(OrcMT)]
        lib.orc_gaussian_r.argtypes = [vp, i, C.POINTER(OrcMT)]
        lib.orc_gaussian_c.argtypes = [vp, i, C.POINTER(OrcMT), i]
        lib.orc_unit_gauge_u1.argtypes = [vp, i, i]
        lib.orc_rand_gauge_u1.argtypes = [vp, i, i, C.POINTER(OrcMT)]
        lib.orc_gauss_gauge_u1.argtypes = [vp, i, i, C.POINTER(OrcMT), d]
        lib.orc_rand_trans_u1.argtypes = [vp, i, i, C.POINTER(OrcMT)]
        lib.orc_apply_gauge_trans_u1.argtypes = [vp, vp, i, i]
        lib.orc_philox4x32_10.argtypes = [vp, C.c_uint32, C.c_uint32]
        lib.orc_philox_gauge_u1.argtypes = [vp, i, i, d, C.c_uint64, C.c_uint32]
        lib.orc_plaquette_u1.restype = C2
        lib.orc_plaquette_u1.argtypes = [vp, i, i]
        lib.orc_read_gauge_u1.restype = i
        lib.orc_read_gauge_u1.argtypes = [vp, i, i, C.c_char_p]
        lib.orc_write_gauge_u1.restype = i
        lib.orc_write_gauge_u1.argtypes = [vp, i, i, C.c_char_p]
        lib.orc_coord_to_index.restype = i
        lib.orc_coord_to_index.argtypes = [vp, i, vp, i]
        lib.orc_index_to_coord.argtypes = [vp, i, i, vp, vp]
        lib.orc_index_is_even.restype = i
        lib.orc_index_is_even.argtypes = [vp, i, i]
        for nm in ("orc_stencil_staggered", "orc_stencil_staggered_gamma5", "orc_stencil_staggered_dagger"):
            getattr(lib, nm).argtypes = [C.POINTER(OrcStencil), C.POINTER(OrcStagOp)]
        lib.orc_staggered_eoprec_prepare.argtypes = [vp, vp, vp]
        lib.orc_staggered_eoprec_reconstruct.argtypes = [vp, vp, vp, vp]
        lib.orc_stencil_eoprec_prepare.argtypes = [vp, vp, vp]
        lib.orc_cstencil_apply.argtypes = [vp, vp, vp, i]
        lib.orc_cstencil_eoprec_prepare.argtypes = [vp, vp, vp]
        lib.orc_cstencil_m2mdeodoe.argtypes = [vp, vp, vp]
        lib.orc_cstencil_eoprec_reconstruct.argtypes = [vp, vp, vp, vp]
        lib.orc_cstencil_generate.argtypes = [vp, vp, vp]
        lib.orc_stencil_eoprec_reconstruct.argtypes = [vp, vp, vp, vp]
        lib.orc_staggered_force.argtypes = [vp, vp, vp, C.POINTER(OrcStagOp)]
        lib.orc_staggered_dS_dm.restype = d
        lib.orc_staggered_dS_dm.argtypes = [vp, C.POINTER(OrcStagOp)]
        _orc = lib
    return _orc


def ref_available() -> bool:
    return os.path.exists(os.path.join(HERE, "_ref", "libref.so"))


def ref_lib():
    """The reference build (None when oracle/_ref was not built here)."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libref.so")
        if not os.path.exists(path):
            return None
        lib = C.CDLL(path)
        vp, i, d = C.c_void_p, C.c_int, C.c_double
        lib.ref_apply.argtypes = [i, vp, vp, vp, d, i, i, i]
        lib.ref_eoprec_prepare.argtypes = [i, vp, vp, vp, d, i, i]
        lib.ref_eoprec_reconstruct.argtypes = [i, vp, vp, vp, vp, d, i, i]
        lib.ref_stencil_coeffs.argtypes = [i, vp, vp, vp, vp, d, i, i]
        lib.ref_solve.argtypes = [i, i, vp, vp, i, i, d, i, vp, d, i, i, i, i, C.c_char_p, i, i, vp, i, vp]
        lib.ref_cg_m.argtypes = [i, vp, vp, i, i, i, i, d, vp, i, vp, d, i, i, i, i, C.c_char_p, i, i, vp, i, vp, vp]
        lib.ref_blas.argtypes = [i, i, vp, vp, i, vp]
        lib.ref_rng_new.restype = vp
        lib.ref_rng_new.argtypes = [C.c_uint]
        lib.ref_rng_free.argtypes = [vp]
        lib.ref_gaussian_r.argtypes = [vp, vp, i]
        lib.ref_gaussian_c.argtypes = [vp, vp, i]
        lib.ref_gauss_gauge_u1.argtypes = [vp, vp, i, i, d]
        lib.ref_rand_gauge_u1.argtypes = [vp, vp, i, i]
        lib.ref_rand_trans_u1.argtypes = [vp, vp, i, i]
        lib.ref_unit_gauge_u1.argtypes = [vp, i, i]
        lib.ref_apply_gauge_trans_u1.argtypes = [vp, vp, i, i]
        lib.ref_plaquette_u1.argtypes = [vp, i, i, vp]
        lib.ref_read_gauge_u1.argtypes = [vp, i, i, C.c_char_p]
        lib.ref_write_gauge_u1.argtypes = [vp, i, i, C.c_char_p]
        lib.ref_ape_smear_u1.argtypes = [vp, vp, i, i, d, i]
        lib.ref_apply_wc.argtypes = [i, vp, vp, vp, d, d, i, i]
        lib.ref_apply_wc.restype = i
        lib.ref_cstencil_apply.argtypes = [i, i, i, i, i, i, vp, vp, vp, vp, vp, vp, vp]
        lib.ref_cstencil_apply.restype = i
        lib.ref_cstencil_generate.argtypes = [i, vp, d, i, i, i, i, vp, vp, vp, vp, i, vp, vp, vp]
        lib.ref_cstencil_generate.restype = i
        lib.ref_topo_u1.argtypes = [vp, i, i]
        lib.ref_topo_u1.restype = d
        _ref = lib
    return _ref


# --------------------------------------------------------------------------
# Operators (oracle restatement).
_ORC_OP_SYMS = {
    "laplace_real": "orc_laplace_real",
    "laplace": "orc_laplace",
    "laplace_u1": "orc_laplace_u1",
    "staggered": "orc_staggered_u1",
    "staggered_dagger": "orc_staggered_dagger_u1",
    "staggered_normal": "orc_staggered_normal_u1",
    "staggered_gamma5": "orc_staggered_gamma5_u1",
    "gamma5": "orc_gamma5",
    "deo": "orc_staggered_deo_u1",
    "doe": "orc_staggered_doe_u1",
    "m2mdeodoe": "orc_staggered_m2mdeodoe_u1",
    "stencil": "orc_stencil_apply",
    "stencil_dagger": "orc_stencil_apply",
    "stencil_gamma5": "orc_stencil_apply",
    "stencil_eo": "orc_stencil_apply_eo",
    "stencil_oe": "orc_stencil_apply_oe",
    "stencil_m2mdeodoe": "orc_stencil_m2mdeodoe",
    "staggered_free": "orc_staggered_free",
    "staggered_gamma5_free": "orc_staggered_gamma5_free",
    "staggered_2linklaplace": "orc_staggered_2linklaplace_u1",
    "symmshift_x": "orc_staggered_symmshift_x",
    "symmshift_y": "orc_staggered_symmshift_y",
    "index_operator": "orc_staggered_index_operator",
}


class OpContext:
    """Keeps the C structs (and the arrays they point into) alive for a call."""

    def __init__(self, op: str, links, mass: float, lx: int, ly: int, nc: int = 1):
        lib = orc_lib()
        self.op = op
        self.links = None if links is None else np.ascontiguousarray(links, dtype=np.complex128)
        self.st = OrcStagOp(self.links.ctypes.data if self.links is not None else None, mass, lx, ly, nc, 0.0)
        self.fn = C.cast(getattr(lib, _ORC_OP_SYMS[op]), C.c_void_p)
        self.extra = C.cast(C.pointer(self.st), C.c_void_p)
        if op.startswith("stencil"):
            V = lx * ly
            self.clover = np.zeros(V, np.complex128)
            self.hopping = np.zeros(4 * V, np.complex128)
            self.sten = OrcStencil(lx, ly, self.clover.ctypes.data, self.hopping.ctypes.data)
            builder = {"stencil_dagger": lib.orc_stencil_staggered_dagger,
                       "stencil_gamma5": lib.orc_stencil_staggered_gamma5}.get(op, lib.orc_stencil_staggered)
            builder(C.byref(self.sten), C.byref(self.st))
            self.extra = C.cast(C.pointer(self.sten), C.c_void_p)


def apply(op: str, rhs, links=None, mass: float = 0.0, lx: int = 0, ly: int = 0, nc: int = 1, impl: str = "orc",
          wilson: float = 0.0):
    """lhs = A rhs with operator `op`; impl 'orc' (restatement) or 'ref';
    `wilson` is staggered_u1_op.wilson_coeff (the two-link Laplacian term)."""
    real = op == "laplace_real"
    rhs = np.ascontiguousarray(rhs, dtype=np.float64 if real else np.complex128)
    lhs = np.empty_like(rhs)
    if impl == "ref":
        lib = ref_lib()
        lk = None if links is None else np.ascontiguousarray(links, dtype=np.complex128)
        if wilson != 0.0:
            rc = lib.ref_apply_wc(OPS[op], lhs.ctypes.data, rhs.ctypes.data, None if lk is None else lk.ctypes.data,
                                  mass, wilson, lx, ly)
        else:
            rc = lib.ref_apply(OPS[op], lhs.ctypes.data, rhs.ctypes.data, None if lk is None else lk.ctypes.data,
                               mass, lx, ly, nc)
        assert rc == 0
        return lhs
    ctx = OpContext(op, links, mass, lx, ly, nc)
    ctx.st.wilson_coeff = wilson
    proto = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.c_void_p)
    proto(ctx.fn.value)(lhs.ctypes.data, rhs.ctypes.data, ctx.extra)
    return lhs


def stencil_coeffs(op: str, links, mass: float, lx: int, ly: int, impl: str = "orc"):
    V = lx * ly
    if impl == "ref":
        clover = np.zeros(V, np.complex128)
        hopping = np.zeros(4 * V, np.complex128)
        sh = np.zeros(6)
        lk = np.ascontiguousarray(links, dtype=np.complex128)
        ref_lib().ref_stencil_coeffs(OPS[op], clover.ctypes.data, hopping.ctypes.data, sh.ctypes.data, lk.ctypes.data,
                                     mass, lx, ly)
        return clover, hopping, sh
    ctx = OpContext(op, links, mass, lx, ly)
    sh = np.array([ctx.sten.shift[0], ctx.sten.shift[1], ctx.sten.eo_shift[0], ctx.sten.eo_shift[1],
                   ctx.sten.dof_shift[0], ctx.sten.dof_shift[1]])
    return ctx.clover.copy(), ctx.hopping.copy(), sh


def eoprec_prepare(rhs, links, mass, lx, ly, stencil=False, impl="orc"):
    rhs = np.ascontiguousarray(rhs, dtype=np.complex128)
    out = np.empty_like(rhs)
    if impl == "ref":
        lk = np.ascontiguousarray(links, dtype=np.complex128)
        ref_lib().ref_eoprec_prepare(int(stencil), out.ctypes.data, rhs.ctypes.data, lk.ctypes.data, mass, lx, ly)
        return out
    ctx = OpContext("stencil" if stencil else "staggered", links, mass, lx, ly)
    fn = orc_lib().orc_stencil_eoprec_prepare if stencil else orc_lib().orc_staggered_eoprec_prepare
    fn(out.ctypes.data, rhs.ctypes.data, ctx.extra)
    return out


def eoprec_reconstruct(lhs_e, rhs_o, links, mass, lx, ly, stencil=False, impl="orc"):
    lhs_e = np.ascontiguousarray(lhs_e, dtype=np.complex128)
    rhs_o = np.ascontiguousarray(rhs_o, dtype=np.complex128)
    out = np.empty_like(lhs_e)
    if impl == "ref":
        lk = np.ascontiguousarray(links, dtype=np.complex128)
        ref_lib().ref_eoprec_reconstruct(int(stencil), out.ctypes.data, lhs_e.ctypes.data, rhs_o.ctypes.data,
                                         lk.ctypes.data, mass, lx, ly)
        return out
    ctx = OpContext("stencil" if stencil else "staggered", links, mass, lx, ly)
    fn = orc_lib().orc_stencil_eoprec_reconstruct if stencil else orc_lib().orc_staggered_eoprec_reconstruct
    fn(out.ctypes.data, lhs_e.ctypes.data, rhs_o.ctypes.data, ctx.extra)
    return out


# --------------------------------------------------------------------------
# Solvers.
def solve(solver: str, op: str, b, x0=None, *, links=None, mass=0.0, lx=0, ly=0, nc=1, max_iter=10000,
          eps=1e-10, restart_freq=0, verb_level=0, prefix="", fmt="%g", impl="orc", hist=False,
          reductions="serial"):
    """Run CG / BiCGStab / GCR (optionally restarted).  Returns (x, SolveInfo[, hist]).

    reductions="extended" (complex operators, impl "orc"): the same solver with
    its dot products and norms accumulated in long double and rounded once
    (orc_dot_cx), instead of the reference's serial double sums."""
    real = op == "laplace_real"
    dt = np.float64 if real else np.complex128
    b = np.ascontiguousarray(b, dtype=dt)
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=dt, copy=True)
    size = b.size
    cap = 1 << 22
    buf = C.create_string_buffer(cap)
    if impl == "ref":
        sci = 1 if "e" in fmt else 0
        prec = int(fmt.split(".")[1][0]) if "." in fmt else 0
        info = np.zeros(4)
        lk = None if links is None else np.ascontiguousarray(links, dtype=np.complex128)
        rc = ref_lib().ref_solve(SOLVERS[solver], OPS[op], x.ctypes.data, b.ctypes.data, size, max_iter, eps,
                                 restart_freq, None if lk is None else lk.ctypes.data, mass, lx, ly, nc, verb_level,
                                 prefix.encode(), sci, prec, buf, cap, info.ctypes.data)
        assert rc == 0
        name = {0: "CG", 1: "BiCGStab", 2: "GCR"}[SOLVERS[solver]]
        if restart_freq > 0:
            name = f"{name}({restart_freq})"
        si = SolveInfo(info[0], int(info[1]), bool(info[2]), int(info[3]), name, buf.value.decode())
        return (x, si, None) if hist else (x, si)
    lib = orc_lib()
    ctx = OpContext(op, links, mass, lx, ly, nc)
    verb = OrcVerb(verb_level, prefix.encode(), fmt.encode(), C.cast(buf, C.c_void_p), cap, 0)
    sfx = "r" if real else "c"
    if reductions == "extended":
        if real:
            raise ValueError("extended reductions exist for the complex solvers only")
        sfx = "cx"
    h = np.zeros(max(max_iter, 1)) if hist else None
    if restart_freq > 0:
        f = getattr(lib, f"orc_{solver}_restart_{sfx}")
        r = f(x.ctypes.data, b.ctypes.data, size, max_iter, eps, restart_freq, ctx.fn, ctx.extra, C.byref(verb))
    else:
        f = getattr(lib, f"orc_{solver}_{sfx}")
        r = f(x.ctypes.data, b.ctypes.data, size, max_iter, eps, ctx.fn, ctx.extra, C.byref(verb),
              None if h is None else h.ctypes.data)
    text = C.string_at(C.cast(buf, C.c_void_p), verb.len).decode()
    si = SolveInfo(r.resSq, r.iter, bool(r.success), r.ops_count, r.name.decode(), text)
    if hist:
        return x, si, h[: si.iter] if restart_freq == 0 else None
    return x, si


def cg_m(op: str, b, shifts, *, links=None, mass=0.0, lx=0, ly=0, nc=1, resid_freq_check=10, max_iter=10000,
         eps=1e-10, worst_first=False, verb_level=0, prefix="", fmt="%g", impl="orc", reductions="serial"):
    """Multishift CG.  Returns (X[n_shift, size], SolveInfo, resSqmrhs, shifts_after).
    reductions="extended": long double reductions, as in solve()."""
    real = op == "laplace_real"
    dt = np.float64 if real else np.complex128
    b = np.ascontiguousarray(b, dtype=dt)
    sh = np.array(shifts, dtype=np.float64, copy=True)
    n = sh.size
    size = b.size
    X = np.full((n, size), np.nan, dtype=dt)
    res = np.zeros(n)
    cap = 1 << 22
    buf = C.create_string_buffer(cap)
    if impl == "ref":
        sci = 1 if "e" in fmt else 0
        prec = int(fmt.split(".")[1][0]) if "." in fmt else 0
        info = np.zeros(4)
        lk = None if links is None else np.ascontiguousarray(links, dtype=np.complex128)
        rc = ref_lib().ref_cg_m(OPS[op], X.ctypes.data, b.ctypes.data, n, size, resid_freq_check, max_iter, eps,
                                sh.ctypes.data, int(worst_first), None if lk is None else lk.ctypes.data, mass, lx, ly,
                                nc, verb_level, prefix.encode(), sci, prec, buf, cap, info.ctypes.data, res.ctypes.data)
        assert rc == 0
        return X, SolveInfo(info[0], int(info[1]), bool(info[2]), int(info[3]), "CG-M", buf.value.decode()), res, sh
    lib = orc_lib()
    ctx = OpContext(op, links, mass, lx, ly, nc)
    verb = OrcVerb(verb_level, prefix.encode(), fmt.encode(), C.cast(buf, C.c_void_p), cap, 0)
    ptrs = (C.c_void_p * n)(*[X[k].ctypes.data for k in range(n)])
    if reductions == "extended" and real:
        raise ValueError("extended reductions exist for the complex solvers only")
    f = lib.orc_cg_m_r if real else (lib.orc_cg_m_cx if reductions == "extended" else lib.orc_cg_m_c)
    r = f(ptrs, b.ctypes.data, n, size, resid_freq_check, max_iter, eps, sh.ctypes.data, ctx.fn, ctx.extra,
          int(worst_first), C.byref(verb), res.ctypes.data)
    text = C.string_at(C.cast(buf, C.c_void_p), verb.len).decode()
    return X, SolveInfo(r.resSq, r.iter, bool(r.success), r.ops_count, "CG-M", text), res, sh


# --------------------------------------------------------------------------
# BLAS-1 and random inputs.
def blas(which: str, a, b=None, impl="orc"):
    """generic_vector.h's dot / norm2sq / diffnorm2sq (serial sums); impl
    "orcx" is the long double variant (complex only)."""
    k = {"dot": 0, "norm2sq": 1, "diffnorm2sq": 2}[which]
    cplx = np.iscomplexobj(a)
    dt = np.complex128 if cplx else np.float64
    a = np.ascontiguousarray(a, dtype=dt)
    b = a if b is None else np.ascontiguousarray(b, dtype=dt)
    n = a.size
    if impl == "ref":
        out = np.zeros(2)
        ref_lib().ref_blas(k, int(cplx), a.ctypes.data, b.ctypes.data, n, out.ctypes.data)
        return complex(out[0], out[1]) if (cplx and k == 0) else out[0]
    lib = orc_lib()
    if impl == "orcx":
        if not cplx:
            raise ValueError("the long double reductions are complex only")
        if k == 0:
            v = lib.orc_dot_cx(a.ctypes.data, b.ctypes.data, n)
            return complex(v.re, v.im)
        if k == 1:
            return lib.orc_norm2sq_cx(a.ctypes.data, n)
        return lib.orc_diffnorm2sq_cx(a.ctypes.data, b.ctypes.data, n)
    if k == 0:
        if cplx:
            v = lib.orc_dot_c(a.ctypes.data, b.ctypes.data, n)
            return complex(v.re, v.im)
        return lib.orc_dot_r(a.ctypes.data, b.ctypes.data, n)
    if k == 1:
        return (lib.orc_norm2sq_c if cplx else lib.orc_norm2sq_r)(a.ctypes.data, n)
    return (lib.orc_diffnorm2sq_c if cplx else lib.orc_diffnorm2sq_r)(a.ctypes.data, b.ctypes.data, n)


class Rng:
    """One seeded std::mt19937 stream, drawn by the oracle or by the reference."""

    # libstdc++/g++ on x86-64 evaluates std::complex's two constructor
    # arguments right to left, so the second draw becomes the real part; this
    # is pinned by tests against the reference build.
    IMAG_FIRST = 1

    def __init__(self, seed: int = 1337, impl: str = "orc"):
        self.impl = impl
        if impl == "ref":
            self.h = ref_lib().ref_rng_new(seed)
        else:
            self.mt = OrcMT()
            orc_lib().orc_mt_seed(C.byref(self.mt), seed)

    def __del__(self):
        if getattr(self, "impl", None) == "ref" and getattr(self, "h", None):
            ref_lib().ref_rng_free(self.h)

    def gaussian_real(self, n: int):
        v = np.zeros(n)
        if self.impl == "ref":
            ref_lib().ref_gaussian_r(self.h, v.ctypes.data, n)
        else:
            orc_lib().orc_gaussian_r(v.ctypes.data, n, C.byref(self.mt))
        return v

    def gaussian_complex(self, n: int):
        v = np.zeros(n, np.complex128)
        if self.impl == "ref":
            ref_lib().ref_gaussian_c(self.h, v.ctypes.data, n)
        else:
            orc_lib().orc_gaussian_c(v.ctypes.data, n, C.byref(self.mt), self.IMAG_FIRST)
        return v

    def gauss_gauge(self, lx: int, ly: int, beta: float):
        u = np.zeros(2 * lx * ly, np.complex128)
        if self.impl == "ref":
            ref_lib().ref_gauss_gauge_u1(self.h, u.ctypes.data, lx, ly, beta)
        else:
            orc_lib().orc_gauss_gauge_u1(u.ctypes.data, lx, ly, C.byref(self.mt), beta)
        return u

    def rand_gauge(self, lx: int, ly: int):
        u = np.zeros(2 * lx * ly, np.complex128)
        if self.impl == "ref":
            ref_lib().ref_rand_gauge_u1(self.h, u.ctypes.data, lx, ly)
        else:
            orc_lib().orc_rand_gauge_u1(u.ctypes.data, lx, ly, C.byref(self.mt))
        return u

    def rand_trans(self, lx: int, ly: int):
        t = np.zeros(lx * ly, np.complex128)
        if self.impl == "ref":
            ref_lib().ref_rand_trans_u1(self.h, t.ctypes.data, lx, ly)
        else:
            orc_lib().orc_rand_trans_u1(t.ctypes.data, lx, ly, C.byref(self.mt))
        return t

    def canonical(self) -> float:
        assert self.impl == "orc"
        return orc_lib().orc_canonical(C.byref(self.mt))


def philox4x32_10(counter, key):
    """Philox4x32-10 block: four 32-bit counter words, two key words."""
    c = np.array(counter, dtype=np.uint32)
    orc_lib().orc_philox4x32_10(c.ctypes.data, int(key[0]), int(key[1]))
    return [int(v) for v in c]


def philox_gauge(lx, ly, beta, seed, stream=0):
    """The library's device field generator, restated on the CPU."""
    u = np.zeros(2 * lx * ly, np.complex128)
    orc_lib().orc_philox_gauge_u1(u.ctypes.data, lx, ly, float(beta), int(seed), int(stream))
    return u


def unit_gauge(lx, ly):
    u = np.zeros(2 * lx * ly, np.complex128)
    orc_lib().orc_unit_gauge_u1(u.ctypes.data, lx, ly)
    return u


def apply_gauge_trans(u, t, lx, ly, impl="orc"):
    u = np.array(u, dtype=np.complex128, copy=True)
    t = np.ascontiguousarray(t, dtype=np.complex128)
    if impl == "ref":
        ref_lib().ref_apply_gauge_trans_u1(u.ctypes.data, t.ctypes.data, lx, ly)
    else:
        orc_lib().orc_apply_gauge_trans_u1(u.ctypes.data, t.ctypes.data, lx, ly)
    return u


def plaquette(u, lx, ly, impl="orc"):
    u = np.ascontiguousarray(u, dtype=np.complex128)
    if impl == "ref":
        out = np.zeros(2)
        ref_lib().ref_plaquette_u1(u.ctypes.data, lx, ly, out.ctypes.data)
        return complex(out[0], out[1])
    v = orc_lib().orc_plaquette_u1(u.ctypes.data, lx, ly)
    return complex(v.re, v.im)


def read_gauge(path, lx, ly, impl="orc"):
    u = np.zeros(2 * lx * ly, np.complex128)
    if impl == "ref":
        ref_lib().ref_read_gauge_u1(u.ctypes.data, lx, ly, str(path).encode())
    else:
        rc = orc_lib().orc_read_gauge_u1(u.ctypes.data, lx, ly, str(path).encode())
        if rc != 0:
            raise IOError(f"read_gauge_u1 failed ({rc}) for {path}")
    return u


def write_gauge(path, u, lx, ly, impl="orc"):
    u = np.ascontiguousarray(u, dtype=np.complex128)
    if impl == "ref":
        ref_lib().ref_write_gauge_u1(u.ctypes.data, lx, ly, str(path).encode())
    else:
        orc_lib().orc_write_gauge_u1(u.ctypes.data, lx, ly, str(path).encode())


def force(chi, psi, links, mass, lx, ly):
    """FP64 CPU gradient dS/dtheta (2V, AoS like the links) and dS/dm."""
    lib = orc_lib()
    links = np.ascontiguousarray(links, dtype=np.complex128)
    chi = np.ascontiguousarray(chi, dtype=np.complex128)
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    st = OrcStagOp(links.ctypes.data, mass, lx, ly, 1, 0.0)
    out = np.zeros(2 * lx * ly)
    lib.orc_staggered_force(out.ctypes.data, chi.ctypes.data, psi.ctypes.data, C.byref(st))
    dm = lib.orc_staggered_dS_dm(chi.ctypes.data, C.byref(st))
    return out, dm


def lattice_index(dims, nc, coord, color):
    d = np.ascontiguousarray(dims, dtype=np.int32)
    c = np.ascontiguousarray(coord, dtype=np.int32)
    return orc_lib().orc_coord_to_index(d.ctypes.data, nc, c.ctypes.data, color)


def lattice_coord(dims, nc, i):
    d = np.ascontiguousarray(dims, dtype=np.int32)
    c = np.zeros(2, np.int32)
    col = np.zeros(1, np.int32)
    orc_lib().orc_index_to_coord(d.ctypes.data, nc, i, c.ctypes.data, col.ctypes.data)
    return (int(c[0]), int(c[1])), int(col[0])


def lattice_is_even(dims, nc, i):
    d = np.ascontiguousarray(dims, dtype=np.int32)
    return bool(orc_lib().orc_index_is_even(d.ctypes.data, nc, i))


# --------------------------------------------------------------------------
# stencil_2d for any Nc, up to two links (coarse_stencil.cpp:12-1627).
PIECES = {"all": 0, "eo": 1, "oe": 2, "tb": 3, "bt": 4}
SDIR = {"all": 0, "clover": 1, "+x": 2, "+y": 3, "-x": 4, "-y": 5, "+2x": 6, "+x+y": 7, "+2y": 8, "-x+y": 9,
        "-2x": 10, "-x-y": 11, "-2y": 12, "+x-y": 13}


class CStencil:
    """Coefficients of a general stencil_2d in the reference layout:
    clover (V nc nc), hopping (4 V nc nc), two_link (8 V nc nc or None);
    element (term, site, row, column) at ((term V + site) nc + row) nc + column."""

    def __init__(self, lx, ly, nc, clover=None, hopping=None, two_link=None, shift=0j, eo_shift=0j, dof_shift=0j,
                 sdir=0, has_two=None):
        self.lx, self.ly, self.nc = lx, ly, nc
        blk = lx * ly * nc * nc
        self.has_two = (two_link is not None) if has_two is None else bool(has_two)
        z = lambda k: np.zeros(k * blk, np.complex128)
        self.clover = z(1) if clover is None else np.ascontiguousarray(clover, dtype=np.complex128)
        self.hopping = z(4) if hopping is None else np.ascontiguousarray(hopping, dtype=np.complex128)
        self.two_link = (z(8) if two_link is None else np.ascontiguousarray(two_link, dtype=np.complex128)) \
            if self.has_two else None
        self.shift, self.eo_shift, self.dof_shift = complex(shift), complex(eo_shift), complex(dof_shift)
        self.sdir = SDIR[sdir] if isinstance(sdir, str) else int(sdir)

    @classmethod
    def random(cls, rng, lx, ly, nc, has_two=False, shifts=(0j, 0j, 0j), sdir=0):
        blk = lx * ly * nc * nc
        g = lambda k: rng.standard_normal(k * blk) + 1j * rng.standard_normal(k * blk)
        return cls(lx, ly, nc, g(1), g(4), g(8) if has_two else None, *shifts, sdir=sdir)

    def shifts6(self):
        return np.array([self.shift.real, self.shift.imag, self.eo_shift.real, self.eo_shift.imag,
                         self.dof_shift.real, self.dof_shift.imag])

    def struct(self):
        st = OrcCStencil()
        st.lx, st.ly, st.nc, st.has_two, st.sdir = self.lx, self.ly, self.nc, int(self.has_two), self.sdir
        st.clover, st.hopping = self.clover.ctypes.data, self.hopping.ctypes.data
        st.two_link = self.two_link.ctypes.data if self.has_two else None
        st.shift[:] = [self.shift.real, self.shift.imag]
        st.eo_shift[:] = [self.eo_shift.real, self.eo_shift.imag]
        st.dof_shift[:] = [self.dof_shift.real, self.dof_shift.imag]
        return st

    @property
    def n(self):
        return self.lx * self.ly * self.nc


def _ref_cstencil(piece, st, lhs, rhs, aux=None):
    rc = ref_lib().ref_cstencil_apply(piece, st.lx, st.ly, st.nc, int(st.has_two), st.sdir, st.clover.ctypes.data,
                                      st.hopping.ctypes.data, st.two_link.ctypes.data if st.has_two else None,
                                      st.shifts6().ctypes.data, lhs.ctypes.data, rhs.ctypes.data,
                                      None if aux is None else aux.ctypes.data)
    assert rc == 0


def cstencil_apply(st, rhs, piece="all", impl="orc", lhs=None):
    """lhs = piece of the stencil applied to rhs; `lhs` (if given) is the
    initial content, which some sdir/piece combinations leave untouched."""
    rhs = np.ascontiguousarray(rhs, dtype=np.complex128)
    out = np.zeros_like(rhs) if lhs is None else np.array(lhs, dtype=np.complex128, copy=True)
    if impl == "ref":
        _ref_cstencil(PIECES[piece], st, out, rhs)
    else:
        s = st.struct()
        orc_lib().orc_cstencil_apply(out.ctypes.data, rhs.ctypes.data, C.byref(s), PIECES[piece])
    return out


def cstencil_m2mdeodoe(st, rhs, impl="orc"):
    rhs = np.ascontiguousarray(rhs, dtype=np.complex128)
    out = np.zeros_like(rhs)
    if impl == "ref":
        _ref_cstencil(5, st, out, rhs)
    else:
        s = st.struct()
        orc_lib().orc_cstencil_m2mdeodoe(out.ctypes.data, rhs.ctypes.data, C.byref(s))
    return out


def cstencil_eoprec_prepare(st, rhs, impl="orc"):
    rhs = np.ascontiguousarray(rhs, dtype=np.complex128)
    out = np.zeros_like(rhs)
    if impl == "ref":
        _ref_cstencil(6, st, out, rhs)
    else:
        s = st.struct()
        orc_lib().orc_cstencil_eoprec_prepare(out.ctypes.data, rhs.ctypes.data, C.byref(s))
    return out


def cstencil_eoprec_reconstruct(st, lhs_e, rhs_o, impl="orc"):
    lhs_e = np.ascontiguousarray(lhs_e, dtype=np.complex128)
    rhs_o = np.ascontiguousarray(rhs_o, dtype=np.complex128)
    out = np.zeros_like(lhs_e)
    if impl == "ref":
        _ref_cstencil(7, st, out, lhs_e, rhs_o)
    else:
        s = st.struct()
        orc_lib().orc_cstencil_eoprec_reconstruct(out.ctypes.data, lhs_e.ctypes.data, rhs_o.ctypes.data, C.byref(s))
    return out


def cstencil_generate(lx, ly, nc, has_two, op=None, links=None, mass=0.0, src=None, impl="orc"):
    """generate_stencil_2d: probe an operator (a function-form op by name, or
    another general stencil `src`) with unit vectors."""
    out = CStencil(lx, ly, nc, two_link=np.zeros(8 * lx * ly * nc * nc, np.complex128) if has_two else None)
    if impl == "ref":
        lk = None if links is None else np.ascontiguousarray(links, dtype=np.complex128)
        z6 = np.zeros(6)
        ref_lib().ref_cstencil_generate(
            OPS[op] if op is not None else -1, None if lk is None else lk.ctypes.data, mass, lx, ly, nc, int(has_two),
            None if src is None else src.clover.ctypes.data, None if src is None else src.hopping.ctypes.data,
            None if src is None or not src.has_two else src.two_link.ctypes.data,
            (z6 if src is None else src.shifts6()).ctypes.data, 0 if src is None else int(src.has_two),
            out.clover.ctypes.data, out.hopping.ctypes.data, out.two_link.ctypes.data if has_two else None)
        return out
    s = out.struct()
    if src is not None:
        ss = src.struct()
        orc_lib().orc_cstencil_generate(C.byref(s), C.cast(orc_lib().orc_cstencil_apply_all, C.c_void_p), C.byref(ss))
    else:
        ctx = OpContext(op, links, mass, lx, ly, nc)
        orc_lib().orc_cstencil_generate(C.byref(s), ctx.fn, ctx.extra)
    return out
