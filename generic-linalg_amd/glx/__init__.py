"""Python binding of libglx.so, the MI355X solver library (ctypes, no torch).

Mirrors the reference's interface for the solver path: ``minv_vector_cg``,
``minv_vector_cg_restart``, ``minv_vector_cg_m``, ``minv_vector_bicgstab``
(+ ``_restart``), ``minv_vector_gcr`` (+ ``_restart``) with the operator
callbacks of ``operators.h`` named as in the reference (``square_staggered_u1``
...).  Vectors are numpy arrays (float64 for the real Laplacian, complex128
otherwise); links are a complex128 array of 2*lx*ly values,
``links[2*(y*lx+x)+mu]``.  Every call runs on the GPU through the C ABI of
include/glx.h; there is no CPU path, and a missing library or GPU raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
# GLX_LIBRARY names another build of the library (e.g. to compare two
# versions of a kernel in one run); the default is the in-tree build.
LIB_PATH = os.environ.get("GLX_LIBRARY") or os.path.join(PKG, "libglx.so")

OP_KINDS = {
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
    "stencil_eo": 14,
    "stencil_oe": 15,
    "stencil_m2mdeodoe": 16,
    "staggered_eo_schur": 17,
}
# C function implementing each operator callback (include/glx.h)
OP_SYMBOLS = {
    "laplace_real": "glx_square_laplace_real",
    "laplace": "glx_square_laplace",
    "laplace_u1": "glx_square_laplace_u1",
    "staggered": "glx_square_staggered_u1",
    "staggered_dagger": "glx_square_staggered_dagger_u1",
    "staggered_normal": "glx_square_staggered_normal_u1",
    "staggered_gamma5": "glx_square_staggered_gamma5_u1",
    "gamma5": "glx_gamma_5",
    "deo": "glx_square_staggered_deo_u1",
    "doe": "glx_square_staggered_doe_u1",
    "m2mdeodoe": "glx_square_staggered_m2mdeodoe_u1",
    "stencil": "glx_apply_stencil_2d",
    "stencil_dagger": "glx_apply_stencil_2d",
    "stencil_gamma5": "glx_apply_stencil_2d",
    "stencil_eo": "glx_apply_stencil_2d_eo",
    "stencil_oe": "glx_apply_stencil_2d_oe",
    "stencil_m2mdeodoe": "glx_apply_square_staggered_m2mdeodoe_stencil",
    "staggered_free": "glx_square_staggered",
    "staggered_gamma5_free": "glx_square_staggered_gamma5",
    "staggered_2linklaplace": "glx_square_staggered_2linklaplace_u1",
    "symmshift_x": "glx_staggered_symmshift_x",
    "symmshift_y": "glx_staggered_symmshift_y",
    "index_operator": "glx_staggered_index_operator",
}
SOLVER_IDS = {"cg": 0, "bicgstab": 1, "gcr": 2}
VERB_NONE, VERB_SUMMARY, VERB_RESTART_DETAIL, VERB_DETAIL = 0, 1, 2, 3
LINKS_F32, LINKS_F16 = 1, 2


class GlxError(RuntimeError):
    pass


class Verbose(C.Structure):
    _fields_ = [("verbosity", C.c_int), ("verb_prefix", C.c_char_p), ("precond_verbosity", C.c_int),
                ("precond_verb_prefix", C.c_char_p), ("out_buf", C.c_void_p), ("out_cap", C.c_size_t),
                ("out_len", C.c_size_t), ("float_format", C.c_int), ("float_precision", C.c_int)]


class InversionInfo(C.Structure):
    _fields_ = [("resSq", C.c_double), ("iter", C.c_int), ("success", C.c_int), ("ops_count", C.c_int),
                ("name", C.c_char * 32), ("resSqmrhs", C.c_void_p), ("n_rhs", C.c_int),
                ("reliable_updates", C.c_int)]


class StaggeredU1Op(C.Structure):
    _fields_ = [("lattice", C.c_void_p), ("mass", C.c_double), ("x_fine", C.c_int), ("y_fine", C.c_int),
                ("Nc", C.c_int), ("wilson_coeff", C.c_double)]


class Stencil2D(C.Structure):
    _fields_ = [("lx", C.c_int), ("ly", C.c_int), ("clover", C.c_void_p), ("hopping", C.c_void_p),
                ("shift", C.c_double * 2), ("eo_shift", C.c_double * 2), ("dof_shift", C.c_double * 2)]


class Stencil2DNc(C.Structure):
    _fields_ = [("lx", C.c_int), ("ly", C.c_int), ("nc", C.c_int), ("has_two", C.c_int), ("sdir", C.c_int),
                ("clover", C.c_void_p), ("hopping", C.c_void_p), ("two_link", C.c_void_p),
                ("shift", C.c_double * 2), ("eo_shift", C.c_double * 2), ("dof_shift", C.c_double * 2)]


@dataclass
class Info:
    """inversion_info (inverter_struct.h:14-89)."""
    resSq: float
    iter: int
    success: bool
    ops_count: int
    name: str
    resSqmrhs: list = field(default_factory=list)
    reliable_updates: int = 0
    text: str = ""


_lib = None


def lib():
    """Load libglx.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GlxError(f"{LIB_PATH} is missing: build it with `make -C {PKG}`")
        L = C.CDLL(LIB_PATH)
        vp, i, d = C.c_void_p, C.c_int, C.c_double
        VP = C.POINTER(Verbose)
        IP = C.POINTER(InversionInfo)
        L.glx_last_error.restype = C.c_char_p
        for nm in ("cg", "bicgstab", "gcr"):
            for t in ("r", "c"):
                f = getattr(L, f"glx_minv_vector_{nm}_{t}")
                f.restype = i
                f.argtypes = [vp, vp, i, i, d, vp, vp, VP, IP]
                f = getattr(L, f"glx_minv_vector_{nm}_restart_{t}")
                f.restype = i
                f.argtypes = [vp, vp, i, i, d, i, vp, vp, VP, IP]
        for t in ("r", "c"):
            f = getattr(L, f"glx_minv_unpreconditioned_{t}")
            f.restype = i
            f.argtypes = [vp, vp, i, i, vp, vp, vp, VP, IP]
        for t in ("r", "c"):
            f = getattr(L, f"glx_minv_vector_cg_m_{t}")
            f.restype = i
            f.argtypes = [vp, vp, i, i, i, i, d, vp, vp, vp, i, VP, IP]
        for sym in set(OP_SYMBOLS.values()):
            getattr(L, sym).argtypes = [vp, vp, vp]
        L.glx_square_staggered_eoprec_prepare.argtypes = [vp, vp, vp]
        L.glx_square_staggered_eoprec_reconstruct.argtypes = [vp, vp, vp, vp]
        L.glx_apply_square_staggered_eoprec_prepare_stencil.argtypes = [vp, vp, vp]
        L.glx_apply_square_staggered_eoprec_reconstruct_stencil.argtypes = [vp, vp, vp, vp]
        for nm in ("glx_get_square_staggered_u1_stencil", "glx_get_square_staggered_gamma5_u1_stencil",
                   "glx_get_square_staggered_dagger_u1_stencil"):
            getattr(L, nm).restype = i
            getattr(L, nm).argtypes = [C.POINTER(Stencil2D), C.POINTER(StaggeredU1Op)]
        L.glx_register_operator.restype = i
        L.glx_register_operator.argtypes = [vp, i, C.POINTER(StaggeredU1Op)]
        L.glx_unregister_operator.restype = i
        L.glx_unregister_operator.argtypes = [vp]
        L.glx_lattice_create.restype = i
        L.glx_lattice_create.argtypes = [C.POINTER(vp), vp, i, i, i]
        L.glx_lattice_update_links.restype = i
        L.glx_lattice_update_links.argtypes = [vp, vp]
        L.glx_lattice_create_gaussian.restype = i
        L.glx_lattice_create_gaussian.argtypes = [C.POINTER(vp), i, i, i, C.c_double, C.c_ulonglong, C.c_uint]
        L.glx_lattice_gaussian_links.restype = i
        L.glx_lattice_gaussian_links.argtypes = [vp, C.c_double, C.c_ulonglong, C.c_uint]
        L.glx_lattice_get_links.restype = i
        L.glx_lattice_get_links.argtypes = [vp, vp]
        L.glx_lattice_plaquette.restype = i
        L.glx_lattice_plaquette.argtypes = [vp, vp]
        for f in ("glx_apply_stencil_2d_nc", "glx_apply_stencil_2d_nc_eo", "glx_apply_stencil_2d_nc_oe",
                  "glx_apply_stencil_2d_nc_tb", "glx_apply_stencil_2d_nc_bt", "glx_apply_stencil_2d_nc_m2mdeodoe"):
            getattr(L, f).restype = None
            getattr(L, f).argtypes = [vp, vp, vp]
        L.glx_stencil_2d_nc_eoprec_prepare.restype = None
        L.glx_stencil_2d_nc_eoprec_prepare.argtypes = [vp, vp, vp]
        L.glx_stencil_2d_nc_eoprec_reconstruct.restype = None
        L.glx_stencil_2d_nc_eoprec_reconstruct.argtypes = [vp, vp, vp, vp]
        L.glx_generate_stencil_2d_nc.restype = i
        L.glx_generate_stencil_2d_nc.argtypes = [vp, vp, vp]
        L.glx_cstencil_create.restype = i
        L.glx_cstencil_create.argtypes = [C.POINTER(vp), vp]
        L.glx_cstencil_destroy.restype = i
        L.glx_cstencil_destroy.argtypes = [vp]
        L.glx_dev_cstencil_apply.restype = i
        L.glx_dev_cstencil_apply.argtypes = [vp, i, i, vp, vp]
        L.glx_dev_cstencil_solve.restype = i
        L.glx_dev_cstencil_solve.argtypes = [i, vp, i, vp, vp, i, C.c_double, vp, vp]
        L.glx_dev_time_cstencil.restype = i
        L.glx_dev_time_cstencil.argtypes = [vp, i, vp, vp, i, C.POINTER(C.c_double)]
        L.glx_lattice_destroy.restype = i
        L.glx_lattice_destroy.argtypes = [vp]
        L.glx_dev_solve.restype = i
        L.glx_dev_solve.argtypes = [i, vp, i, d, i, vp, vp, i, d, VP, IP]
        L.glx_dev_cg_m.restype = i
        L.glx_dev_cg_m.argtypes = [vp, i, d, i, vp, vp, i, vp, i, i, d, i, VP, IP]
        L.glx_dev_cg_mixed.restype = i
        L.glx_dev_cg_mixed.argtypes = [vp, i, d, vp, vp, i, d, d, i, VP, IP]
        if hasattr(L, "glx_dev_cg_mixed_noalloc"):  # absent from builds that predate it
            L.glx_dev_cg_mixed_noalloc.restype = i
            L.glx_dev_cg_mixed_noalloc.argtypes = [vp, i, d, vp, vp, i, d, d, i, VP, IP]
        L.glx_dev_cg_fused.restype = i
        L.glx_dev_cg_fused.argtypes = [vp, i, d, vp, vp, i, d, VP, IP]
        L.glx_set_tuning.restype = i
        L.glx_set_tuning.argtypes = [i, i]
        L.glx_set_sstep.restype = i
        L.glx_set_sstep.argtypes = [i, i]
        L.glx_get_sstep.restype = i
        L.glx_get_sstep.argtypes = [C.POINTER(i), C.POINTER(i)]
        L.glx_set_sstep_rows.restype = i
        L.glx_set_sstep_rows.argtypes = [i, i, i]
        L.glx_get_sstep_rows.restype = i
        L.glx_get_sstep_rows.argtypes = [i, i, C.POINTER(i)]
        if hasattr(L, "glx_set_sstep_form"):  # absent from builds that predate it
            L.glx_set_sstep_form.restype = i
            L.glx_set_sstep_form.argtypes = [i, i]
            L.glx_get_sstep_form.restype = i
            L.glx_get_sstep_form.argtypes = [C.POINTER(i)]
            L.glx_set_sstep_form_entry.restype = i
            L.glx_set_sstep_form_entry.argtypes = [i, i, i]
            L.glx_get_sstep_form_for.restype = i
            L.glx_get_sstep_form_for.argtypes = [i, i, C.POINTER(i)]
        L.glx_autotune_sstep.restype = i
        L.glx_autotune_sstep.argtypes = [vp, i, C.POINTER(i), C.POINTER(d)]
        L.glx_dev_time_cg_kernel.restype = i
        L.glx_dev_time_cg_kernel.argtypes = [vp, d, i, i, C.POINTER(d), C.POINTER(d)]
        L.glx_dev_staggered_grad.restype = i
        L.glx_dev_staggered_grad.argtypes = [vp, d, vp, i, i, d, vp, vp, C.POINTER(d), C.POINTER(d), C.POINTER(d), IP]
        L.glx_dev_cg_m_batched.restype = i
        L.glx_dev_cg_m_batched.argtypes = [i, vp, i, vp, i, i, d, i, i, vp]
        if hasattr(L, "glx_dev_cg_m_batched_ex"):
            L.glx_dev_cg_m_batched_ex.restype = i
            L.glx_dev_cg_m_batched_ex.argtypes = [i, vp, i, vp, i, i, d, i, i, i, vp, vp]
        L.glx_dev_solve_staggered_eo.restype = i
        L.glx_dev_solve_staggered_eo.argtypes = [vp, d, vp, vp, i, i, d, C.POINTER(d), IP]
        L.glx_dev_eo_prepare.restype = i
        L.glx_dev_eo_prepare.argtypes = [vp, d, vp, vp]
        L.glx_dev_eo_reconstruct.restype = i
        L.glx_dev_eo_reconstruct.argtypes = [vp, d, vp, vp, vp]
        L.glx_set_normal_rows.restype = i
        L.glx_set_normal_rows.argtypes = [i, i, i, i]
        L.glx_set_dwave_rows.restype = i
        L.glx_set_dwave_rows.argtypes = [i]
        L.glx_set_ddwave_rows.restype = i
        L.glx_set_ddwave_rows.argtypes = [i]
        L.glx_get_ddwave_rows.restype = i
        L.glx_get_ddwave_rows.argtypes = [i, i, i, C.POINTER(i)]
        L.glx_get_dwave_rows.restype = i
        L.glx_get_dwave_rows.argtypes = [i, i, i, C.POINTER(i)]
        L.glx_get_normal_rows.restype = i
        L.glx_get_normal_rows.argtypes = [i, i, i, C.POINTER(i)]
        L.glx_autotune_normal.restype = i
        L.glx_autotune_normal.argtypes = [vp, i, C.POINTER(i), C.POINTER(i), C.POINTER(d), C.POINTER(d)]
        L.glx_dev_apply.restype = i
        L.glx_dev_apply.argtypes = [vp, i, d, i, i, i, vp, vp]
        if hasattr(L, "glx_dev_time_sscg_traffic"):
            L.glx_dev_time_sscg_traffic.restype = i
            L.glx_dev_time_sscg_traffic.argtypes = [C.c_void_p, i, i, C.POINTER(d)]
        if hasattr(L, "glx_dev_time_stream"):
            L.glx_dev_time_stream.restype = i
            L.glx_dev_time_stream.argtypes = [i, C.c_size_t, i, C.POINTER(d), C.POINTER(d)]
        L.glx_dev_time_apply.restype = i
        L.glx_dev_time_apply.argtypes = [vp, i, d, i, i, vp, vp, i, C.POINTER(C.c_double)]
        L.glx_dev_alloc.restype = i
        L.glx_dev_alloc.argtypes = [C.POINTER(vp), C.c_size_t]
        L.glx_dev_free.restype = i
        L.glx_dev_free.argtypes = [vp]
        L.glx_memcpy_htod.restype = i
        L.glx_memcpy_htod.argtypes = [vp, vp, C.c_size_t]
        L.glx_memcpy_dtoh.restype = i
        L.glx_memcpy_dtoh.argtypes = [vp, vp, C.c_size_t]
        L.glx_memcpy_dtod.restype = i
        L.glx_memcpy_dtod.argtypes = [vp, vp, C.c_size_t]
        L.glx_device_synchronize.restype = i
        L.glx_device_count.restype = i
        L.glx_device_count.argtypes = [C.POINTER(C.c_int)]
        L.glx_release_cache.restype = i
        _lib = L
    return _lib


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise GlxError(f"{what} failed ({rc}): {lib().glx_last_error().decode()}")


def device_count() -> int:
    n = C.c_int(0)
    rc = lib().glx_device_count(C.byref(n))
    return n.value if rc == 0 else 0


def op_callback(op: str) -> int:
    """Address of the C callback for an operator name."""
    return C.cast(getattr(lib(), OP_SYMBOLS[op]), C.c_void_p).value


class _Extra:
    """Keeps the extra_info struct (and the arrays it points at) alive."""

    def __init__(self, op: str, links, mass: float, lx: int, ly: int, nc: int = 1, wilson: float = 0.0):
        self.links = None if links is None else np.ascontiguousarray(links, dtype=np.complex128)
        self.st = StaggeredU1Op(None if self.links is None else self.links.ctypes.data, mass, lx, ly, nc, wilson)
        self.ptr = C.cast(C.pointer(self.st), C.c_void_p)
        if op.startswith("stencil"):
            V = lx * ly
            self.clover = np.zeros(V, np.complex128)
            self.hopping = np.zeros(4 * V, np.complex128)
            self.sten = Stencil2D(lx, ly, self.clover.ctypes.data, self.hopping.ctypes.data)
            builder = {"stencil_dagger": "glx_get_square_staggered_dagger_u1_stencil",
                       "stencil_gamma5": "glx_get_square_staggered_gamma5_u1_stencil"}.get(
                op, "glx_get_square_staggered_u1_stencil")
            _check(getattr(lib(), builder)(C.byref(self.sten), C.byref(self.st)), builder)
            self.ptr = C.cast(C.pointer(self.sten), C.c_void_p)


class _Text:
    def __init__(self, level: int, prefix: str, fmt: str):
        # 4 MiB for verbose output; a silent call prints nothing (allocating
        # and zeroing the large buffer cost ~0.7 ms per call)
        self.cap = (1 << 22) if level > 0 else 16
        self.buf = C.create_string_buffer(self.cap)
        self.prefix = prefix.encode()
        sci = "e" in fmt
        prec = int(fmt.split(".")[1].rstrip("eg")) if "." in fmt else 6
        self.v = Verbose(level, self.prefix, 0, b"", C.cast(self.buf, C.c_void_p), self.cap, 0, 1 if sci else 0, prec)

    def text(self) -> str:
        return C.string_at(C.cast(self.buf, C.c_void_p), self.v.out_len).decode()


def _info(ci: InversionInfo, text: str = "", res=None) -> Info:
    return Info(ci.resSq, ci.iter, bool(ci.success), ci.ops_count, ci.name.decode(),
                [] if res is None else list(res), ci.reliable_updates, text)


# --------------------------------------------------------------------------
def apply(op: str, rhs, links=None, mass: float = 0.0, lx: int = 0, ly: int = 0, nc: int = 1, wilson: float = 0.0):
    """lhs = A rhs on the GPU through the registered operator callback
    (`wilson`: staggered_u1_op.wilson_coeff)."""
    real = op == "laplace_real"
    rhs = np.ascontiguousarray(rhs, dtype=np.float64 if real else np.complex128)
    lhs = np.empty_like(rhs)
    ex = _Extra(op, links, mass, lx, ly, nc, wilson)
    getattr(lib(), OP_SYMBOLS[op])(lhs.ctypes.data, rhs.ctypes.data, ex.ptr)
    if np.isnan(lhs).any():
        raise GlxError(f"{op} failed: {lib().glx_last_error().decode()}")
    return lhs


def eoprec_prepare(rhs, links, mass, lx, ly, stencil=False):
    rhs = np.ascontiguousarray(rhs, dtype=np.complex128)
    out = np.empty_like(rhs)
    ex = _Extra("stencil" if stencil else "staggered", links, mass, lx, ly)
    fn = lib().glx_apply_square_staggered_eoprec_prepare_stencil if stencil else lib().glx_square_staggered_eoprec_prepare
    fn(out.ctypes.data, rhs.ctypes.data, ex.ptr)
    return out


def eoprec_reconstruct(lhs_e, rhs_o, links, mass, lx, ly, stencil=False):
    lhs_e = np.ascontiguousarray(lhs_e, dtype=np.complex128)
    rhs_o = np.ascontiguousarray(rhs_o, dtype=np.complex128)
    out = np.empty_like(lhs_e)
    ex = _Extra("stencil" if stencil else "staggered", links, mass, lx, ly)
    fn = (lib().glx_apply_square_staggered_eoprec_reconstruct_stencil if stencil
          else lib().glx_square_staggered_eoprec_reconstruct)
    fn(out.ctypes.data, lhs_e.ctypes.data, rhs_o.ctypes.data, ex.ptr)
    return out


def stencil_coeffs(op: str, links, mass: float, lx: int, ly: int):
    """(clover, hopping, shifts6) of get_square_staggered_*_u1_stencil."""
    ex = _Extra(op, links, mass, lx, ly)
    s = ex.sten
    sh = np.array([s.shift[0], s.shift[1], s.eo_shift[0], s.eo_shift[1], s.dof_shift[0], s.dof_shift[1]])
    return ex.clover.copy(), ex.hopping.copy(), sh


def solve(solver: str, op: str, b, x0=None, *, links=None, mass=0.0, lx=0, ly=0, nc=1, max_iter=10000, eps=1e-10,
          restart_freq=0, verb_level=0, prefix="", fmt="%g", callback=None, extra=None, wilson=0.0):
    """minv_vector_{cg,bicgstab,gcr}[_restart] with host vectors.

    `callback` / `extra` override the operator with a raw callback address and
    extra_info pointer (e.g. a ctypes host callback)."""
    real = op == "laplace_real"
    dt = np.float64 if real else np.complex128
    b = np.ascontiguousarray(b, dtype=dt)
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=dt, copy=True)
    ex = None if callback is not None else _Extra(op, links, mass, lx, ly, nc, wilson)
    cb = callback if callback is not None else op_callback(op)
    ep = extra if callback is not None else ex.ptr
    tx = _Text(verb_level, prefix, fmt)
    ci = InversionInfo()
    t = "r" if real else "c"
    L = lib()
    if restart_freq > 0:
        rc = getattr(L, f"glx_minv_vector_{solver}_restart_{t}")(
            x.ctypes.data, b.ctypes.data, b.size, max_iter, eps, restart_freq, cb, ep, C.byref(tx.v), C.byref(ci))
    else:
        rc = getattr(L, f"glx_minv_vector_{solver}_{t}")(
            x.ctypes.data, b.ctypes.data, b.size, max_iter, eps, cb, ep, C.byref(tx.v), C.byref(ci))
    _check(rc, f"minv_vector_{solver}")
    return x, _info(ci, tx.text())


# minv_inverter (generic_inverters.h:70-81)
MINV = {"cg": 0, "cr": 1, "gcr": 2, "bicgstab": 3, "bicgstab_l": 4, "gmres": 5, "sor": 6, "minres": 7,
        "invalid": -1}


class InverterParams(C.Structure):
    """minv_inverter_params (generic_inverters.h:86-105), field by field."""
    _fields_ = [("tol", C.c_double), ("max_iters", C.c_int), ("restart", C.c_int), ("restart_freq", C.c_int),
                ("sor_omega", C.c_double), ("minres_omega", C.c_double), ("bicgstabl_l", C.c_int)]


def minv_unpreconditioned(solver: str, op: str, b, x0=None, *, links=None, mass=0.0, lx=0, ly=0, nc=1, tol=1e-10,
                          max_iters=10000, restart=False, restart_freq=0, verb_level=0, prefix="", fmt="%g",
                          callback=None, extra=None):
    """The reference's dispatcher minv_unpreconditioned
    (generic_inverter.cpp:18-188) with host vectors; `callback` / `extra`
    as in solve()."""
    real = op == "laplace_real"
    dt = np.float64 if real else np.complex128
    b = np.ascontiguousarray(b, dtype=dt)
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=dt, copy=True)
    ex = None if callback is not None else _Extra(op, links, mass, lx, ly, nc)
    cb = callback if callback is not None else op_callback(op)
    ep = extra if callback is not None else ex.ptr
    prm = InverterParams(tol, max_iters, int(restart), restart_freq, 1.0, 1.0, 2)
    tx = _Text(verb_level, prefix, fmt)
    ci = InversionInfo()
    f = lib().glx_minv_unpreconditioned_r if real else lib().glx_minv_unpreconditioned_c
    _check(f(x.ctypes.data, b.ctypes.data, b.size, MINV[solver], C.byref(prm), cb, ep, C.byref(tx.v), C.byref(ci)),
           "minv_unpreconditioned")
    return x, _info(ci, tx.text())


def minv_vector_cg(*a, **k):
    return solve("cg", *a, **k)


def minv_vector_bicgstab(*a, **k):
    return solve("bicgstab", *a, **k)


def minv_vector_gcr(*a, **k):
    return solve("gcr", *a, **k)


def minv_vector_cg_m(op: str, b, shifts, *, links=None, mass=0.0, lx=0, ly=0, nc=1, resid_freq_check=10,
                     max_iter=10000, eps=1e-10, worst_first=False, verb_level=0, prefix="", fmt="%g"):
    """Multishift CG.  Returns (X[n_shift, size], Info with resSqmrhs, shifts)."""
    real = op == "laplace_real"
    dt = np.float64 if real else np.complex128
    b = np.ascontiguousarray(b, dtype=dt)
    sh = np.array(shifts, dtype=np.float64, copy=True)
    n = sh.size
    X = np.zeros((n, b.size), dtype=dt)
    res = np.zeros(n)
    ex = _Extra(op, links, mass, lx, ly, nc)
    tx = _Text(verb_level, prefix, fmt)
    ci = InversionInfo()
    ci.resSqmrhs = res.ctypes.data
    ci.n_rhs = n
    ptrs = (C.c_void_p * n)(*[X[k].ctypes.data for k in range(n)])
    f = lib().glx_minv_vector_cg_m_r if real else lib().glx_minv_vector_cg_m_c
    _check(f(ptrs, b.ctypes.data, n, b.size, resid_freq_check, max_iter, eps, sh.ctypes.data, op_callback(op), ex.ptr,
             int(worst_first), C.byref(tx.v), C.byref(ci)), "minv_vector_cg_m")
    return X, _info(ci, tx.text(), res), sh


def register_operator(callback_addr: int, op: str, lx=0, ly=0, mass=0.0, nc=1, links=None):
    """Bind a host callback to a device operator (glx_register_operator)."""
    st = None
    keep = None
    if lx > 0:
        keep = None if links is None else np.ascontiguousarray(links, dtype=np.complex128)
        st = StaggeredU1Op(None if keep is None else keep.ctypes.data, mass, lx, ly, nc, 0.0)
    _check(lib().glx_register_operator(callback_addr, OP_KINDS[op], None if st is None else C.byref(st)),
           "glx_register_operator")
    return keep


def unregister_operator(callback_addr: int):
    _check(lib().glx_unregister_operator(callback_addr), "glx_unregister_operator")


# ---------------------------------------------------------------- device --
class DeviceArray:
    """A device buffer holding a numpy array's bytes."""

    def __init__(self, arr=None, *, nbytes: int = 0, dtype=np.complex128, shape=None):
        self.dtype = np.dtype(dtype if arr is None else np.asarray(arr).dtype)
        self.shape = shape if arr is None else np.asarray(arr).shape
        self.nbytes = nbytes if arr is None else np.asarray(arr).nbytes
        if self.shape is None:
            self.shape = (self.nbytes // self.dtype.itemsize,)
        if int(np.prod(self.shape)) * self.dtype.itemsize != self.nbytes:
            raise ValueError("DeviceArray: shape and dtype do not cover nbytes")
        p = C.c_void_p()
        _check(lib().glx_dev_alloc(C.byref(p), max(self.nbytes, 1)), "glx_dev_alloc")
        self.ptr = p.value
        if arr is not None:
            self.upload(arr)

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        assert a.nbytes == self.nbytes
        _check(lib().glx_memcpy_htod(self.ptr, a.ctypes.data, a.nbytes), "glx_memcpy_htod")

    def download(self):
        out = np.empty(self.shape, dtype=self.dtype)
        assert out.nbytes == self.nbytes
        _check(lib().glx_memcpy_dtoh(out.ctypes.data, self.ptr, self.nbytes), "glx_memcpy_dtoh")
        return out

    def free(self):
        if self.ptr:
            lib().glx_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Lattice:
    """A gauge field resident in HBM (FP64 + optional FP32 / FP16 copies)."""

    def __init__(self, links, lx: int, ly: int, f32: bool = True, f16: bool = True):
        self.lx, self.ly = lx, ly
        self.links = np.ascontiguousarray(links, dtype=np.complex128)
        h = C.c_void_p()
        flags = (LINKS_F32 if f32 else 0) | (LINKS_F16 if f16 else 0)
        _check(lib().glx_lattice_create(C.byref(h), self.links.ctypes.data, lx, ly, flags), "glx_lattice_create")
        self.h = h.value

    @classmethod
    def gaussian(cls, lx: int, ly: int, beta: float, seed: int, stream: int = 0, f32: bool = True,
                 f16: bool = True):
        """Links generated in HBM: theta ~ N(0, 1/|beta|) (uniform for beta = 0)
        from Philox4x32-10 with counter {site, stream, 0, 0} and key seed."""
        self = cls.__new__(cls)
        self.lx, self.ly, self.links = lx, ly, None
        h = C.c_void_p()
        flags = (LINKS_F32 if f32 else 0) | (LINKS_F16 if f16 else 0)
        _check(lib().glx_lattice_create_gaussian(C.byref(h), lx, ly, flags, float(beta), int(seed), int(stream)),
               "glx_lattice_create_gaussian")
        self.h = h.value
        return self

    def regenerate(self, beta: float, seed: int, stream: int = 0):
        self.links = None
        _check(lib().glx_lattice_gaussian_links(self.h, float(beta), int(seed), int(stream)),
               "glx_lattice_gaussian_links")

    def get_links(self):
        out = np.empty(2 * self.lx * self.ly, np.complex128)
        _check(lib().glx_lattice_get_links(self.h, out.ctypes.data), "glx_lattice_get_links")
        return out

    def plaquette(self) -> complex:
        out = np.zeros(2)
        _check(lib().glx_lattice_plaquette(self.h, out.ctypes.data), "glx_lattice_plaquette")
        return complex(out[0], out[1])

    def update(self, links):
        self.links = np.ascontiguousarray(links, dtype=np.complex128)
        _check(lib().glx_lattice_update_links(self.h, self.links.ctypes.data), "glx_lattice_update_links")

    def close(self):
        if getattr(self, "h", None):
            lib().glx_lattice_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def dev_solve(solver: str, lat, op: str, mass: float, x: DeviceArray, b: DeviceArray, max_iter=10000, eps=1e-10,
              nc=1, verb_level=0, prefix="", fmt="%g") -> Info:
    tx = _Text(verb_level, prefix, fmt)
    ci = InversionInfo()
    _check(lib().glx_dev_solve(SOLVER_IDS[solver], None if lat is None else lat.h, OP_KINDS[op], mass, nc, x.ptr,
                               b.ptr, max_iter, eps, C.byref(tx.v), C.byref(ci)), "glx_dev_solve")
    return _info(ci, tx.text())


def dev_cg_mixed(lat, op: str, mass: float, x: DeviceArray, b: DeviceArray, max_iter=100000, eps=1e-10, delta=0.1,
                 link_prec=LINKS_F32, verb_level=0) -> Info:
    tx = _Text(verb_level, "", "%g")
    ci = InversionInfo()
    _check(lib().glx_dev_cg_mixed(lat.h, OP_KINDS[op], mass, x.ptr, b.ptr, max_iter, eps, delta, link_prec,
                                  C.byref(tx.v), C.byref(ci)), "glx_dev_cg_mixed")
    return _info(ci, tx.text())


def dev_cg_mixed_noalloc(lat, op: str, mass: float, x: DeviceArray, b: DeviceArray, max_iter=100000, eps=1e-10,
                         delta=0.1, link_prec=LINKS_F32, verb_level=0) -> Info:
    """dev_cg_mixed without device allocation (glx_dev_cg_mixed_noalloc):
    the library's fixed scratch array, at most 2^20 sites."""
    tx = _Text(verb_level, "", "%g")
    ci = InversionInfo()
    _check(lib().glx_dev_cg_mixed_noalloc(lat.h, OP_KINDS[op], mass, x.ptr, b.ptr, max_iter, eps, delta, link_prec,
                                          C.byref(tx.v), C.byref(ci)), "glx_dev_cg_mixed_noalloc")
    return _info(ci, tx.text())


def dev_cg_fused(lat, op: str, mass: float, x: DeviceArray, b: DeviceArray, max_iter=100000, eps=1e-10,
                 verb_level=0) -> Info:
    tx = _Text(verb_level, "", "%g")
    ci = InversionInfo()
    _check(lib().glx_dev_cg_fused(lat.h, OP_KINDS[op], mass, x.ptr, b.ptr, max_iter, eps, C.byref(tx.v),
                                  C.byref(ci)), "glx_dev_cg_fused")
    return _info(ci, tx.text())


@dataclass
class Grad:
    grad: np.ndarray      # dS/dtheta, 2V doubles in the links' layout
    dS_dm: float
    S: float              # Re phi^dag chi
    S_robust: float       # 2 Re phi^dag chi - |D chi|^2
    info: "Info"


def staggered_grad(lat, mass: float, phi: DeviceArray, precision: int = 0, max_iter: int = 100000,
                   eps: float = 1e-10, grad: DeviceArray | None = None, chi: DeviceArray | None = None,
                   download: bool = True) -> Grad:
    """Gradient of S = phi^dag (D^dag D)^-1 phi with respect to the link angles
    and the mass (precision 0 FP64; LINKS_F32 / LINKS_F16 mixed-precision
    solve).  Runs on the device; the gradient is downloaded unless
    download=False (then Grad.grad is None and `grad` holds it)."""
    V = lat.lx * lat.ly
    g = grad if grad is not None else DeviceArray(nbytes=2 * V * 8, dtype=np.float64, shape=(2 * V,))
    dm, S, Sr = C.c_double(0.0), C.c_double(0.0), C.c_double(0.0)
    ci = InversionInfo()
    _check(lib().glx_dev_staggered_grad(lat.h, mass, phi.ptr, precision, max_iter, eps, g.ptr,
                                        None if chi is None else chi.ptr, C.byref(dm), C.byref(S), C.byref(Sr),
                                        C.byref(ci)), "glx_dev_staggered_grad")
    return Grad(g.download() if download else None, dm.value, S.value, Sr.value, _info(ci, ""))


def set_tuning(group: int = 16, use_graph: bool = False):
    _check(lib().glx_set_tuning(group, int(use_graph)), "glx_set_tuning")


def set_sstep(s: int = 2, rows: int = 0):
    """Mixed-precision CG on D^dag D: s = 2 (default) runs two CG iterations
    per launch (s-step form) where the lattice allows, s = 1 one; rows per
    block of the s-step kernel for every size, or 0 (default) per size."""
    _check(lib().glx_set_sstep(s, rows), "glx_set_sstep")


def set_sstep_rows(lx: int, ly: int, rows: int):
    """Per-size entry: s-step rows per block, 0 = one-iteration form, -1 = remove."""
    _check(lib().glx_set_sstep_rows(lx, ly, rows), "glx_set_sstep_rows")


def set_sstep_form(form: int = -1, band_rows: int = 0):
    """The s-step launch on the full-lattice D^dag D: form 0 blocks of four
    waves marching a strip together through LDS rings, form 1 every wave
    marching a strip of its own in registers (band_rows rows per wave, 0: by
    lattice size); -1 (default): per lattice size."""
    _check(lib().glx_set_sstep_form(form, band_rows), "glx_set_sstep_form")


def get_sstep_form() -> int:
    f = C.c_int(0)
    _check(lib().glx_get_sstep_form(C.byref(f)), "glx_get_sstep_form")
    return f.value


def set_sstep_form_entry(lx: int, ly: int, form: int):
    """Per-size entry of the s-step form (0 or 1), -1 = remove."""
    _check(lib().glx_set_sstep_form_entry(lx, ly, form), "glx_set_sstep_form_entry")


def get_sstep_form_for(lx: int, ly: int) -> int:
    f = C.c_int(0)
    _check(lib().glx_get_sstep_form_for(lx, ly, C.byref(f)), "glx_get_sstep_form_for")
    return f.value


def get_sstep_rows(lx: int, ly: int) -> int:
    r = C.c_int(0)
    _check(lib().glx_get_sstep_rows(lx, ly, C.byref(r)), "glx_get_sstep_rows")
    return r.value


def autotune_sstep(lat, reps: int = 40) -> dict:
    """Time the mixed CG on this lattice in one-iteration form, in s-step form
    0 at 8 / 16 / 32 / 64 / 128 rows per block (and 256 where ly >= 2048) and
    in s-step form 1 (wave strips); install and return the fastest (rows 0:
    the one-iteration form; form: the s-step launch's form)."""
    r, ms = C.c_int(0), C.c_double(0.0)
    _check(lib().glx_autotune_sstep(lat.h, reps, C.byref(r), C.byref(ms)), "glx_autotune_sstep")
    form = get_sstep_form_for(lat.lx, lat.ly) if hasattr(lib(), "glx_get_sstep_form_for") else 0
    return {"rows": r.value, "form": form, "ms_per_iteration": ms.value}


def get_sstep():
    s, rows = C.c_int(), C.c_int()
    _check(lib().glx_get_sstep(C.byref(s), C.byref(rows)), "glx_get_sstep")
    return s.value, rows.value


def set_dwave_rows(rows: int = -1):
    """Rows per wave of the wave-march kernel of the plain one-hop staggered
    kinds (D, D^dag, gamma5 D, D_eo, D_oe)
    (lattices with lx % 64 == 0, ly >= 4): -1 by lattice size (the default),
    0 the tiled kernel always, > 0 that many."""
    _check(lib().glx_set_dwave_rows(rows), "glx_set_dwave_rows")


def set_ddwave_rows(rows: int = -1):
    """Rows per wave of the wave-march kernel of the plain D^dag D: -1 by
    lattice size (the default), 0 the row-marching or tiled kernel always,
    > 0 that many."""
    _check(lib().glx_set_ddwave_rows(rows), "glx_set_ddwave_rows")


def get_ddwave_rows(lx: int, ly: int, esize: int = 8) -> int:
    """Rows per wave a plain lx x ly D^dag D apply would use (0: another kernel)."""
    r = C.c_int(0)
    _check(lib().glx_get_ddwave_rows(lx, ly, esize, C.byref(r)), "glx_get_ddwave_rows")
    return r.value


def get_dwave_rows(lx: int, ly: int, esize: int = 8) -> int:
    """Rows per wave an lx x ly apply of D or D^dag with vector element size
    esize (8 complex FP32, 16 complex FP64) would use (0: tiled kernel)."""
    r = C.c_int(0)
    _check(lib().glx_get_dwave_rows(lx, ly, esize, C.byref(r)), "glx_get_dwave_rows")
    return r.value


def set_normal_rows(lx: int, ly: int, esize: int, rows: int):
    """Rows per block of the row-marching D^dag D kernel for an lx x ly
    lattice and element size (16 complex FP64, 8 complex FP32); 0 = tiled
    kernel; lx = ly = 0 sets the default (-1, the initial one: the fewest of
    16/32/64/128 rows that keep the grid within 512 blocks, else 256)."""
    _check(lib().glx_set_normal_rows(lx, ly, esize, rows), "glx_set_normal_rows")


def get_normal_rows(lx: int, ly: int, esize: int) -> int:
    r = C.c_int(0)
    _check(lib().glx_get_normal_rows(lx, ly, esize, C.byref(r)), "glx_get_normal_rows")
    return r.value


def autotune_normal(lat, reps: int = 20, cache: str | None = None) -> dict:
    """Pick the fastest D^dag D launch shape for this lattice (FP64 and FP32)
    and install it.  With `cache`, a JSON file of earlier results is read
    first (an entry for this size is installed without timing) and updated."""
    import json
    key = f"{lat.lx}x{lat.ly}"
    table = {}
    if cache and os.path.exists(cache):
        with open(cache) as f:
            table = json.load(f)
    if key in table:
        e = table[key]
        set_normal_rows(lat.lx, lat.ly, 16, int(e["rows64"]))
        if e.get("rows32") is not None:
            set_normal_rows(lat.lx, lat.ly, 8, int(e["rows32"]))
        return dict(e, cached=True)
    r64, r32 = C.c_int(-1), C.c_int(-1)
    m64, m32 = C.c_double(0.0), C.c_double(0.0)
    _check(lib().glx_autotune_normal(lat.h, reps, C.byref(r64), C.byref(r32), C.byref(m64), C.byref(m32)),
           "glx_autotune_normal")
    e = {"rows64": r64.value, "ms64": m64.value,
         "rows32": r32.value if r32.value >= 0 else None, "ms32": m32.value if r32.value >= 0 else None}
    if cache:
        table[key] = e
        tmp = cache + ".tmp"
        with open(tmp, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
        os.replace(tmp, cache)
    return dict(e, cached=False)


def dev_cg_m(lat, op: str, mass: float, xs, b: DeviceArray, shifts, resid_freq_check=10, max_iter=10000, eps=1e-10,
             worst_first=False, nc=1) -> Info:
    sh = np.ascontiguousarray(shifts, dtype=np.float64)
    n = sh.size
    res = np.zeros(n)
    ci = InversionInfo()
    ci.resSqmrhs = res.ctypes.data
    ci.n_rhs = n
    ptrs = (C.c_void_p * n)(*[x.ptr for x in xs])
    _check(lib().glx_dev_cg_m(None if lat is None else lat.h, OP_KINDS[op], mass, nc, ptrs, b.ptr, n, sh.ctypes.data,
                              resid_freq_check, max_iter, eps, int(worst_first), None, C.byref(ci)), "glx_dev_cg_m")
    return _info(ci, "", res)


EO_METHODS = {"cg": 0, "mixed": LINKS_F32, "mixed_f16": LINKS_F16, "cg_fused": 3}


def dev_solve_staggered_eo(lat, mass: float, x: DeviceArray, b: DeviceArray, method: str = "cg",
                           max_iter: int = 100000, eps: float = 1e-10):
    """D x = b by the even/odd pipeline on compacted vectors; returns
    (Info of the even-site solve, |D x - b|^2)."""
    r2 = C.c_double(0.0)
    ci = InversionInfo()
    _check(lib().glx_dev_solve_staggered_eo(lat.h, mass, x.ptr, b.ptr, EO_METHODS[method], max_iter, eps,
                                            C.byref(r2), C.byref(ci)), "glx_dev_solve_staggered_eo")
    return _info(ci), r2.value


def dev_eo_prepare(lat, mass: float, be: DeviceArray, b: DeviceArray):
    _check(lib().glx_dev_eo_prepare(lat.h, mass, be.ptr, b.ptr), "glx_dev_eo_prepare")


def dev_eo_reconstruct(lat, mass: float, x: DeviceArray, xe: DeviceArray, b: DeviceArray):
    _check(lib().glx_dev_eo_reconstruct(lat.h, mass, x.ptr, xe.ptr, b.ptr), "glx_dev_eo_reconstruct")


class BatchSystem(C.Structure):
    _fields_ = [("lat", C.c_void_p), ("mass", C.c_double), ("b_dev", C.c_void_p), ("x_dev", C.c_void_p)]


class BatchStats(C.Structure):
    _fields_ = [("bytes_stencil", C.c_double), ("bytes_r_pass", C.c_double), ("bytes_shift_pass", C.c_double),
                ("ms_stencil", C.c_double), ("ms_r_pass", C.c_double), ("ms_shift_pass", C.c_double),
                ("iterations", C.c_longlong), ("shift_updates_x", C.c_longlong), ("shift_updates_p", C.c_longlong)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


def dev_cg_m_batched(systems, shifts, resid_freq_check=10, max_iter=10000, eps=1e-10, worst_first=False,
                     use_graph=True, stats=False, time_phases=False):
    """Batched multishift CG on the staggered normal operator: systems is a
    list of (lattice, mass, b: DeviceArray, xs: list of DeviceArray), all with
    the same shifts.  Returns one Info per system (resSqmrhs per shift); with
    stats, (infos, dict of the per-phase algorithmic bytes and, with
    time_phases, HIP-event times) from glx_dev_cg_m_batched_ex."""
    sh = np.ascontiguousarray(shifts, dtype=np.float64)
    n = sh.size
    ns = len(systems)
    arr = (BatchSystem * max(ns, 1))()
    keep = []
    for s, (lat, mass, b, xs) in enumerate(systems):
        if len(xs) != n:
            raise GlxError("each system needs one solution vector per shift")
        ptrs = (C.c_void_p * n)(*[x.ptr for x in xs])
        keep.append(ptrs)
        arr[s] = BatchSystem(lat.h, mass, b.ptr, C.cast(ptrs, C.c_void_p))
    infos = (InversionInfo * max(ns, 1))()
    res = [np.zeros(n) for _ in range(ns)]
    for s in range(ns):
        infos[s].resSqmrhs = res[s].ctypes.data
        infos[s].n_rhs = n
    if stats or time_phases:
        st = BatchStats()
        _check(lib().glx_dev_cg_m_batched_ex(ns, C.cast(arr, C.c_void_p), n, sh.ctypes.data, resid_freq_check,
                                             max_iter, eps, int(worst_first), int(use_graph), int(time_phases),
                                             C.cast(infos, C.c_void_p), C.byref(st)), "glx_dev_cg_m_batched_ex")
        return [_info(infos[s], "", res[s]) for s in range(ns)], st.as_dict()
    _check(lib().glx_dev_cg_m_batched(ns, C.cast(arr, C.c_void_p), n, sh.ctypes.data, resid_freq_check, max_iter,
                                      eps, int(worst_first), int(use_graph), C.cast(infos, C.c_void_p)),
           "glx_dev_cg_m_batched")
    return [_info(infos[s], "", res[s]) for s in range(ns)]


def dev_apply(lat, op: str, mass: float, out: DeviceArray, inp: DeviceArray, fp32=False, link_prec=LINKS_F32, nc=1):
    _check(lib().glx_dev_apply(None if lat is None else lat.h, OP_KINDS[op], mass, nc, int(fp32), link_prec, out.ptr,
                               inp.ptr), "glx_dev_apply")


def time_apply(lat, op: str, mass: float, out: DeviceArray, inp: DeviceArray, reps=100, fp32=False,
               link_prec=LINKS_F32) -> float:
    ms = C.c_double(0.0)
    _check(lib().glx_dev_time_apply(lat.h, OP_KINDS[op], mass, int(fp32), link_prec, out.ptr, inp.ptr, reps,
                                    C.byref(ms)), "glx_dev_time_apply")
    return ms.value


STREAM_MODES = {"read": 0, "write": 1, "copy": 2, "read3_write1": 3}


def time_stream(mode: str, nbytes: int, reps: int = 20):
    """(ms per pass, bytes per pass) of a plain streaming kernel over a
    footprint of nbytes (glx_dev_time_stream)."""
    ms, by = C.c_double(0.0), C.c_double(0.0)
    _check(lib().glx_dev_time_stream(STREAM_MODES[mode], nbytes, reps, C.byref(ms), C.byref(by)),
           "glx_dev_time_stream")
    return ms.value, by.value


def time_cg_kernel(lat, mass: float, prec: int = 0, reps: int = 100):
    """(ms per launch, algorithmic bytes per site) of the single-pass CG's
    fused iteration kernel: prec 0 FP64, LINKS_F32 / LINKS_F16 mixed FP32;
    3 / 4 the mixed solver's s-step kernel (two iterations per launch) with
    FP32 / FP16 links."""
    ms, by = C.c_double(0.0), C.c_double(0.0)
    _check(lib().glx_dev_time_cg_kernel(lat.h, mass, prec, reps, C.byref(ms), C.byref(by)),
           "glx_dev_time_cg_kernel")
    return ms.value, by.value


def time_sscg_traffic(lat, link_prec: int = LINKS_F32, reps: int = 100) -> float:
    """ms per launch of the wave-strip s-step launch's memory traffic alone
    (the same loads and stores without the arithmetic; glx_dev_time_sscg_traffic)."""
    ms = C.c_double(0.0)
    _check(lib().glx_dev_time_sscg_traffic(lat.h, link_prec, reps, C.byref(ms)), "glx_dev_time_sscg_traffic")
    return ms.value


def copy_device(dst: DeviceArray, src: DeviceArray):
    """dst = src on the device (hipMemcpy device to device; sizes must match)."""
    if dst.nbytes != src.nbytes:
        raise GlxError("copy_device: sizes differ")
    _check(lib().glx_memcpy_dtod(dst.ptr, src.ptr, src.nbytes), "glx_memcpy_dtod")


def synchronize():
    _check(lib().glx_device_synchronize(), "glx_device_synchronize")


def release_cache():
    _check(lib().glx_release_cache(), "glx_release_cache")


# ----------------------------------------------------- general stencil_2d --
CST_KINDS = {"all": 18, "eo": 19, "oe": 20, "tb": 21, "bt": 22, "m2mdeodoe": 23}
CST_CALLBACKS = {"all": "glx_apply_stencil_2d_nc", "eo": "glx_apply_stencil_2d_nc_eo",
                 "oe": "glx_apply_stencil_2d_nc_oe", "tb": "glx_apply_stencil_2d_nc_tb",
                 "bt": "glx_apply_stencil_2d_nc_bt", "m2mdeodoe": "glx_apply_stencil_2d_nc_m2mdeodoe"}


class StencilNc:
    """stencil_2d for any nc (coarse_stencil.h:33-154): coefficient arrays in
    the reference layout (clover V nc^2, hopping 4 V nc^2, two_link 8 V nc^2 or
    None; element (term, site, row, column) at ((term V + site) nc + row) nc +
    column), complex shifts and the stencil_dir `sdir`."""

    def __init__(self, lx, ly, nc, clover, hopping, two_link=None, shift=0j, eo_shift=0j, dof_shift=0j, sdir=0):
        self.lx, self.ly, self.nc, self.sdir = lx, ly, nc, int(sdir)
        self.clover = np.ascontiguousarray(clover, dtype=np.complex128)
        self.hopping = np.ascontiguousarray(hopping, dtype=np.complex128)
        self.two_link = None if two_link is None else np.ascontiguousarray(two_link, dtype=np.complex128)
        self.shift, self.eo_shift, self.dof_shift = complex(shift), complex(eo_shift), complex(dof_shift)
        blk = lx * ly * nc * nc
        if self.clover.size != blk or self.hopping.size != 4 * blk or \
                (self.two_link is not None and self.two_link.size != 8 * blk):
            raise ValueError("coefficient arrays do not match lx * ly * nc^2 blocks")

    @property
    def n(self):
        return self.lx * self.ly * self.nc

    def struct(self) -> Stencil2DNc:
        st = Stencil2DNc()
        st.lx, st.ly, st.nc, st.sdir = self.lx, self.ly, self.nc, self.sdir
        st.has_two = int(self.two_link is not None)
        st.clover, st.hopping = self.clover.ctypes.data, self.hopping.ctypes.data
        st.two_link = None if self.two_link is None else self.two_link.ctypes.data
        st.shift[:] = [self.shift.real, self.shift.imag]
        st.eo_shift[:] = [self.eo_shift.real, self.eo_shift.imag]
        st.dof_shift[:] = [self.dof_shift.real, self.dof_shift.imag]
        return st


def cstencil_apply(st: StencilNc, rhs, piece="all", lhs=None):
    """The callback form (host vectors; the GPU does the work).  `lhs` is the
    initial content of the output, which a two-link sdir on a stencil without
    two-link terms leaves as it is."""
    rhs = np.ascontiguousarray(rhs, dtype=np.complex128)
    out = np.zeros_like(rhs) if lhs is None else np.array(lhs, dtype=np.complex128, copy=True)
    cs = st.struct()
    fn = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.c_void_p)(op_callback_addr(CST_CALLBACKS[piece]))
    fn(out.ctypes.data, rhs.ctypes.data, C.cast(C.pointer(cs), C.c_void_p))
    return out


def op_callback_addr(symbol: str) -> int:
    return C.cast(getattr(lib(), symbol), C.c_void_p).value


def cstencil_eoprec_prepare(st: StencilNc, rhs):
    rhs = np.ascontiguousarray(rhs, dtype=np.complex128)
    out = np.zeros_like(rhs)
    cs = st.struct()
    lib().glx_stencil_2d_nc_eoprec_prepare(out.ctypes.data, rhs.ctypes.data, C.byref(cs))
    return out


def cstencil_eoprec_reconstruct(st: StencilNc, lhs_e, rhs_o):
    lhs_e = np.ascontiguousarray(lhs_e, dtype=np.complex128)
    rhs_o = np.ascontiguousarray(rhs_o, dtype=np.complex128)
    out = np.zeros_like(lhs_e)
    cs = st.struct()
    lib().glx_stencil_2d_nc_eoprec_reconstruct(out.ctypes.data, lhs_e.ctypes.data, rhs_o.ctypes.data, C.byref(cs))
    return out


def cstencil_generate(lx, ly, nc, has_two, callback: int, extra) -> StencilNc:
    """generate_stencil_2d: probe the operator at address `callback` (with
    extra_info pointer `extra`) with unit vectors."""
    blk = lx * ly * nc * nc
    out = StencilNc(lx, ly, nc, np.zeros(blk, np.complex128), np.zeros(4 * blk, np.complex128),
                    np.zeros(8 * blk, np.complex128) if has_two else None)
    cs = out.struct()
    _check(lib().glx_generate_stencil_2d_nc(C.byref(cs), callback, extra), "glx_generate_stencil_2d_nc")
    return out


def cstencil_solve(solver: str, st: StencilNc, b, piece="all", x0=None, max_iter=10000, eps=1e-10, restart_freq=0):
    """minv_vector_* with a general-stencil callback: the solve runs on the GPU."""
    cs = st.struct()
    return solve(solver, "stencil", b, x0, max_iter=max_iter, eps=eps, restart_freq=restart_freq,
                 callback=op_callback_addr(CST_CALLBACKS[piece]), extra=C.cast(C.pointer(cs), C.c_void_p).value)


class CStencilDev:
    """A general stencil resident in HBM (glx_cstencil_create)."""

    def __init__(self, st: StencilNc):
        self.lx, self.ly, self.nc, self.sdir = st.lx, st.ly, st.nc, st.sdir
        h = C.c_void_p()
        cs = st.struct()
        _check(lib().glx_cstencil_create(C.byref(h), C.byref(cs)), "glx_cstencil_create")
        self.h = h.value

    def apply(self, out: DeviceArray, inp: DeviceArray, piece="all", sdir=None):
        _check(lib().glx_dev_cstencil_apply(self.h, CST_KINDS[piece], self.sdir if sdir is None else int(sdir),
                                            out.ptr, inp.ptr), "glx_dev_cstencil_apply")

    def solve(self, solver: str, x: DeviceArray, b: DeviceArray, piece="all", max_iter=10000, eps=1e-10):
        ci = InversionInfo()
        _check(lib().glx_dev_cstencil_solve(SOLVER_IDS[solver], self.h, CST_KINDS[piece], x.ptr, b.ptr, max_iter, eps,
                                            None, C.byref(ci)), "glx_dev_cstencil_solve")
        return _info(ci)

    def time(self, out: DeviceArray, inp: DeviceArray, piece="all", reps=100) -> float:
        ms = C.c_double()
        _check(lib().glx_dev_time_cstencil(self.h, CST_KINDS[piece], out.ptr, inp.ptr, reps, C.byref(ms)),
               "glx_dev_time_cstencil")
        return ms.value

    def close(self):
        if getattr(self, "h", None):
            lib().glx_cstencil_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
