"""The C ABI library without a GPU (CPU): it loads, exports every entry
point include/glx.h declares, reports a missing device as an error rather
than falling back to the CPU, and the reference-style C++ driver compiles
against the compat headers and links against it.  No compute calls."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "glx.h")
LIB = os.path.join(ROOT, "generic-linalg_amd", "libglx.so")


def declared_functions():
    text = open(HDR).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(glx_\w+)\s*\(", text)) - {"glx_matvec"})


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build with `make -C generic-linalg_amd`"
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (\w+)", out))
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing, missing
    assert len(declared_functions()) > 50


def test_cpp_dropin_symbols_exported():
    out = subprocess.run(["nm", "-D", "-C", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    for sig in ["minv_vector_cg(std::complex<double>*", "minv_vector_cg_m(double**", "minv_vector_bicgstab_restart(",
                "minv_vector_gcr(double*", "square_staggered_normal_u1(std::complex<double>*",
                "square_staggered_eoprec_reconstruct(", "square_laplace(double*"]:
        assert sig in out, sig


def test_loads_and_reports_no_device():
    lib = C.CDLL(LIB)
    lib.glx_abi_version.restype = C.c_int
    assert lib.glx_abi_version() == 1
    lib.glx_last_error.restype = C.c_char_p
    n = C.c_int(-1)
    rc = lib.glx_device_count(C.byref(n))
    if rc != 0:  # no GPU here: an error, never a silent CPU path
        assert rc == 2 and b"device" in lib.glx_last_error().lower()


def test_argument_errors_without_device():
    """Argument checks run before any device work."""
    lib = C.CDLL(LIB)
    lib.glx_last_error.restype = C.c_char_p
    f = lib.glx_minv_vector_cg_restart_c
    f.restype = C.c_int
    vp = C.c_void_p
    f.argtypes = [vp, vp, C.c_int, C.c_int, C.c_double, C.c_int, vp, vp, vp, vp]
    rc = f(None, None, 4, 10, 1e-10, 0, None, None, None, None)
    assert rc == 1 and b"restart_freq" in lib.glx_last_error()
    g = lib.glx_register_operator
    g.restype = C.c_int
    g.argtypes = [vp, C.c_int, vp]
    assert g(None, 3, None) == 1
    # an in-place operator application is refused (every kernel reads sites another wave writes)
    a = lib.glx_dev_apply
    a.restype = C.c_int
    a.argtypes = [vp, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, vp, vp]
    buf = C.create_string_buffer(64)
    assert a(None, 3, 0.1, 1, 0, 1, C.addressof(buf), C.addressof(buf)) == 1
    assert b"different buffers" in lib.glx_last_error()


class _Params(C.Structure):
    _fields_ = [("tol", C.c_double), ("max_iters", C.c_int), ("restart", C.c_int), ("restart_freq", C.c_int),
                ("sor_omega", C.c_double), ("minres_omega", C.c_double), ("bicgstabl_l", C.c_int)]


class _Info(C.Structure):
    _fields_ = [("resSq", C.c_double), ("iter", C.c_int), ("success", C.c_int), ("ops_count", C.c_int),
                ("name", C.c_char * 32), ("resSqmrhs", C.c_void_p), ("n_rhs", C.c_int),
                ("reliable_updates", C.c_int)]


def test_dispatcher_without_device():
    """minv_unpreconditioned (generic_inverter.cpp:18-188): MINV_INVALID
    returns a default info as the reference does; solvers outside this
    library are refused with a message, before any device work."""
    lib = C.CDLL(LIB)
    lib.glx_last_error.restype = C.c_char_p
    vp = C.c_void_p
    for f in (lib.glx_minv_unpreconditioned_r, lib.glx_minv_unpreconditioned_c):
        f.restype = C.c_int
        f.argtypes = [vp, vp, C.c_int, C.c_int, C.POINTER(_Params), vp, vp, vp, C.POINTER(_Info)]
        p = _Params(1e-10, 100, 0, 10, 1.0, 1.0, 2)
        info = _Info()
        info.iter = 77
        assert f(None, None, 4, -1, C.byref(p), None, None, None, C.byref(info)) == 0
        assert info.iter == 0 and info.success == 0 and info.name == b""
        for t in (1, 4, 5, 6, 7):
            assert f(None, None, 4, t, C.byref(p), None, None, None, C.byref(info)) == 1
            assert b"not part of this library" in lib.glx_last_error()
        assert f(None, None, 4, 0, None, None, None, None, C.byref(info)) == 1


def test_dropin_driver_builds(tmp_path):
    d = os.path.join(ROOT, "tests", "dropin")
    out = subprocess.run(["make", "-C", d, "-B"], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    for exe in ("square_laplace_driver", "u1_driver", "stencil_driver", "c_abi_check"):
        assert os.path.exists(os.path.join(d, exe)), exe


def test_header_is_plain_c():
    """include/glx.h compiles as C99 with -pedantic -Werror (c_abi_check.c)
    and the program runs without a device."""
    d = os.path.join(ROOT, "tests", "dropin")
    subprocess.run(["make", "-s", "-C", d, "c_abi_check"], check=True)
    out = subprocess.run([os.path.join(d, "c_abi_check")], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert re.match(r"abi \d+ 1 1 1", out.stdout)


def test_python_binding_imports():
    import glx
    assert glx.lib() is not None
    assert set(glx.OP_KINDS) >= {"staggered", "staggered_normal", "laplace_real", "m2mdeodoe"}


def test_normal_rows_default_without_device():
    """Rows per block of the row-marching D^dag D kernels for sizes not in the
    table: the fewest of 16/32/64/128 that keep the grid within 512 blocks
    (64 lanes of 1 FP64 or 2 FP32 sites), else 256 -- the shapes the kernel
    timings favour at 1024^2, 2048^2 and 4096^2; an explicit default
    overrides it and -1 restores it (no device work)."""
    lib = C.CDLL(LIB)
    lib.glx_last_error.restype = C.c_char_p
    r = C.c_int()
    assert lib.glx_get_normal_rows(0, 0, 16, C.byref(r)) == 0 and r.value == -1
    want = {(256, 256, 16): 16, (512, 512, 16): 16, (1024, 1024, 16): 32, (1024, 1024, 8): 16,
            (2048, 2048, 16): 128, (2048, 2048, 8): 64, (4096, 4096, 16): 256, (4096, 4096, 8): 256,
            (1024, 256, 16): 16, (256, 4096, 16): 32}
    for (lx, ly, es), rows in want.items():
        # 512 blocks: lx / (64 cpl) strips x ceil(ly / rows) bands
        assert (lx // (64 * (1 if es == 16 else 2))) * -(-ly // rows) <= 512 or rows == 256
        assert lib.glx_get_normal_rows(lx, ly, es, C.byref(r)) == 0 and r.value == rows, (lx, ly, es, r.value)
    assert lib.glx_set_normal_rows(0, 0, 0, 16) == 0
    assert lib.glx_get_normal_rows(2048, 2048, 16, C.byref(r)) == 0 and r.value == 16
    assert lib.glx_set_normal_rows(2048, 2048, 16, -2) == 1 and b"rows must be" in lib.glx_last_error()
    assert lib.glx_set_normal_rows(0, 0, 0, -1) == 0
    assert lib.glx_get_normal_rows(2048, 2048, 16, C.byref(r)) == 0 and r.value == 128
    # a size's entry overrides the default until -1 removes it
    assert lib.glx_set_normal_rows(2048, 2048, 16, 8) == 0
    assert lib.glx_get_normal_rows(2048, 2048, 16, C.byref(r)) == 0 and r.value == 8
    assert lib.glx_set_normal_rows(2048, 2048, 16, -1) == 0
    assert lib.glx_get_normal_rows(2048, 2048, 16, C.byref(r)) == 0 and r.value == 128


def test_dwave_rows_setting_without_device():
    """glx_set_dwave_rows validates its argument; glx_get_dwave_rows gives
    the measured shapes (16 rows FP32; FP64 18, or 16 on large lattices;
    else about one wave per SIMD, at least 4 rows) and 0 (the tiled kernel)
    for widths that are not whole 64-column strips; likewise the D^dag D
    kernel, which leaves applications above 256 MiB to the row-marching
    kernel (no device work)."""
    lib = C.CDLL(LIB)
    lib.glx_last_error.restype = C.c_char_p
    r = C.c_int()
    want = {(1024, 1024, 8): 16, (1024, 1024, 16): 18, (2048, 2048, 8): 16, (2048, 2048, 16): 16,
            (4096, 4096, 16): 16,
            (512, 512, 8): 4, (768, 768, 8): 9, (64, 64, 16): 4, (96, 64, 8): 0, (64, 2, 8): 0,
            (16384, 16384, 8): 0}
    for (lx, ly, es), rows in want.items():
        assert lib.glx_get_dwave_rows(lx, ly, es, C.byref(r)) == 0 and r.value == rows, (lx, ly, es, r.value)
    assert lib.glx_set_dwave_rows(7) == 0
    assert lib.glx_get_dwave_rows(1024, 1024, 8, C.byref(r)) == 0 and r.value == 7
    assert lib.glx_set_dwave_rows(0) == 0
    assert lib.glx_get_dwave_rows(1024, 1024, 16, C.byref(r)) == 0 and r.value == 0
    assert lib.glx_set_dwave_rows(-2) == 1 and b"rows must be" in lib.glx_last_error()
    assert lib.glx_get_dwave_rows(0, 8, 8, C.byref(r)) == 1
    assert lib.glx_get_dwave_rows(64, 8, 4, C.byref(r)) == 1 and b"esize" in lib.glx_last_error()
    assert lib.glx_set_dwave_rows(-1) == 0
    # the D^dag D kernel: 60 owned columns per wave, 7 k - 4 rows where the lattice is large enough
    want = {(2048, 2048, 8): 24, (2048, 2048, 16): 24, (4096, 4096, 8): 0, (4096, 4096, 16): 0,
            (1024, 1024, 8): 10, (1024, 1024, 16): 10,
            (512, 512, 8): 5, (64, 64, 16): 3, (63, 64, 8): 0, (64, 3, 8): 0}
    for (lx, ly, es), rows in want.items():
        assert lib.glx_get_ddwave_rows(lx, ly, es, C.byref(r)) == 0 and r.value == rows, (lx, ly, es, r.value)
    assert lib.glx_set_ddwave_rows(0) == 0
    assert lib.glx_get_ddwave_rows(1024, 1024, 8, C.byref(r)) == 0 and r.value == 0
    assert lib.glx_set_ddwave_rows(-2) == 1 and b"rows must be" in lib.glx_last_error()
    assert lib.glx_set_ddwave_rows(-1) == 0


def test_sstep_setting_without_device():
    """glx_set_sstep validates its arguments and glx_get_sstep reads back the
    process default (no device work)."""
    lib = C.CDLL(LIB)
    lib.glx_last_error.restype = C.c_char_p
    s, rows = C.c_int(), C.c_int()
    assert lib.glx_get_sstep(C.byref(s), C.byref(rows)) == 0 and (s.value, rows.value) == (2, 0)
    assert lib.glx_set_sstep(3, 0) == 1 and b"s must be" in lib.glx_last_error()
    assert lib.glx_set_sstep(2, 6) == 1 and b"rows" in lib.glx_last_error()
    assert lib.glx_set_sstep(1, 32) == 0
    assert lib.glx_get_sstep(C.byref(s), C.byref(rows)) == 0 and (s.value, rows.value) == (1, 32)
    assert lib.glx_set_sstep(2, 0) == 0
    assert lib.glx_get_sstep(C.byref(s), C.byref(rows)) == 0 and (s.value, rows.value) == (2, 0)
    # per size: the automatic choice by block count, the table, removal
    r = C.c_int()
    for (lx, ly), want in {(256, 256): 8, (512, 512): 8, (768, 768): 16, (1024, 1024): 16, (2048, 2048): 64,
                           (4096, 4096): 256}.items():
        assert lib.glx_get_sstep_rows(lx, ly, C.byref(r)) == 0 and r.value == want, (lx, r.value)
    assert lib.glx_set_sstep_rows(1024, 1024, 0) == 0
    assert lib.glx_get_sstep_rows(1024, 1024, C.byref(r)) == 0 and r.value == 0
    assert lib.glx_set_sstep_rows(1024, 1024, 6) == 1
    assert lib.glx_set_sstep_rows(1024, 1024, -1) == 0
    assert lib.glx_get_sstep_rows(1024, 1024, C.byref(r)) == 0 and r.value == 16
    assert lib.glx_set_sstep(2, 24) == 0
    assert lib.glx_get_sstep_rows(256, 256, C.byref(r)) == 0 and r.value == 24
    assert lib.glx_set_sstep(2, 0) == 0


def test_sstep_form_setting_without_device():
    """The s-step form: per size by default (wave strips from 512^2 sites up
    to lattices within 1024 waves of 52-row bands), a per-size table entry,
    and a process-wide choice; invalid values are refused (no device work)."""
    lib = C.CDLL(LIB)
    lib.glx_last_error.restype = C.c_char_p
    f = C.c_int()
    assert lib.glx_get_sstep_form(C.byref(f)) == 0 and f.value == -1
    for (lx, ly), want in {(256, 256): 0, (512, 512): 1, (768, 768): 1, (1024, 1024): 1, (1536, 1536): 1,
                           (2048, 2048): 0, (4096, 4096): 0, (1024, 64): 0, (2048, 512): 1}.items():
        assert lib.glx_get_sstep_form_for(lx, ly, C.byref(f)) == 0 and f.value == want, (lx, ly, f.value)
    assert lib.glx_set_sstep_form_entry(1024, 1024, 0) == 0
    assert lib.glx_get_sstep_form_for(1024, 1024, C.byref(f)) == 0 and f.value == 0
    assert lib.glx_set_sstep_form_entry(1024, 1024, 2) == 1 and b"entry" in lib.glx_last_error()
    assert lib.glx_set_sstep_form_entry(1024, 1024, -1) == 0
    assert lib.glx_get_sstep_form_for(1024, 1024, C.byref(f)) == 0 and f.value == 1
    assert lib.glx_set_sstep_form(0, 0) == 0
    assert lib.glx_get_sstep_form_for(1024, 1024, C.byref(f)) == 0 and f.value == 0
    assert lib.glx_set_sstep_form(1, 57) == 1 and b"band rows" in lib.glx_last_error()
    assert lib.glx_set_sstep_form(2, 0) == 1 and b"form must be" in lib.glx_last_error()
    assert lib.glx_set_sstep_form(-1, 0) == 0
    assert lib.glx_get_sstep_form(C.byref(f)) == 0 and f.value == -1
