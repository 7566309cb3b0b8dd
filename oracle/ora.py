"""ctypes wrapper of liboracle.so, the CPU FP64 restatement (ds_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
benchmark's CPU-baseline leg as the checker, never by the product.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

FORM_SN = 0
FORM_W = 1
TABLE_N = 200

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class Covar(C.Structure):
    _fields_ = [("var_s", C.c_double), ("c_perp", C.c_double),
                ("c_para", C.c_double), ("table_n", C.c_int),
                ("dx_inv", C.c_double), ("table", C.c_double * TABLE_N)]


def build():
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.ora_covar_init.argtypes = [C.POINTER(Covar), C.c_double, C.c_double,
                                     C.c_double]
        L.ora_linterp.argtypes = [C.c_double, C.c_int, C.c_double, _dp]
        L.ora_linterp.restype = C.c_double
        for f in (L.ora_signal_covar, L.ora_signal_covar_exact):
            f.argtypes = [C.c_double, C.c_double, C.POINTER(Covar)]
            f.restype = C.c_double
        L.ora_A_elem.argtypes = [C.c_int, C.c_int, _dp, C.POINTER(Covar),
                                 C.c_int, C.c_int, C.c_int]
        L.ora_A_elem.restype = C.c_double
        for f in (L.ora_A_product, L.ora_A_residual, L.ora_precond):
            f.argtypes = [C.c_int, C.c_int, _dp, C.POINTER(Covar), _dp, _dp]
        L.ora_norm_b.argtypes = [C.c_int, C.c_int, _dp]
        L.ora_norm_b.restype = C.c_double
        L.ora_solve_pcg.argtypes = [C.c_int, C.c_int, _dp, C.POINTER(Covar),
                                    C.c_int, C.c_int, C.c_double, _dp, _ip,
                                    _dp, _ip]
        L.ora_solve_pcg.restype = C.c_int
        L.ora_A_dense.argtypes = [C.c_int, C.c_int, _dp, C.POINTER(Covar),
                                  C.c_int, _dp]
        L.ora_cholesky_solve.argtypes = [C.c_int, _dp, _dp, _dp]
        L.ora_cholesky_solve.restype = C.c_int
        L.ora_solve_cf.argtypes = [C.c_int, C.c_int, _dp, C.POINTER(Covar),
                                   C.c_int, _dp]
        L.ora_solve_cf.restype = C.c_int
        L.ora_smp_product.argtypes = [C.c_int, _dp, C.c_int, C.c_int, _dp,
                                      C.c_int, C.POINTER(Covar), _dp, _dp]
        L.ora_sigma_m_noise_dom.argtypes = [C.c_int, _dp, C.c_int, _dp,
                                            C.c_int, C.POINTER(Covar), _dp]
        L.ora_spm_column.argtypes = [C.c_int, _dp, _dp, C.POINTER(Covar), _dp]
        L.ora_map_covar_diag.argtypes = [C.c_int, _dp, C.c_int, _dp, C.c_int,
                                         C.POINTER(Covar), C.c_int, C.c_int,
                                         C.c_double, _dp, _ip, _ip]
        L.ora_init_uniform_map_coords.argtypes = [C.c_double] * 3 + \
            [C.c_int] * 3 + [_dp]
        for f in (L.ora_pixel_to_npixel, L.ora_npixel_to_pixel,
                  L.ora_pixel_to_wpixel, L.ora_wpixel_to_pixel):
            f.argtypes = [C.c_int, _dp, _dp]
        L.ora_num_threads.restype = C.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def covar(var_s, l_perp, l_para) -> Covar:
    c = Covar()
    lib().ora_covar_init(C.byref(c), var_s, l_perp, l_para)
    return c


def signal_covar(xp2, xl2, c: Covar, exact=False):
    f = lib().ora_signal_covar_exact if exact else lib().ora_signal_covar
    xp2 = np.atleast_1d(np.asarray(xp2, dtype=np.float64))
    xl2 = np.broadcast_to(np.asarray(xl2, dtype=np.float64), xp2.shape)
    return np.array([f(a, b, C.byref(c)) for a, b in zip(xp2.ravel(),
                                                       xl2.ravel())]
                    ).reshape(xp2.shape)


def to_form(pixels, form):
    """Raw Pixel records -> NPixel (form SN) or WPixel (form W) records."""
    p = _f64(pixels).copy()
    out = np.empty_like(p)
    n = len(p)
    if form == FORM_W:
        lib().ora_pixel_to_wpixel(n, _p(p), _p(out))
    else:
        lib().ora_pixel_to_npixel(n, _p(p), _p(out))
    return out


def A_product(form, fpix, c, a):
    fpix, a = _f64(fpix), _f64(a)
    b = np.empty(len(fpix))
    lib().ora_A_product(form, len(fpix), _p(fpix), C.byref(c), _p(a), _p(b))
    return b


def A_residual(form, fpix, c, x):
    fpix, x = _f64(fpix), _f64(x)
    r = np.empty(len(fpix))
    lib().ora_A_residual(form, len(fpix), _p(fpix), C.byref(c), _p(x), _p(r))
    return r


def precond(form, fpix, c, b):
    fpix, b = _f64(fpix), _f64(b)
    x = np.empty(len(fpix))
    lib().ora_precond(form, len(fpix), _p(fpix), C.byref(c), _p(b), _p(x))
    return x


def norm_b(form, fpix):
    fpix = _f64(fpix)
    return lib().ora_norm_b(form, len(fpix), _p(fpix))


def solve_pcg(form, fpix, c, max_iter, step_r, tol, x0=None):
    """Returns (x, success, num_iters, residual_norms)."""
    fpix = _f64(fpix)
    n = len(fpix)
    x = np.zeros(n) if x0 is None else _f64(x0).copy()
    norms = np.empty(max(max_iter, 1) + 1)
    it = C.c_int(0)
    nn = C.c_int(0)
    ok = lib().ora_solve_pcg(form, n, _p(fpix), C.byref(c), max_iter, step_r,
                             tol, _p(x), C.byref(it), _p(norms), C.byref(nn))
    return x, bool(ok), it.value, norms[:nn.value].copy()


def A_dense(form, fpix, c, exact=False):
    fpix = _f64(fpix)
    n = len(fpix)
    A = np.empty((n, n))
    lib().ora_A_dense(form, n, _p(fpix), C.byref(c), int(exact), _p(A))
    return A


def solve_cf(form, fpix, c, exact=False):
    fpix = _f64(fpix)
    x = np.empty(len(fpix))
    rc = lib().ora_solve_cf(form, len(fpix), _p(fpix), C.byref(c), int(exact),
                            _p(x))
    if rc != 0:
        raise np.linalg.LinAlgError("matrix not positive definite")
    return x


def map_coords(box, shape):
    nx, ny, nz = shape
    out = np.empty((nx * ny * nz, 3))
    lib().ora_init_uniform_map_coords(float(box[0]), float(box[1]),
                                      float(box[2]), nx, ny, nz, _p(out))
    return out


def smp_product(pts, mpts, c, x):
    pts, mpts, x = _f64(pts), _f64(mpts), _f64(x)
    m = np.empty(len(mpts))
    lib().ora_smp_product(len(pts), _p(pts), pts.shape[1], len(mpts), _p(mpts),
                          mpts.shape[1], C.byref(c), _p(x), _p(m))
    return m


def sigma_m_noise_dom(npix, mpts, c):
    npix, mpts = _f64(npix), _f64(mpts)
    s = np.empty(len(mpts))
    lib().ora_sigma_m_noise_dom(len(npix), _p(npix), len(mpts), _p(mpts),
                                mpts.shape[1], C.byref(c), _p(s))
    return s


def spm_column(npix, mpt, c):
    npix, mpt = _f64(npix), _f64(mpt)
    col = np.empty(len(npix))
    lib().ora_spm_column(len(npix), _p(npix), _p(mpt), C.byref(c), _p(col))
    return col


def map_covar_diag(npix, mpts, c, max_iter, step_r, tol):
    npix, mpts = _f64(npix), _f64(mpts)
    nm = len(mpts)
    s = np.empty(nm)
    it = np.zeros(nm, dtype=np.int32)
    ok = np.zeros(nm, dtype=np.int32)
    lib().ora_map_covar_diag(len(npix), _p(npix), nm, _p(mpts), mpts.shape[1],
                             C.byref(c), max_iter, step_r, tol, _p(s),
                             it.ctypes.data_as(_ip), ok.ctypes.data_as(_ip))
    return s, it, ok.astype(bool)


def num_threads():
    return lib().ora_num_threads()
