"""Python mirror of dachshund's library interface over libdachshund_gpu.so.

Names, argument meaning and return conventions follow the reference's C++
API (lib/*.h); the arithmetic runs in the gfx950 kernels.  Pixel arrays are
numpy arrays of shape (n, 5) holding the reference's AoS records:
Pixel {x, y, z, n, d}, NPixel {x, y, z, N, d}, WPixel {x, y, z, w, wd}
(lib/pixel.h:16-45); points are (n, 3) arrays (lib/pixel.h:8-12).
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math
import os

import numpy as np

from . import _lib
from ._lib import (DS_F32_MFMA, DS_F32_MFMA_CLAMPED, DS_F32_VALU, DS_FORM_SN,
                   DS_FORM_W,
                   DS_PRECOND_BLOCK, DS_PRECOND_JACOBI,
                   DS_PREC_F32, DS_PREC_F64, DsCovar, DsPcgParams,
                   DsPcgResult, DsProblemStats, check, lib)

DEFAULT_TABLE_X_MAX = 4.0   # lib/signal_covar.h:7
DEFAULT_TABLE_N = 200       # lib/signal_covar.h:8


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None and a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


class SignalCovarParams:
    """lib/signal_covar.h:10-46: Gaussian covariance constants and the
    200-point table f_i = exp(-i * 8/199) (built with the C library's exp,
    like the reference's)."""

    def __init__(self, var_s: float, l_perp: float, l_para: float):
        self.var_s = float(var_s)
        self.gaussian_const_perp = 1.0 / (2.0 * l_perp * l_perp)
        self.gaussian_const_para = 1.0 / (2.0 * l_para * l_para)
        fx_max = DEFAULT_TABLE_X_MAX * DEFAULT_TABLE_X_MAX / 2.0
        self.func_table_n = DEFAULT_TABLE_N
        ft_dx = fx_max / (self.func_table_n - 1)
        self.func_table_dx_inv = 1.0 / ft_dx
        self.func_table = np.array([math.exp(-(ft_dx * i))
                                    for i in range(self.func_table_n)])
        self.l_perp = float(l_perp)
        self.l_para = float(l_para)

    def c_struct(self) -> DsCovar:
        self._keep = np.ascontiguousarray(self.func_table, dtype=np.float64)
        return DsCovar(self.var_s, self.gaussian_const_perp,
                       self.gaussian_const_para, self.func_table_dx_inv,
                       self.func_table_n, _dp(self._keep))


# ---- pixel formats (lib/pixel.cc:9-71), host side ------------------------

def points_from_pixels(pixels):
    """lib/pixel.cc:9-19."""
    return _f64(pixels)[:, :3].copy()


def pixel_to_npixel(p):
    """lib/pixel.cc:23-33: N = n*n."""
    p = _f64(p).copy()
    p[:, 3] = p[:, 3] * p[:, 3]
    return p


def npixel_to_pixel(p):
    """lib/pixel.cc:35-45: n = sqrt(N)."""
    p = _f64(p).copy()
    p[:, 3] = np.sqrt(p[:, 3])
    return p


def pixel_to_wpixel(p):
    """lib/pixel.cc:47-58: w = 1/n, wd = d/n."""
    p = _f64(p).copy()
    n = p[:, 3].copy()
    p[:, 3] = 1.0 / n
    p[:, 4] = p[:, 4] / n
    return p


def wpixel_to_pixel(p):
    """lib/pixel.cc:60-71: n = 1/w, d = wd/w."""
    p = _f64(p).copy()
    w = p[:, 3].copy()
    p[:, 3] = 1.0 / w
    p[:, 4] = p[:, 4] / w
    return p


def read_pixel_data(path, num_pixels):
    """lib/util.cc:11-23: raw little-endian float64 Pixel records."""
    if not os.path.exists(path):
        raise FileNotFoundError(f"Could not open file {path}")
    a = np.fromfile(path, dtype="<f8", count=5 * num_pixels)
    if a.size != 5 * num_pixels:
        raise ValueError(f"{path}: expected {num_pixels} pixels, "
                         f"found {a.size // 5}")
    return a.reshape(num_pixels, 5)


def init_uniform_map_coords(lx, ly, lz, nx, ny, nz):
    """lib/util.cc:25-46: cell centres, index (ix*ny + iy)*nz + iz."""
    dx, dy, dz = lx / nx, ly / ny, lz / nz
    x = dx * (np.arange(nx) + 0.5)
    y = dy * (np.arange(ny) + 0.5)
    z = dz * (np.arange(nz) + 0.5)
    out = np.empty((nx, ny, nz, 3))
    out[..., 0] = x[:, None, None]
    out[..., 1] = y[None, :, None]
    out[..., 2] = z[None, None, :]
    return out.reshape(-1, 3)


# ---- PCG types (lib/pcg.h:13-25) -----------------------------------------

@dataclasses.dataclass
class PCGParams:
    n: int
    max_iter: int
    step_r: int
    tol: float
    verbose: bool = False

    def c_struct(self):
        return DsPcgParams(self.n, self.max_iter, self.step_r, self.tol,
                           int(self.verbose))


@dataclasses.dataclass
class PCGResult:
    success: bool
    num_iters: int
    residual_norms: np.ndarray
    n_true: int = 0
    n_matvec64: int = 0
    n_matvec32: int = 0
    true_norm: float = 0.0
    seconds: float = 0.0


# ---- device context ------------------------------------------------------

class Context:
    """One HIP stream on one device (ds_ctx)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        rc = lib().ds_ctx_create(device, C.byref(h))
        if rc != _lib.DS_OK:
            raise _lib.DsError(rc, f"cannot create a context on device "
                               f"{device}")
        self.h = h

    @property
    def num_cu(self):
        return lib().ds_ctx_num_cu(self.h)

    def synchronize(self):
        """Wait for the work enqueued on the context's stream."""
        check(lib().ds_ctx_synchronize(self.h), self.h, "synchronize")

    def set_host_geometry(self, on: bool):
        """Build later problems' ordering and lists with the host threads
        (the device builder's reference; identical arrays) instead of on
        the device."""
        check(lib().ds_ctx_set_host_geometry(self.h, int(bool(on))), self.h)

    def close(self):
        if self.h:
            lib().ds_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


# ---- WFVector (lib/wf_vector.h:33-96) ------------------------------------

class _WFVector:
    FORM = None

    def __init__(self, pixels, s_params: SignalCovarParams, ctx=None):
        self.ctx = ctx or default_context()
        self.pixels = _f64(pixels)
        if self.pixels.ndim != 2 or self.pixels.shape[1] != 5:
            raise ValueError("pixels must have shape (n, 5)")
        self.n = len(self.pixels)
        self.s_params = s_params
        cov = s_params.c_struct()
        h = C.c_void_p()
        check(lib().ds_problem_create(self.ctx.h, self.FORM, self.n,
                                      _dp(self.pixels), C.byref(cov),
                                      C.byref(h)), self.ctx.h,
              "problem_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().ds_problem_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stats(self) -> DsProblemStats:
        s = DsProblemStats()
        check(lib().ds_problem_get_stats(self.h, C.byref(s)))
        return s

    def set_waves(self, waves: int):
        check(lib().ds_problem_set_waves(self.h, waves), self.ctx.h)

    # ds_problem_get_geometry: array id -> element dtype (and width)
    GEOMETRY = {
        "perm": (0, np.int32, 1), "pos": (1, np.float64, 4),
        "rhs": (2, np.float64, 1), "diag": (3, np.float64, 1),
        "wvec": (4, np.float64, 1), "loc32": (5, np.float32, 4),
        "tgt32": (6, np.float32, 4), "tcount": (7, np.int32, 1),
        "tbox_lo": (8, np.float64, 3), "tbox_hi": (9, np.float64, 3),
        "tptr": (10, np.int32, 1), "tidx": (11, np.int32, 1),
        "tmask": (12, np.uint8, 1), "toff32": (13, np.float32, 4),
        "hptr": (14, np.int32, 1), "hrec": (15, np.float32, 4),
        "gptr": (16, np.int32, 1), "gslot": (17, np.int32, 1),
        "blk_order": (18, np.int32, 1), "ord32": (19, np.int32, 1),
        "knn": (20, np.int32, 32), "minv": (21, np.float32, 33 * 64),
    }

    def geometry(self, name: str) -> np.ndarray:
        """One of the problem's geometry arrays as the kernels read it."""
        which, dt, width = self.GEOMETRY[name]
        nb = C.c_longlong()
        check(lib().ds_problem_get_geometry(self.h, which, None, 0,
                                            C.byref(nb)), self.ctx.h)
        out = np.empty(nb.value // np.dtype(dt).itemsize, dtype=dt)
        check(lib().ds_problem_get_geometry(
            self.h, which, out.ctypes.data_as(C.c_void_p), nb.value,
            C.byref(nb)), self.ctx.h)
        return out.reshape(-1, width) if width > 1 else out

    def _vec(self, a):
        a = _f64(a)
        if a.shape != (self.n,):
            raise ValueError(f"vector must have shape ({self.n},)")
        return a

    def A_product(self, a, prec=DS_PREC_F64):
        a = self._vec(a)
        b = np.empty(self.n)
        check(lib().ds_A_product(self.h, prec, _dp(a), _dp(b)), self.ctx.h)
        return b

    def A_residual(self, x):
        x = self._vec(x)
        r = np.empty(self.n)
        check(lib().ds_A_residual(self.h, _dp(x), _dp(r)), self.ctx.h)
        return r

    def preconditioner_solve(self, b):
        b = self._vec(b)
        x = np.empty(self.n)
        check(lib().ds_precond(self.h, _dp(b), _dp(x)), self.ctx.h)
        return x

    def preconditioner_solve_r32(self, b):
        """M^-1 b as the mixed solver applies the block-Jacobi inverse (b
        rounded to FP32); DS_PRECOND_BLOCK must be selected."""
        b = self._vec(b)
        x = np.empty(self.n)
        check(lib().ds_precond_r32(self.h, _dp(b), _dp(x)), self.ctx.h)
        return x

    def norm_b(self):
        v = C.c_double()
        check(lib().ds_norm_b(self.h, C.byref(v)), self.ctx.h)
        return v.value

    def _result(self, r: DsPcgResult, norms):
        return PCGResult(bool(r.success), r.num_iters, norms, r.n_true,
                         r.n_matvec64, r.n_matvec32, r.true_norm, r.seconds)

    def solve_pcg(self, params: PCGParams, x):
        """WFVector*::solve_pcg: x is updated in place (in: initial guess, W
        space for the W form; out: the solution).  Returns PCGResult."""
        if not (isinstance(x, np.ndarray) and x.dtype == np.float64
                and x.flags.c_contiguous and x.shape == (self.n,)):
            raise ValueError("x must be a contiguous float64 array of n")
        cap = max(params.max_iter, 1) + 1
        norms = np.empty(cap)
        res = DsPcgResult()
        prm = params.c_struct()
        check(lib().ds_solve_pcg(self.h, C.byref(prm), _dp(x), _dp(norms),
                                 cap, C.byref(res)), self.ctx.h, "solve_pcg")
        return self._result(res, norms[:res.n_norms].copy())

    def solve_mixed(self, params: PCGParams, x, refine_drop=1e-3):
        """Mixed FP32/FP64 solve (no reference counterpart), same in/out
        convention as solve_pcg; converged on the FP64 true residual."""
        if not (isinstance(x, np.ndarray) and x.dtype == np.float64
                and x.flags.c_contiguous and x.shape == (self.n,)):
            raise ValueError("x must be a contiguous float64 array of n")
        cap = max(params.max_iter, 1) + 1
        norms = np.empty(cap)
        res = DsPcgResult()
        prm = params.c_struct()
        check(lib().ds_solve_mixed(self.h, C.byref(prm), refine_drop, _dp(x),
                                   _dp(norms), cap, C.byref(res)), self.ctx.h,
              "solve_mixed")
        return self._result(res, norms[:res.n_norms].copy())

    # device-resident solution
    def set_x(self, x):
        x = self._vec(x)
        check(lib().ds_problem_set_x(self.h, _dp(x)), self.ctx.h)

    def get_x(self):
        x = np.empty(self.n)
        check(lib().ds_problem_get_x(self.h, _dp(x)), self.ctx.h)
        return x

    def set_launch(self, waves, order_by_cost=False):
        check(lib().ds_problem_set_launch(self.h, waves, int(order_by_cost)),
              self.ctx.h)

    def set_f32_kernel(self, kind):
        """DS_F32_MFMA (default) or DS_F32_VALU form of the FP32 kernel."""
        check(lib().ds_problem_set_f32_kernel(self.h, kind), self.ctx.h)

    def set_tile(self, f64_rows):
        """FP64 pair-kernel tile: rows per group (64, 32 or 16) x 16
        source slots; 32 and 16 skip dead 16-row sub-blocks."""
        check(lib().ds_problem_set_tile(self.h, f64_rows), self.ctx.h)

    def launch(self):
        """(waves per block, longest-first order, FP64 tile rows, FP32
        kernel form) in use."""
        v = [C.c_int() for _ in range(4)]
        check(lib().ds_problem_get_launch(self.h, *[C.byref(x) for x in v]),
              self.ctx.h)
        return tuple(x.value for x in v)

    def set_precond(self, kind):
        """DS_PRECOND_JACOBI (the reference's; default) or DS_PRECOND_BLOCK
        (exact inverses of the 64-row target blocks) for this problem's
        solves and preconditioner_solve."""
        check(lib().ds_problem_set_precond(self.h, kind), self.ctx.h)

    def solve_resident(self, params: PCGParams, mixed=False,
                       refine_drop=1e-3):
        res = DsPcgResult()
        prm = params.c_struct()
        if mixed:
            rc = lib().ds_solve_mixed_resident(self.h, C.byref(prm),
                                               refine_drop, C.byref(res))
        else:
            rc = lib().ds_solve_pcg_resident(self.h, C.byref(prm),
                                             C.byref(res))
        check(rc, self.ctx.h, "solve")
        return self._result(res, None)

    def fill_x(self, value=0.0):
        check(lib().ds_problem_fill_x(self.h, float(value)), self.ctx.h)

    def time_matvec(self, prec=DS_PREC_F64, reps=10):
        """Average pair-kernel launch duration (ms), HIP events."""
        ms = C.c_double()
        check(lib().ds_problem_time_matvec(self.h, prec, reps, C.byref(ms)),
              self.ctx.h)
        return ms.value

    def solve_cf(self, exact=False):
        """WFVector*::solve_cf on the GPU (dense Cholesky); returns x."""
        x = np.empty(self.n)
        check(lib().ds_solve_cf(self.h, int(bool(exact)), _dp(x)), self.ctx.h,
              "solve_cf")
        return x

    def smp_grid(self, lx, ly, lz, nx, ny, nz):
        """smp_product (lib/map.cc:20-43) of the resident solution onto the
        uniform grid (lib/util.cc:25-46)."""
        m = np.empty(nx * ny * nz)
        check(lib().ds_problem_smp_grid(self.h, lx, ly, lz, nx, ny, nz,
                                        _dp(m)), self.ctx.h, "smp_grid")
        return m


    def smp_grid_noalloc(self, lx, ly, lz, nx, ny, nz):
        """smp_grid with every per-call device buffer in the library's
        static scratch array (ds_problem_smp_grid_noalloc): no device
        allocation; the same map bit for bit."""
        m = np.empty(nx * ny * nz)
        check(lib().ds_problem_smp_grid_noalloc(self.h, lx, ly, lz, nx, ny,
                                                nz, _dp(m)), self.ctx.h,
              "smp_grid_noalloc")
        return m


class BatchWF(_WFVector):
    """Many independent subvolumes solved together (ds_batch_*): every pair
    and CG kernel processes all of them in one launch; each has its own
    covariance, CG scalars and stopping test.  `pixels` is a list of (n_k,
    5) record arrays in the given form, `s_params` a list of
    SignalCovarParams (all with the same table)."""

    def __init__(self, pixels, s_params, form="w", ctx=None):
        self.ctx = ctx or default_context()
        self.FORM = DS_FORM_W if form == "w" else DS_FORM_SN
        parts = [_f64(p) for p in pixels]
        self.sub_n = np.array([len(p) for p in parts], dtype=np.int32)
        self.pixels = np.ascontiguousarray(np.concatenate(parts))
        self.n = len(self.pixels)
        self.n_sub = len(parts)
        self.s_params = list(s_params)
        covs = (DsCovar * self.n_sub)(*[sp.c_struct() for sp in s_params])
        h = C.c_void_p()
        check(lib().ds_batch_create(self.ctx.h, self.FORM, self.n_sub,
                                    self.sub_n.ctypes.data_as(
                                        C.POINTER(C.c_int)),
                                    _dp(self.pixels), covs, C.byref(h)),
              self.ctx.h, "batch_create")
        self.h = h
        self._dense = False
        self.offsets = np.concatenate([[0], np.cumsum(self.sub_n)])

    def split(self, v):
        return [v[a:b] for a, b in zip(self.offsets[:-1], self.offsets[1:])]

    def solve(self, params: PCGParams, mixed=True, refine_drop=None,
              use_graph=True):
        """Solve all subvolumes from the resident x; one PCGResult each.
        refine_drop defaults to 0.1 with the explicit BF16 block as inner
        operator (relative error ~2^-9) and 1e-3 with the FP32 one."""
        if refine_drop is None:
            refine_drop = 0.1 if self._dense else 1e-3
        res = (DsPcgResult * self.n_sub)()
        prm = params.c_struct()
        check(lib().ds_batch_solve(self.h, C.byref(prm), int(mixed),
                                   refine_drop, int(use_graph), res),
              self.ctx.h, "batch_solve")
        return [self._result(r, None) for r in res]

    def set_launch(self, waves, order_by_cost):
        check(lib().ds_problem_set_launch(self.h, waves, int(order_by_cost)),
              self.ctx.h)

    def set_f32_kernel(self, kind):
        check(lib().ds_problem_set_f32_kernel(self.h, kind), self.ctx.h)

    def set_precond(self, kind):
        check(lib().ds_problem_set_precond(self.h, kind), self.ctx.h)

    def sub_pairs(self):
        """Per subvolume: evaluated pairs per matvec of the FP64 kernel and
        of the MFMA FP32 kernel."""
        p64 = np.zeros(self.n_sub, dtype=np.int64)
        p32 = np.zeros(self.n_sub, dtype=np.int64)
        check(lib().ds_batch_sub_pairs(
            self.h, p64.ctypes.data_as(C.POINTER(C.c_longlong)),
            p32.ctypes.data_as(C.POINTER(C.c_longlong))), self.ctx.h)
        return p64, p32

    def build_dense_bf16(self):
        """Explicit BF16 S for subvolumes sharing one geometry; the mixed
        solver then uses one MFMA GEMM for all right-hand sides."""
        check(lib().ds_problem_build_dense_bf16(self.h), self.ctx.h,
              "build_dense_bf16")
        self._dense = True

    def set_dense(self, on):
        check(lib().ds_problem_set_dense(self.h, int(bool(on))), self.ctx.h)
        self._dense = bool(on)

    def set_fused(self, on):
        """Fused multi-RHS operator (S tiles on the fly x all right-hand
        sides on MFMA, BF16); switches the explicit block off."""
        check(lib().ds_problem_set_fused(self.h, int(bool(on))), self.ctx.h)
        self._dense = bool(on)

    def dense_product(self, a):
        a = self._vec(a)
        b = np.empty(self.n)
        check(lib().ds_problem_dense_product(self.h, _dp(a), _dp(b)),
              self.ctx.h, "dense_product")
        return b

    def time_dense(self, reps=10):
        ms = C.c_double()
        check(lib().ds_problem_time_dense(self.h, reps, C.byref(ms)),
              self.ctx.h)
        return ms.value

    def autotune(self, prec=DS_PREC_F32, reps=5):
        w, o, ms = C.c_int(), C.c_int(), C.c_double()
        check(lib().ds_problem_autotune(self.h, prec, reps, C.byref(w),
                                        C.byref(o), C.byref(ms)), self.ctx.h)
        return w.value, bool(o.value), ms.value


class WFVectorSN(_WFVector):
    """x = (S + N)^-1 d with NPixel records (lib/wf_vector.cc:16-155)."""
    FORM = DS_FORM_SN


class WFVectorW(_WFVector):
    """x = w (w S w + I)^-1 w d with WPixel records
    (lib/wf_vector.cc:157-319)."""
    FORM = DS_FORM_W


# ---- map-side products ---------------------------------------------------

def smp_product(pixel_coords, map_coords, s_params: SignalCovarParams, x,
                ctx=None):
    """lib/map.cc:20-43: m_i = sum_j S(map_i, pix_j) x_j."""
    ctx = ctx or default_context()
    pc = _f64(pixel_coords)
    mc = _f64(map_coords)
    x = _f64(x)
    if pc.ndim != 2 or pc.shape[1] != 3 or mc.ndim != 2 or mc.shape[1] != 3:
        raise ValueError("coordinates must have shape (n, 3)")
    if x.shape != (len(pc),):
        raise ValueError("x must have one entry per pixel")
    m = np.empty(len(mc))
    cov = s_params.c_struct()
    check(lib().ds_smp_product(ctx.h, len(pc), _dp(pc), len(mc), _dp(mc),
                               C.byref(cov), _dp(x), _dp(m)), ctx.h,
          "smp_product")
    return m


def smp_product_grid(pixel_coords, box, shape, s_params: SignalCovarParams,
                     x, ctx=None):
    """smp_product onto init_uniform_map_coords(box, shape) cells."""
    ctx = ctx or default_context()
    pc = _f64(pixel_coords)
    x = _f64(x)
    if pc.ndim != 2 or pc.shape[1] != 3 or x.shape != (len(pc),):
        raise ValueError("bad pixel coordinates or x")
    nx, ny, nz = (int(v) for v in shape)
    m = np.empty(nx * ny * nz)
    cov = s_params.c_struct()
    check(lib().ds_smp_product_grid(ctx.h, len(pc), _dp(pc), float(box[0]),
                                    float(box[1]), float(box[2]), nx, ny, nz,
                                    C.byref(cov), _dp(x), _dp(m)), ctx.h,
          "smp_product_grid")
    return m


def sigma_m_noise_dom(npixels, map_coords, s_params: SignalCovarParams,
                      ctx=None):
    """lib/map_covar.cc:11-38: sigma_i = sum_j S_ij^2 / N_j."""
    ctx = ctx or default_context()
    npx = _f64(npixels)
    mc = _f64(map_coords)
    if npx.ndim != 2 or npx.shape[1] != 5 or mc.ndim != 2 or mc.shape[1] != 3:
        raise ValueError("bad shapes")
    s = np.empty(len(mc))
    cov = s_params.c_struct()
    check(lib().ds_sigma_m_noise_dom(ctx.h, len(npx), _dp(npx), len(mc),
                                     _dp(mc), C.byref(cov), _dp(s)), ctx.h,
          "sigma_m_noise_dom")
    return s


def map_covar_diag(npixels, map_coords, s_params: SignalCovarParams,
                   params: PCGParams, ctx=None):
    """MapCovarDiag::solve_pcg (lib/map_covar.cc:73-113) on the GPU.
    Returns (sigma, num_iters, success)."""
    ctx = ctx or default_context()
    npx = _f64(npixels)
    mc = _f64(map_coords)
    nm = len(mc)
    s = np.empty(nm)
    it = np.zeros(nm, dtype=np.int32)
    ok = np.zeros(nm, dtype=np.int32)
    cov = s_params.c_struct()
    prm = params.c_struct()
    ip = C.POINTER(C.c_int)
    check(lib().ds_map_covar_diag(ctx.h, len(npx), _dp(npx), nm, _dp(mc),
                                  C.byref(cov), C.byref(prm), _dp(s),
                                  it.ctypes.data_as(ip),
                                  ok.ctypes.data_as(ip)), ctx.h,
          "map_covar_diag")
    return s, it, ok.astype(bool)


def plan_tiles(points, s_params: SignalCovarParams):
    """Host-only: the ordering and culled (block, tile) lists of a problem
    over these points.  Returns (perm, tile_ptr, tile_idx)."""
    pts = _f64(points)
    stride = pts.shape[1]
    cov = s_params.c_struct()
    nb, ni = C.c_int(), C.c_int()
    ip = C.POINTER(C.c_int)
    check(lib().ds_plan_tiles(len(pts), _dp(pts), stride, C.byref(cov),
                              None, None, None, 0, C.byref(nb), C.byref(ni)))
    perm = np.empty(nb.value * 64, dtype=np.int32)
    ptr = np.empty(nb.value + 1, dtype=np.int32)
    idx = np.empty(max(ni.value, 1), dtype=np.int32)
    check(lib().ds_plan_tiles(len(pts), _dp(pts), stride, C.byref(cov),
                              perm.ctypes.data_as(ip), ptr.ctypes.data_as(ip),
                              idx.ctypes.data_as(ip), max(ni.value, 1),
                              C.byref(nb), C.byref(ni)))
    return perm, ptr, idx[:ni.value]
