"""CPU oracle for the sparse-solver path (TEST INFRASTRUCTURE ONLY).

ctypes wrapper over liboracle.so, a C++ restatement of the reference
algorithms (oracle/oracle.cpp, each function cites the reference file it
follows). Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
may import this package, and only as the checker / CPU baseline; the product
library never calls it.

Parity pins of every function are listed in oracle/oracle.cpp's header and
exercised by tests/test_oracle.py.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "oracle.cpp")
LIB = os.path.join(HERE, "liboracle.so")

_P = C.c_void_p
_lib = None


def build():
    """Compiles liboracle.so with g++ (strict FP order, no contraction)."""
    tmp = LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
           "-fopenmp", "-fPIC", "-shared", "-o", tmp, SRC]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


class PBiCGStabCfg(C.Structure):
    _fields_ = [("absTolerance", C.c_float), ("relTolerance", C.c_float),
                ("relResidual", C.c_float), ("maxIterations", C.c_int32),
                ("minIterations", C.c_int32), ("working64", C.c_int32),
                ("precond", C.c_int32), ("factor64", C.c_int32),
                ("gsSweeps", C.c_int32), ("gsSymmetric", C.c_int32),
                ("bNormFromX", C.c_int32), ("norm", C.c_int32)]


class MPIRCfg(C.Structure):
    _fields_ = [("absTolerance", C.c_float), ("relTolerance", C.c_float),
                ("relResidual", C.c_float), ("maxIterations", C.c_int32),
                ("minIterations", C.c_int32), ("normFromRounded", C.c_int32),
                ("inner", PBiCGStabCfg), ("norm", C.c_int32)]


# norm codes of the oracle (L1: sum of magnitudes, LINF: largest magnitude)
NORMS = {"L2": 0, "L1": 1, "LINF": 2}


def vec_norm(v, tile_start, norm="L2"):
    """The oracle's vector norm (float32 or float64, in the reference's
    reduction order for L2 and L1)."""
    ts = np.ascontiguousarray(tile_start, np.int64)
    v = np.ascontiguousarray(v)
    f = lib().orc_vec_norm_f32 if v.dtype == np.float32 else lib().orc_vec_norm_f64
    return f(_p(v), _i64(len(v)), C.c_int32(len(ts) - 1), _p(ts), C.c_int32(NORMS[norm]))


class Stats(C.Structure):
    _fields_ = [("initialResidual", C.c_double), ("finalResidual", C.c_double),
                ("iterations", C.c_int32), ("converged", C.c_int32),
                ("singular", C.c_int32), ("pad", C.c_int32)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(SRC):
            build()
        L = C.CDLL(LIB)
        L.orc_poisson7.restype = C.c_int64
        L.orc_coo_to_crs.restype = C.c_int64
        L.orc_tile_layout.restype = C.c_int64
        L.orc_levels.restype = C.c_int64
        L.orc_ldu_build.restype = C.c_int64
        L.orc_reduce_add_f32.restype = C.c_float
        L.orc_reduce_add_f64.restype = C.c_double
        L.orc_system_create.restype = _P
        L.orc_pbicgstab_f32.restype = C.c_int32
        L.orc_pbicgstab_f64.restype = C.c_int32
        L.orc_mpir.restype = C.c_int32
        L.orc_mpir_dw.restype = C.c_int32
        L.orc_max_threads.restype = C.c_int32
        L.orc_vec_norm_f32.restype = C.c_float
        L.orc_vec_norm_f64.restype = C.c_double
        _lib = L
    return _lib


def set_threads(n):
    """Threads of the oracle's parallel loops (results do not depend on it)."""
    lib().orc_set_threads(C.c_int32(int(n)))


def max_threads():
    return int(lib().orc_max_threads())


def _p(a):
    return a.ctypes.data_as(_P)


def _i64(x):
    return C.c_int64(int(x))


def uniform_pm1(seed, n):
    out = np.empty(n, np.float32)
    lib().orc_rng_uniform_pm1(C.c_uint64(seed), _i64(n), _p(out))
    return out


def rng_u64(seed, n):
    out = np.empty(n, np.uint64)
    lib().orc_rng_u64(C.c_uint64(seed), _i64(n), _p(out))
    return out


def poisson7(nx, ny, nz):
    c = lib().orc_poisson7(_i64(nx), _i64(ny), _i64(nz), None, None, None)
    rows = np.empty(c, np.int64)
    cols = np.empty(c, np.int64)
    vals = np.empty(c, np.float32)
    lib().orc_poisson7(_i64(nx), _i64(ny), _i64(nz), _p(rows), _p(cols), _p(vals))
    return rows, cols, vals


def coo_to_crs(n, rows, cols, vals):
    rows = np.ascontiguousarray(rows, np.int64)
    cols = np.ascontiguousarray(cols, np.int64)
    vals = np.ascontiguousarray(vals, np.float32)
    nnz = len(vals)
    rowPtr = np.empty(n + 1, np.int64)
    col = np.empty(max(nnz, 1), np.int32)
    off = np.empty(max(nnz, 1), np.float32)
    diag = np.empty(n, np.float32)
    m = lib().orc_coo_to_crs(_i64(n), _i64(nnz), _p(rows), _p(cols), _p(vals),
                             _p(rowPtr), _p(col), _p(off), _p(diag))
    return rowPtr, col[:m].copy(), off[:m].copy(), diag


def tile_layout(n, rowPtr, col, off, diag, T, rowToTile):
    """Returns the tile-major layout dict (see oracle.cpp orc_tile_layout)."""
    nnz = int(rowPtr[-1])
    rowPtr = np.ascontiguousarray(rowPtr, np.int64)
    col = np.ascontiguousarray(col, np.int32)
    off = np.ascontiguousarray(off, np.float32)
    diag = np.ascontiguousarray(diag, np.float32)
    r2t = np.ascontiguousarray(rowToTile, np.int32)
    o = dict(tile_start=np.empty(T + 1, np.int64), new_to_old=np.empty(n, np.int64),
             num_interior=np.empty(T, np.int32), num_separator_regions=np.empty(T, np.int32),
             halo_start=np.empty(T + 1, np.int64), halo_global=np.empty(max(nnz, 1), np.int64),
             row_ptr=np.empty(n + 1, np.int64), col_local=np.empty(max(nnz, 1), np.int32),
             col=np.empty(max(nnz, 1), np.int64), off=np.empty(max(nnz, 1), np.float32),
             diag=np.empty(n, np.float32))
    h = lib().orc_tile_layout(
        _i64(n), _p(rowPtr), _p(col), _p(off), _p(diag), C.c_int32(T), _p(r2t),
        _p(o["tile_start"]), _p(o["new_to_old"]), _p(o["num_interior"]),
        _p(o["num_separator_regions"]), _p(o["halo_start"]), _p(o["halo_global"]),
        _p(o["row_ptr"]), _p(o["col_local"]), _p(o["col"]), _p(o["off"]), _p(o["diag"]))
    if h < 0:
        raise ValueError("invalid partition")
    o["halo_global"] = o["halo_global"][:h].copy()
    for k in ("col_local", "col", "off"):
        o[k] = o[k][:nnz].copy()
    o["T"] = T
    o["n"] = n
    return o


def levels(n, T, tile_start, row_ptr, col):
    tile_start = np.ascontiguousarray(tile_start, np.int64)
    row_ptr = np.ascontiguousarray(row_ptr, np.int64)
    col = np.ascontiguousarray(col, np.int64)
    level = np.empty(n, np.int32)
    addr = np.empty(n, np.int32)
    start = np.empty(n + T + 1, np.int64)
    numc = np.empty(T, np.int32)
    cs = lib().orc_levels(_i64(n), C.c_int32(T), _p(tile_start), _p(row_ptr), _p(col),
                          _p(level), _p(addr), _p(start), _p(numc))
    return dict(level=level, color_sort_addr=addr, color_start=start[:cs].copy(),
                num_colors=numc)


def _sys_arrays(L):
    return (np.ascontiguousarray(L["row_ptr"], np.int64),
            np.ascontiguousarray(L["col"], np.int64),
            np.ascontiguousarray(L["off"], np.float32),
            np.ascontiguousarray(L["diag"], np.float32))


def spmv(L, x):
    rp, col, off, diag = _sys_arrays(L)
    n = len(diag)
    if x.dtype == np.float32:
        y = np.empty(n, np.float32)
        lib().orc_spmv_f32(_i64(n), _p(rp), _p(col), _p(off), _p(diag),
                           _p(np.ascontiguousarray(x)), _p(y))
    else:
        y = np.empty(n, np.float64)
        lib().orc_spmv_f64(_i64(n), _p(rp), _p(col), _p(off), _p(diag),
                           _p(np.ascontiguousarray(x, np.float64)), _p(y))
    return y


def residual_f32(L, x, b):
    rp, col, off, diag = _sys_arrays(L)
    n = len(diag)
    r = np.empty(n, np.float32)
    lib().orc_residual_f32(_i64(n), _p(rp), _p(col), _p(off), _p(diag),
                           _p(np.ascontiguousarray(x, np.float32)),
                           _p(np.ascontiguousarray(b, np.float32)), _p(r))
    return r


def residual_mixed(L, x64, b32):
    rp, col, off, diag = _sys_arrays(L)
    n = len(diag)
    r32 = np.empty(n, np.float32)
    r64 = np.empty(n, np.float64)
    lib().orc_residual_mixed(_i64(n), _p(rp), _p(col), _p(off), _p(diag),
                             _p(np.ascontiguousarray(x64, np.float64)),
                             _p(np.ascontiguousarray(b32, np.float32)), _p(r32), _p(r64))
    return r32, r64


TWOFLOAT = np.dtype([("hi", "<f4"), ("lo", "<f4")])


def _pairs(a):
    """(n, 2) float32 view of double words given as (n, 2) floats or the
    structured TWOFLOAT dtype."""
    a = np.ascontiguousarray(a)
    if a.dtype == TWOFLOAT:
        a = a.view(np.float32).reshape(-1, 2)
    return np.ascontiguousarray(a, np.float32).reshape(-1, 2)


def dw_two_sum(a, b):
    out = np.empty(2, np.float32)
    lib().orc_dw_two_sum(C.c_float(a), C.c_float(b), _p(out))
    return out


def dw_two_prod(a, b):
    out = np.empty(2, np.float32)
    lib().orc_dw_two_prod(C.c_float(a), C.c_float(b), _p(out))
    return out


def dw_add(x, y):
    """Double-word sums of (n, 2) float32 arrays (hi, lo)."""
    x, y = _pairs(x), _pairs(y)
    z = np.empty_like(x)
    lib().orc_dw_add(_i64(len(x)), _p(x), _p(y), _p(z))
    return z


def dw_mul(x, y):
    x, y = _pairs(x), _pairs(y)
    z = np.empty_like(x)
    lib().orc_dw_mul(_i64(len(x)), _p(x), _p(y), _p(z))
    return z


def residual_dw(L, xhl, b32):
    """Double-word residual (CRS order): r32 = fl32(b - A x), r64 = the same
    rounded to float64; x as (n, 2) float pairs."""
    rp, col, off, diag = _sys_arrays(L)
    n = len(diag)
    x = _pairs(xhl)
    r32 = np.empty(n, np.float32)
    r64 = np.empty(n, np.float64)
    lib().orc_residual_dw(_i64(n), _p(rp), _p(col), _p(off), _p(diag), _p(x),
                          _p(np.ascontiguousarray(b32, np.float32)), _p(r32), _p(r64))
    return r32, r64


def ldu_residual_dw(U, xhl, b32):
    n = len(U["diag"])
    x = _pairs(xhl)
    r32 = np.empty(n, np.float32)
    r64 = np.empty(n, np.float64)
    lib().orc_ldu_residual_dw(_i64(n), _p(U["lower_addr"]), _p(U["upper_addr"]),
                              _p(U["owner_start"]), _p(U["neighbour_start"]),
                              _p(U["neighbour_col_ind"]), _p(U["upper"]), _p(U["lower"]),
                              _p(U["diag"]), _p(x), _p(np.ascontiguousarray(b32, np.float32)),
                              _p(r32), _p(r64))
    return r32, r64


def ldu_build(L):
    rp, col, off, diag = _sys_arrays(L)
    n = len(diag)
    nnz = len(off)
    F = nnz // 2 + 1
    o = dict(lower_addr=np.empty(F, np.int64), upper_addr=np.empty(F, np.int64),
             owner_start=np.empty(n + 1, np.int64), neighbour_start=np.empty(n + 1, np.int64),
             neighbour_col_ind=np.empty(F, np.int64), upper=np.empty(F, np.float32),
             lower=np.empty(F, np.float32))
    f = lib().orc_ldu_build(_i64(n), _p(rp), _p(col), _p(off), _p(o["lower_addr"]),
                            _p(o["upper_addr"]), _p(o["owner_start"]),
                            _p(o["neighbour_start"]), _p(o["neighbour_col_ind"]),
                            _p(o["upper"]), _p(o["lower"]))
    if f < 0:
        raise ValueError("pattern not structurally symmetric")
    for k in ("lower_addr", "upper_addr", "neighbour_col_ind", "upper", "lower"):
        o[k] = o[k][:f].copy()
    o["num_faces"] = f
    o["diag"] = diag
    return o


def ldu_spmv_f32(U, x):
    n = len(U["diag"])
    y = np.empty(n, np.float32)
    lib().orc_ldu_spmv_f32(_i64(n), _p(U["lower_addr"]), _p(U["upper_addr"]),
                           _p(U["owner_start"]), _p(U["neighbour_start"]),
                           _p(U["neighbour_col_ind"]), _p(U["upper"]), _p(U["lower"]),
                           _p(U["diag"]), _p(np.ascontiguousarray(x, np.float32)), _p(y))
    return y


def ldu_residual_mixed(U, x64, b32):
    n = len(U["diag"])
    r32 = np.empty(n, np.float32)
    r64 = np.empty(n, np.float64)
    lib().orc_ldu_residual_mixed(_i64(n), _p(U["lower_addr"]), _p(U["upper_addr"]),
                                 _p(U["owner_start"]), _p(U["neighbour_start"]),
                                 _p(U["neighbour_col_ind"]), _p(U["upper"]),
                                 _p(U["lower"]), _p(U["diag"]),
                                 _p(np.ascontiguousarray(x64, np.float64)),
                                 _p(np.ascontiguousarray(b32, np.float32)), _p(r32), _p(r64))
    return r32, r64


def ilu0_factor(L, dtype=np.float32, diagonalBased=False):
    """ILU(0) factors (lu, inverted diag); DILU: lu = A's off-diagonals."""
    rp, col, off, diag = _sys_arrays(L)
    n = len(diag)
    ts = np.ascontiguousarray(L["tile_start"], np.int64)
    T = len(ts) - 1
    lu = off.astype(dtype).copy()
    d = diag.astype(dtype).copy()
    if diagonalBased:
        f = lib().orc_dilu_factor_f32 if dtype == np.float32 else lib().orc_dilu_factor_f64
    else:
        f = lib().orc_ilu0_factor_f32 if dtype == np.float32 else lib().orc_ilu0_factor_f64
    f(_i64(n), C.c_int32(T), _p(ts), _p(rp), _p(col), _p(lu), _p(d))
    return lu, d


def ilu0_apply(L, lu, invD, b, diagonalBased=False):
    rp, col, _, _ = _sys_arrays(L)
    ts = np.ascontiguousarray(L["tile_start"], np.int64)
    T = len(ts) - 1
    x = np.zeros(len(b), lu.dtype)
    b = np.ascontiguousarray(b, lu.dtype)
    if diagonalBased:
        f = lib().orc_dilu_apply_f32 if lu.dtype == np.float32 else lib().orc_dilu_apply_f64
    else:
        f = lib().orc_ilu0_apply_f32 if lu.dtype == np.float32 else lib().orc_ilu0_apply_f64
    f(C.c_int32(T), _p(ts), _p(rp), _p(col), _p(lu), _p(invD), _p(b), _p(x))
    return x


def gs_sweep(L, b, x, backward=False):
    rp, col, off, diag = _sys_arrays(L)
    n = len(diag)
    ts = np.ascontiguousarray(L["tile_start"], np.int64)
    T = len(ts) - 1
    x = np.array(x, copy=True)
    b = np.ascontiguousarray(b, x.dtype)
    f = lib().orc_gs_sweep_f32 if x.dtype == np.float32 else lib().orc_gs_sweep_f64
    f(_i64(n), C.c_int32(T), _p(ts), _p(rp), _p(col), _p(off), _p(diag), _p(b), _p(x),
      C.c_int32(1 if backward else 0))
    return x


def reduce_add(v, tile_start, square=False):
    ts = np.ascontiguousarray(tile_start, np.int64)
    T = len(ts) - 1
    v = np.ascontiguousarray(v)
    if v.dtype == np.float32:
        return lib().orc_reduce_add_f32(_p(v), C.c_int32(T), _p(ts), C.c_int32(int(square)))
    return lib().orc_reduce_add_f64(_p(np.ascontiguousarray(v, np.float64)), C.c_int32(T),
                                    _p(ts), C.c_int32(int(square)))


class System:
    """Oracle solver context on a tile-major layout dict."""

    def __init__(self, L):
        rp, col, off, diag = _sys_arrays(L)
        ts = np.ascontiguousarray(L["tile_start"], np.int64)
        self.n = len(diag)
        self._keep = (rp, col, off, diag, ts)
        self._h = lib().orc_system_create(_i64(self.n), C.c_int32(len(ts) - 1), _p(ts),
                                          _p(rp), _p(col), _p(off), _p(diag))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_system_destroy(C.c_void_p(self._h))
            self._h = None

    @staticmethod
    def pbicgstab_cfg(relResidual=0.0, maxIterations=1000, precond="ilu", working64=False,
                      factor64=None, gsSweeps=1, gsSymmetric=False, absTolerance=0.0,
                      relTolerance=0.0, minIterations=0, bNormFromX=True, norm="L2"):
        c = PBiCGStabCfg()
        c.absTolerance = absTolerance
        c.relTolerance = relTolerance
        c.relResidual = relResidual
        c.maxIterations = maxIterations
        c.minIterations = minIterations
        c.working64 = int(working64)
        c.precond = {"none": 0, None: 0, "ilu": 1, "gs": 2, "dilu": 3}[precond]
        c.factor64 = int(working64 if factor64 is None else factor64)
        c.gsSweeps = gsSweeps
        c.gsSymmetric = int(gsSymmetric)
        c.bNormFromX = int(bNormFromX)
        c.norm = NORMS[norm]
        return c

    def pbicgstab(self, x, b, cfg):
        st = Stats()
        x = np.array(x, copy=True)
        b = np.ascontiguousarray(b, x.dtype)
        f = lib().orc_pbicgstab_f64 if x.dtype == np.float64 else lib().orc_pbicgstab_f32
        f(C.c_void_p(self._h), C.byref(cfg), _p(x), _p(b), C.byref(st))
        return x, st

    def mpir(self, b32, relResidual=1e-10, maxIterations=100, inner=None,
             normFromRounded=False, phi0=None, norm="L2"):
        c = MPIRCfg()
        c.norm = NORMS[norm]
        c.relResidual = relResidual
        c.maxIterations = maxIterations
        c.normFromRounded = int(normFromRounded)
        c.inner = inner if inner is not None else self.pbicgstab_cfg(1e-3, 100)
        st = Stats()
        phi = np.zeros(self.n, np.float64) if phi0 is None else np.array(phi0, np.float64)
        inner_it = C.c_int32(0)
        lib().orc_mpir(C.c_void_p(self._h), C.byref(c), _p(phi),
                       _p(np.ascontiguousarray(b32, np.float32)), C.byref(st),
                       C.byref(inner_it))
        return phi, st, inner_it.value

    def mpir_dw(self, b32, relResidual=1e-10, maxIterations=100, inner=None,
                normFromRounded=False, norm="L2"):
        """MPIR with twofloat32 extended precision: phi returned as (n, 2)
        float32 (hi, lo) pairs."""
        c = MPIRCfg()
        c.norm = NORMS[norm]
        c.relResidual = relResidual
        c.maxIterations = maxIterations
        c.normFromRounded = int(normFromRounded)
        c.inner = inner if inner is not None else self.pbicgstab_cfg(1e-3, 100)
        st = Stats()
        phi = np.zeros((self.n, 2), np.float32)
        inner_it = C.c_int32(0)
        lib().orc_mpir_dw(C.c_void_p(self._h), C.byref(c), _p(phi),
                          _p(np.ascontiguousarray(b32, np.float32)), C.byref(st),
                          C.byref(inner_it))
        return phi, st, inner_it.value
