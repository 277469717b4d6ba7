"""graphene_amd: MI355X-native sparse-solver hot path of Graphene.

Python mirror of the reference's host API (lg/matrix/host/HostMatrix.hpp,
lg/matrix/Matrix.hpp, lg/matrix/solver/Configuration.hpp) over the C-ABI
library libgraphene_amd.so (include/graphene_amd.h). All compute runs in the
library's gfx950 kernels; there is no fallback: if the library is missing,
importing the device parts raises.

    A_host = generate3DPoissonMatrix(32, 32, 32, numTiles=1)
    A = A_host.copyToTile()
    x = A.createVector(np.float32); b = A * x
    stats = A.solve(x, b, {"type": "PBiCGStab", ...})
"""
import ctypes as C
import json as _json
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgraphene_amd.so")

FLOAT16, FLOAT32, FLOAT64, TWOFLOAT32 = 1, 2, 3, 4
CRS, LDU = 0, 1
# twofloat32 (two<float>): per element a float pair whose value is hi + lo
TWOFLOAT = np.dtype([("hi", "<f4"), ("lo", "<f4")])
_NP2T = {np.dtype(np.float16): FLOAT16, np.dtype(np.float32): FLOAT32,
         np.dtype(np.float64): FLOAT64, TWOFLOAT: TWOFLOAT32}
_T2NP = {v: k for k, v in _NP2T.items()}
_NAME2T = {"float16": FLOAT16, "float32": FLOAT32, "float64": FLOAT64,
           "twofloat32": TWOFLOAT32}


class GrapheneError(RuntimeError):
    """Raised when a library call returns a non-zero status."""

    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class _Info(C.Structure):
    _fields_ = [("nrows", C.c_int64), ("nnz_off", C.c_int64),
                ("num_halo", C.c_int64), ("num_tiles", C.c_int32),
                ("max_colors", C.c_int32), ("num_color_entries", C.c_int64),
                ("num_upper_color_entries", C.c_int64),
                ("multicolor_recommended", C.c_int32),
                ("max_tile_rows", C.c_int32)]


class _Stats(C.Structure):
    _fields_ = [("initial_residual", C.c_double), ("final_residual", C.c_double),
                ("iterations", C.c_int64), ("inner_iterations", C.c_int64),
                ("converged", C.c_int32), ("singular", C.c_int32),
                ("seconds", C.c_double), ("name", C.c_char * 96)]


class _ReqResult(C.Structure):
    _fields_ = [("latency", C.c_double), ("wait", C.c_double), ("service", C.c_double),
                ("iterations", C.c_int64), ("inner_iterations", C.c_int64),
                ("rel_residual", C.c_double), ("converged", C.c_int32), ("pad", C.c_int32)]


_lib = None

_P = C.c_void_p
_I64P = C.POINTER(C.c_int64)
_SIGS = {
    "gph_last_error": ([], C.c_char_p),
    "gph_version": ([], C.c_int32),
    "gph_generate_poisson3d": ([C.c_int64] * 3 + [_I64P, _P, _P, _P], C.c_int),
    "gph_generate_convdiff3d27": ([C.c_int64] * 3 + [C.c_double, C.c_uint64, _I64P, _P, _P, _P], C.c_int),
    "gph_generate_convdiff2d9": ([C.c_int64] * 2 + [C.c_double, C.c_double, C.c_uint64, _I64P, _P, _P, _P], C.c_int),
    "gph_generate_powerlaw": ([C.c_int64, C.c_int32, C.c_int32, C.c_double, C.c_int32, C.c_double, C.c_uint64, _I64P, _P, _P, _P], C.c_int),
    "gph_uniform_pm1": ([C.c_uint64, C.c_int64, _P], C.c_int),
    "gph_partition_grid3d": ([C.c_int64] * 6 + [_P, C.POINTER(C.c_int32)], C.c_int),
    "gph_partition_graph": ([C.c_int64, C.c_int64, _P, _P, C.c_int32, _P], C.c_int),
    "gph_host_matrix_from_coo": ([C.c_int64, C.c_int64, _P, _P, _P, C.c_int32, _P, C.POINTER(_P)], C.c_int),
    "gph_host_matrix_poisson3d": ([C.c_int64] * 3 + [C.c_int32, _P, C.POINTER(_P)], C.c_int),
    "gph_host_matrix_convdiff3d27": ([C.c_int64] * 3 + [C.c_double, C.c_uint64, C.c_int32, _P, C.POINTER(_P)], C.c_int),
    "gph_host_matrix_powerlaw": ([C.c_int64, C.c_int32, C.c_int32, C.c_double, C.c_int32, C.c_double, C.c_uint64, C.c_int32, _P, C.POINTER(_P)], C.c_int),
    "gph_host_matrix_load_mtx": ([C.c_char_p, C.c_int32, _P, C.POINTER(_P)], C.c_int),
    "gph_host_matrix_destroy": ([_P], None),
    "gph_host_matrix_get_info": ([_P, C.POINTER(_Info)], C.c_int),
    "gph_host_matrix_export": ([_P] * 12, C.c_int),
    "gph_host_matrix_export_levels": ([_P] * 5, C.c_int),
    "gph_host_matrix_export_ldu": ([_P, _I64P] + [_P] * 7, C.c_int),
    "gph_host_matrix_to_local": ([_P, C.c_int, _P, _P], C.c_int),
    "gph_host_matrix_to_global": ([_P, C.c_int, _P, _P], C.c_int),
    "gph_device_count": ([C.POINTER(C.c_int32)], C.c_int),
    "gph_malloc": ([C.c_size_t, C.POINTER(_P)], C.c_int),
    "gph_free": ([_P], C.c_int),
    "gph_memcpy_htod": ([_P, _P, C.c_size_t], C.c_int),
    "gph_memcpy_dtoh": ([_P, _P, C.c_size_t], C.c_int),
    "gph_memset": ([_P, C.c_int, C.c_size_t], C.c_int),
    "gph_synchronize": ([], C.c_int),
    "gph_vector_cast": ([C.c_int, _P, C.c_int, _P, C.c_int64], C.c_int),
    "gph_matrix_create": ([_P, C.c_int, C.c_int32, C.POINTER(_P)], C.c_int),
    "gph_matrix_destroy": ([_P], None),
    "gph_matrix_set_sliced": ([_P, C.c_int32], C.c_int),
    "gph_matrix_set_stream": ([_P, _P, C.c_int32], C.c_int),
    "gph_stream_create": ([C.POINTER(_P)], C.c_int),
    "gph_stream_destroy": ([_P], C.c_int),
    "gph_stream_synchronize": ([_P], C.c_int),
    "gph_matrix_device_bytes": ([_P, C.POINTER(C.c_size_t)], C.c_int),
    "gph_spmv": ([_P, C.c_int, C.c_int, _P, C.c_int, _P, C.c_int], C.c_int),
    "gph_residual": ([_P, C.c_int, C.c_int, _P, C.c_int, _P, C.c_int, _P, C.c_int, _P], C.c_int),
    "gph_vector_norm": ([_P, C.c_int, _P, C.POINTER(C.c_double)], C.c_int),
    "gph_vector_norm_noalloc": ([_P, C.c_int, _P, C.POINTER(C.c_double)], C.c_int),
    "gph_vector_norm_kind": ([_P, C.c_int, C.c_int, _P, C.POINTER(C.c_double)], C.c_int),
    "gph_bench_spmv": ([_P, C.c_int, C.c_int, _P, _P, C.c_int32, C.POINTER(C.c_double)], C.c_int),
    "gph_solver_create": ([_P, C.c_char_p, C.POINTER(_P)], C.c_int),
    "gph_solver_destroy": ([_P], None),
    "gph_solve": ([_P, C.c_int, _P, C.c_int, _P, C.POINTER(_Stats)], C.c_int),
    "gph_bench_apply": ([_P, C.c_int, _P, _P, C.c_int32, C.POINTER(C.c_double)], C.c_int),
    "gph_solver_name": ([_P, C.c_char_p, C.c_size_t], C.c_int),
    "gph_solver_uses_initial_guess": ([_P, C.POINTER(C.c_int32)], C.c_int),
    "gph_config_check": ([C.c_char_p, C.c_char_p, C.c_size_t], C.c_int),
    "gph_server_create": ([C.c_int32, C.c_size_t, C.POINTER(_P)], C.c_int),
    "gph_batch_create": ([_P, C.c_int32, _P, C.c_char_p, C.POINTER(_P)], C.c_int),
    "gph_batch_destroy": ([_P], None),
    "gph_batch_create_rhs": ([_P, C.c_int32, C.c_char_p, C.POINTER(_P)], C.c_int),
    "gph_batch_info": ([_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                        C.POINTER(C.c_int32)], C.c_int),
    "gph_batch_solve": ([_P, C.c_int, _P, _P, _P, C.POINTER(C.c_double)], C.c_int),
    "gph_server_destroy": ([_P], None),
    "gph_server_add_system": ([_P, _P, C.c_char_p, C.POINTER(C.c_int32)], C.c_int),
    "gph_server_footprint": ([_P, C.c_int32, C.POINTER(C.c_size_t)], C.c_int),
    "gph_server_resident_bytes": ([_P, C.POINTER(C.c_size_t)], C.c_int),
    "gph_server_run": ([_P, C.c_int64, _P, _P, _P, _P, _P, C.POINTER(_ReqResult)], C.c_int),
}


def lib():
    """Loads libgraphene_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m graphene_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _check(code):
    if code != 0:
        raise GrapheneError(code, lib().gph_last_error().decode())


def _ptr(a):
    return a.ctypes.data_as(_P) if a is not None else None


def exported_symbols():
    """Names declared in include/graphene_amd.h that the library must export."""
    return sorted(_SIGS)


def device_count():
    n = C.c_int32(0)
    _check(lib().gph_device_count(C.byref(n)))
    return n.value


# ----------------------------------------------------------------------------
# generators and partitions
# ----------------------------------------------------------------------------
def _triplets(fn, *args):
    cnt = C.c_int64(0)
    _check(fn(*args, C.byref(cnt), None, None, None))
    n = cnt.value
    rows = np.empty(n, np.int64)
    cols = np.empty(n, np.int64)
    vals = np.empty(n, np.float32)
    _check(fn(*args, C.byref(cnt), _ptr(rows), _ptr(cols), _ptr(vals)))
    return rows, cols, vals


def generate3DPoissonTriplets(nx, ny, nz):
    """generate3DPoissonTripletMatrix (lg/matrix/host/details/Poisson.cpp:26-69)."""
    return _triplets(lib().gph_generate_poisson3d, nx, ny, nz)


def generateConvectionDiffusion27Triplets(nx, ny, nz, peclet=10.0, seed=1):
    return _triplets(lib().gph_generate_convdiff3d27, nx, ny, nz, float(peclet), int(seed))


def generateConvectionDiffusion9Triplets(nx, ny, diffusion=1.0, peclet=10.0, seed=1):
    return _triplets(lib().gph_generate_convdiff2d9, nx, ny, float(diffusion),
                     float(peclet), int(seed))


POWERLAW_DEFAULTS = dict(min_len=5, max_len=64, exponent=1.85, band=2000,
                         local_fraction=0.9)


def generatePowerLawTriplets(n, seed=2, **kw):
    p = dict(POWERLAW_DEFAULTS, **kw)
    return _triplets(lib().gph_generate_powerlaw, n, p["min_len"], p["max_len"],
                     float(p["exponent"]), p["band"], float(p["local_fraction"]),
                     int(seed))


def uniform_pm1(seed, n):
    out = np.empty(n, np.float32)
    _check(lib().gph_uniform_pm1(int(seed), n, _ptr(out)))
    return out


def partition_grid3d(nx, ny, nz, bx, by, bz):
    r2t = np.empty(nx * ny * nz, np.int32)
    T = C.c_int32(0)
    _check(lib().gph_partition_grid3d(nx, ny, nz, bx, by, bz, _ptr(r2t), C.byref(T)))
    return r2t, T.value


def partition_graph(nrows, rows, cols, num_tiles):
    """Deterministic graph partition (recursive BFS bisection of the
    symmetrized pattern, weights row length + 1): the row_to_tile array."""
    r = np.ascontiguousarray(rows, np.int64)
    c = np.ascontiguousarray(cols, np.int64)
    out = np.empty(nrows, np.int32)
    _check(lib().gph_partition_graph(int(nrows), len(r), _ptr(r), _ptr(c), int(num_tiles),
                                     _ptr(out)))
    return out


# ----------------------------------------------------------------------------
# host matrices
# ----------------------------------------------------------------------------
class HostMatrix:
    """Prepared host matrix (CRSHostMatrix, CRSHostMatrix.cpp:63-85)."""

    def __init__(self, handle):
        self._h = handle
        info = _Info()
        _check(lib().gph_host_matrix_get_info(self._h, C.byref(info)))
        self.info = info
        self.nrows = info.nrows
        self.nnz_off = info.nnz_off
        self.num_tiles = info.num_tiles

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.gph_host_matrix_destroy(self._h)
            self._h = None

    @staticmethod
    def from_coo(nrows, rows, cols, vals, numTiles=1, rowToTile=None):
        rows = np.ascontiguousarray(rows, np.int64)
        cols = np.ascontiguousarray(cols, np.int64)
        vals = np.ascontiguousarray(vals, np.float32)
        r2t = None if rowToTile is None else np.ascontiguousarray(rowToTile, np.int32)
        h = _P()
        _check(lib().gph_host_matrix_from_coo(nrows, len(vals), _ptr(rows), _ptr(cols),
                                              _ptr(vals), numTiles, _ptr(r2t), C.byref(h)))
        return HostMatrix(h)

    def export(self):
        n, nnz, T, nh = self.nrows, self.nnz_off, self.num_tiles, self.info.num_halo
        o = dict(tile_start=np.empty(T + 1, np.int64), new_to_old=np.empty(n, np.int64),
                 num_interior=np.empty(T, np.int32),
                 num_separator_regions=np.empty(T, np.int32),
                 row_ptr=np.empty(n + 1, np.int64), col=np.empty(nnz, np.int32),
                 col_local=np.empty(nnz, np.int32), off=np.empty(nnz, np.float32),
                 diag=np.empty(n, np.float32), halo_start=np.empty(T + 1, np.int64),
                 halo_global=np.empty(max(nh, 1), np.int64))
        keys = ["tile_start", "new_to_old", "num_interior", "num_separator_regions",
                "row_ptr", "col", "col_local", "off", "diag", "halo_start", "halo_global"]
        _check(lib().gph_host_matrix_export(self._h, *[_ptr(o[k]) for k in keys]))
        o["halo_global"] = o["halo_global"][:nh]
        return o

    def levels(self):
        n, T = self.nrows, self.num_tiles
        o = dict(level=np.empty(n, np.int32), color_sort_addr=np.empty(n, np.int32),
                 color_start=np.empty(self.info.num_color_entries, np.int64),
                 num_colors=np.empty(T, np.int32))
        _check(lib().gph_host_matrix_export_levels(
            self._h, _ptr(o["level"]), _ptr(o["color_sort_addr"]),
            _ptr(o["color_start"]), _ptr(o["num_colors"])))
        return o

    def ldu(self):
        nf = self.nnz_off // 2 + 1
        n = self.nrows
        F = C.c_int64(0)
        o = dict(lower_addr=np.empty(nf, np.int32), upper_addr=np.empty(nf, np.int32),
                 owner_start=np.empty(n + 1, np.int32),
                 neighbour_start=np.empty(n + 1, np.int32),
                 neighbour_col_ind=np.empty(nf, np.int32), upper=np.empty(nf, np.float32),
                 lower=np.empty(nf, np.float32))
        keys = ["lower_addr", "upper_addr", "owner_start", "neighbour_start",
                "neighbour_col_ind", "upper", "lower"]
        _check(lib().gph_host_matrix_export_ldu(self._h, C.byref(F),
                                                *[_ptr(o[k]) for k in keys]))
        for k in ["lower_addr", "upper_addr", "neighbour_col_ind", "upper", "lower"]:
            o[k] = o[k][:F.value]
        o["num_faces"] = F.value
        return o

    def to_local(self, v):
        v = np.ascontiguousarray(v)
        out = np.empty_like(v)
        _check(lib().gph_host_matrix_to_local(self._h, _NP2T[v.dtype], _ptr(v), _ptr(out)))
        return out

    def to_global(self, v):
        v = np.ascontiguousarray(v)
        out = np.empty_like(v)
        _check(lib().gph_host_matrix_to_global(self._h, _NP2T[v.dtype], _ptr(v), _ptr(out)))
        return out

    def copyToTile(self, fmt=CRS, rowsPerBlock=0):
        """HostMatrix::copyToTile (lg/matrix/host/HostMatrix.hpp:53)."""
        return Matrix(self, fmt, rowsPerBlock)


def _host(fn, *args, numTiles=1, rowToTile=None):
    r2t = None if rowToTile is None else np.ascontiguousarray(rowToTile, np.int32)
    h = _P()
    _check(fn(*args, numTiles, _ptr(r2t), C.byref(h)))
    return HostMatrix(h)


def generate3DPoissonMatrix(nx, ny, nz, numTiles=1, rowToTile=None):
    """host::generate3DPoissonMatrix (lg/matrix/host/HostMatrix.hpp:92-96)."""
    return _host(lib().gph_host_matrix_poisson3d, nx, ny, nz, numTiles=numTiles,
                 rowToTile=rowToTile)


def generateConvectionDiffusion27Matrix(nx, ny, nz, peclet=10.0, seed=1, numTiles=1,
                                        rowToTile=None):
    return _host(lib().gph_host_matrix_convdiff3d27, nx, ny, nz, float(peclet),
                 int(seed), numTiles=numTiles, rowToTile=rowToTile)


def generatePowerLawMatrix(n, seed=2, numTiles=1, rowToTile=None, **kw):
    p = dict(POWERLAW_DEFAULTS, **kw)
    return _host(lib().gph_host_matrix_powerlaw, n, p["min_len"], p["max_len"],
                 float(p["exponent"]), p["band"], float(p["local_fraction"]), int(seed),
                 numTiles=numTiles, rowToTile=rowToTile)


def loadMatrixFromFile(path, numTiles=1, rowToTile=None):
    """host::loadMatrixFromFile (lg/matrix/host/HostMatrix.hpp:88-91)."""
    return _host(lib().gph_host_matrix_load_mtx, os.fsencode(path), numTiles=numTiles,
                 rowToTile=rowToTile)


# ----------------------------------------------------------------------------
# device memory, matrices, solvers
# ----------------------------------------------------------------------------
class DeviceVector:
    """A device vector in the matrix's tile-major order (the reference's Tensor)."""

    def __init__(self, n, dtype):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.type = _NP2T[self.dtype]
        p = _P()
        _check(lib().gph_malloc(max(self.n, 1) * self.dtype.itemsize, C.byref(p)))
        self.ptr = p

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.gph_free(self.ptr)
            self.ptr = None

    @staticmethod
    def from_numpy(a):
        a = np.ascontiguousarray(a)
        v = DeviceVector(a.size, a.dtype)
        v.upload(a)
        return v

    def upload(self, a):
        a = np.ascontiguousarray(a, self.dtype)
        assert a.size == self.n
        _check(lib().gph_memcpy_htod(self.ptr, _ptr(a), a.nbytes))

    def numpy(self):
        out = np.empty(self.n, self.dtype)
        _check(lib().gph_memcpy_dtoh(_ptr(out), self.ptr, out.nbytes))
        return out

    def fill(self, value):
        if self.dtype == TWOFLOAT:
            a = np.zeros(self.n, TWOFLOAT)
            a["hi"] = value
            self.upload(a)
        else:
            self.upload(np.full(self.n, value, self.dtype))

    def cast(self, dtype):
        """Expression::cast (libgraphene/dsl/tensor/Expression.hpp:152) on the device."""
        out = DeviceVector(self.n, dtype)
        _check(lib().gph_vector_cast(self.type, self.ptr, out.type, out.ptr, self.n))
        return out


class Stream:
    """A HIP stream (gph_stream_*), for ordering a matrix's work after other
    work on the same stream."""

    def __init__(self):
        p = _P()
        _check(lib().gph_stream_create(C.byref(p)))
        self.ptr = p

    def synchronize(self):
        _check(lib().gph_stream_synchronize(self.ptr))

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.gph_stream_destroy(self.ptr)
            self.ptr = None


class SolverStats:
    """lg/matrix/solver/SolverStats.hpp:29-69."""

    def __init__(self, s):
        self.initialResidual = s.initial_residual
        self.finalResidual = s.final_residual
        self.iterations = s.iterations
        self.innerIterations = s.inner_iterations
        self.converged = bool(s.converged)
        self.singular = bool(s.singular)
        self.seconds = s.seconds
        self.name = s.name.decode()

    def __repr__(self):
        return (f"SolverStats({self.name}: r0={self.initialResidual:.6g}, "
                f"r={self.finalResidual:.6g}, it={self.iterations}, "
                f"inner={self.innerIterations}, converged={self.converged}, "
                f"singular={self.singular}, {self.seconds*1e3:.3f} ms)")


def _cfg_text(config):
    return config if isinstance(config, str) else _json.dumps(config)


def check_config(config):
    """solver::Configuration::fromJSON without a device; returns the solver name."""
    buf = C.create_string_buffer(128)
    _check(lib().gph_config_check(_cfg_text(config).encode(), buf, 128))
    return buf.value.decode()


class Solver:
    """Configuration::fromJSON + Solver::createSolver; solve(x, b)."""

    def __init__(self, matrix, config):
        self.matrix = matrix
        h = _P()
        _check(lib().gph_solver_create(matrix._A, _cfg_text(config).encode(), C.byref(h)))
        self._S = h

    def __del__(self):
        if getattr(self, "_S", None) and _lib is not None:
            _lib.gph_solver_destroy(self._S)
            self._S = None

    def name(self):
        buf = C.create_string_buffer(128)
        _check(lib().gph_solver_name(self._S, buf, 128))
        return buf.value.decode()

    def usesInitialGuess(self):
        o = C.c_int32()
        _check(lib().gph_solver_uses_initial_guess(self._S, C.byref(o)))
        return bool(o.value)

    def bench_apply(self, rhs, out, iters=20):
        """Mean device ms of one preconditioner application out = M^-1 rhs."""
        ms = C.c_double()
        _check(lib().gph_bench_apply(self._S, rhs.type, rhs.ptr, out.ptr, iters,
                                     C.byref(ms)))
        return ms.value

    def solve(self, x, b):
        st = _Stats()
        _check(lib().gph_solve(self._S, x.type, x.ptr, b.type, b.ptr, C.byref(st)))
        return SolverStats(st)


class Matrix:
    """Device matrix (lg/matrix/Matrix.hpp:47-147)."""

    def __init__(self, host, fmt=CRS, rowsPerBlock=0):
        self.host = host
        self.fmt = fmt
        h = _P()
        _check(lib().gph_matrix_create(host._h, fmt, rowsPerBlock, C.byref(h)))
        self._A = h
        self.nrows = host.nrows

    def __del__(self):
        if getattr(self, "_A", None) and _lib is not None:
            _lib.gph_matrix_destroy(self._A)
            self._A = None

    def numTiles(self):
        return self.host.num_tiles

    def set_sliced(self, sliced=True):
        """Row-gather kernel form: sliced ELL or LDS row blocks (without a
        call: chosen per matrix, row blocks for irregular rows with far
        columns)."""
        _check(lib().gph_matrix_set_sliced(self._A, int(bool(sliced))))

    def getFormat(self):
        return self.fmt

    def device_bytes(self):
        b = C.c_size_t()
        _check(lib().gph_matrix_device_bytes(self._A, C.byref(b)))
        return b.value

    def set_stream(self, stream, synchronous=True):
        """Enqueue this matrix's work (and its solvers') on `stream` (a Stream
        or None for the matrix's own); synchronous=False: spmv and residual
        return once enqueued."""
        _check(lib().gph_matrix_set_stream(self._A, stream.ptr if stream else None,
                                           int(bool(synchronous))))

    def createUninitializedVector(self, dtype=np.float32, withHalo=False):
        """Matrix::createUninitializedVector (halo cells need no storage here)."""
        return DeviceVector(self.nrows, dtype)

    def spmv(self, x, y=None, acc=None, value_type=FLOAT32):
        y = y or DeviceVector(self.nrows, x.dtype)
        acc = acc or x.type
        _check(lib().gph_spmv(self._A, value_type, x.type, x.ptr, y.type, y.ptr, acc))
        return y

    def __mul__(self, x):
        """Matrix::operator* (lg/matrix/Matrix.hpp:95-97)."""
        return self.spmv(x)

    def residual(self, x, b, r=None, destType=None, intermediateType=None,
                 value_type=FLOAT32, r64=None):
        """Matrix::residual (lg/matrix/Matrix.hpp:85-93)."""
        dest = destType or x.type
        inter = intermediateType or dest
        r = r or DeviceVector(self.nrows, _T2NP[dest])
        _check(lib().gph_residual(self._A, value_type, x.type, x.ptr, b.type, b.ptr,
                                  r.type, r.ptr, inter, r64.ptr if r64 else None))
        return r

    def vectorNorm(self, x, norm="L2"):
        """Matrix::vectorNorm (lg/matrix/Matrix.hpp:115-117): "L2", "L1"
        (sum of magnitudes) or "LINF" (largest magnitude)."""
        out = C.c_double()
        if norm == "L2":
            _check(lib().gph_vector_norm(self._A, x.type, x.ptr, C.byref(out)))
        else:
            code = {"L1": 1, "LINF": 2}[norm]
            _check(lib().gph_vector_norm_kind(self._A, code, x.type, x.ptr, C.byref(out)))
        return out.value

    def bench_spmv(self, x, y, iters=50, value_type=FLOAT32):
        ms = C.c_double()
        _check(lib().gph_bench_spmv(self._A, value_type, x.type, x.ptr, y.ptr, iters,
                                    C.byref(ms)))
        return ms.value

    def solve(self, x, b, config):
        """Matrix::solve (lg/matrix/Matrix.cpp:26-30)."""
        return Solver(self, config).solve(x, b)


class Batch:
    """Batched solves (include/graphene_amd.h, gph_batch_*): the systems are
    consecutive tile ranges of one block-diagonal device matrix
    (system_tile_start, S+1 entries), solved together from x = 0."""

    def __init__(self, matrix, system_tile_start, config, n_rhs=None):
        """n_rhs: instead of system tile ranges, n_rhs right-hand sides of the
        one system `matrix` (vectors of n_rhs * nrows elements, block j =
        right-hand side j; gph_batch_create_rhs)."""
        h = _P()
        if n_rhs is not None:
            _check(lib().gph_batch_create_rhs(matrix._A, int(n_rhs),
                                              _cfg_text(config).encode(), C.byref(h)))
        else:
            st = np.ascontiguousarray(system_tile_start, np.int32)
            _check(lib().gph_batch_create(matrix._A, len(st) - 1, _ptr(st),
                                          _cfg_text(config).encode(), C.byref(h)))
        self._h = h
        self._A = matrix
        n, ref, rpw = C.c_int32(), C.c_int32(), C.c_int32()
        _check(lib().gph_batch_info(h, C.byref(n), C.byref(ref), C.byref(rpw)))
        self.num_systems = n.value
        self.refinement = bool(ref.value)
        self.rows_per_wave = rpw.value

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.gph_batch_destroy(self._h)
            self._h = None

    def solve(self, x, b):
        """x: float64 (refinement) or float32 DeviceVector, overwritten; b:
        float32. Returns (per-system SolverStats list, seconds)."""
        stats = (_Stats * self.num_systems)()
        sec = C.c_double()
        _check(lib().gph_batch_solve(self._h, x.type, x.ptr, b.ptr, stats, C.byref(sec)))
        return [SolverStats(stats[k]) for k in range(self.num_systems)], sec.value


def batch_union(systems, block=(32, 32)):
    """Block-diagonal union of 2D systems for Batch: systems is a list of
    (nx, ny, rows, cols, vals) triplets (natural order per system); every
    system is tiled into bx x by blocks. Returns (HostMatrix,
    system_tile_start)."""
    bx, by = block
    rows_l, cols_l, vals_l, r2t_l, tile_start = [], [], [], [], [0]
    off = 0
    for nx, ny, r, c, v in systems:
        r2t, T = partition_grid3d(nx, ny, 1, bx, by, 1)
        rows_l.append(np.asarray(r, np.int64) + off)
        cols_l.append(np.asarray(c, np.int64) + off)
        vals_l.append(np.asarray(v, np.float32))
        r2t_l.append(r2t + tile_start[-1])
        tile_start.append(tile_start[-1] + T)
        off += nx * ny
    h = HostMatrix.from_coo(off, np.concatenate(rows_l), np.concatenate(cols_l),
                            np.concatenate(vals_l), tile_start[-1], np.concatenate(r2t_l))
    return h, np.asarray(tile_start, np.int32)


class Server:
    """Request server with continuous batching (include/graphene_amd.h,
    gph_server_*): resident systems, requests admitted by HBM footprint,
    kernels of concurrent requests overlapping on the device."""

    def __init__(self, workers=4, hbm_budget_bytes=64 << 30):
        h = _P()
        _check(lib().gph_server_create(int(workers), int(hbm_budget_bytes), C.byref(h)))
        self._h = h
        self._keep = []

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.gph_server_destroy(self._h)
            self._h = None

    def add_system(self, matrix, config):
        sid = C.c_int32()
        _check(lib().gph_server_add_system(self._h, matrix._A, _cfg_text(config).encode(),
                                           C.byref(sid)))
        self._keep.append(matrix)
        return sid.value

    def footprint(self, sid):
        b = C.c_size_t()
        _check(lib().gph_server_footprint(self._h, int(sid), C.byref(b)))
        return b.value

    def resident_bytes(self):
        b = C.c_size_t()
        _check(lib().gph_server_resident_bytes(self._h, C.byref(b)))
        return b.value

    def run(self, system, n_rhs, arrival, seed, keep_solution=None):
        """Runs the requests; returns (results dict of arrays, solutions).
        keep_solution: optional list of (request index, n_rows) whose first
        right-hand side's solution is returned (tile-major order)."""
        n = len(system)
        sysa = np.ascontiguousarray(system, np.int32)
        rhsa = np.ascontiguousarray(n_rhs, np.int32)
        arra = np.ascontiguousarray(arrival, np.float64)
        seeda = np.ascontiguousarray(seed, np.uint64)
        res = (_ReqResult * max(n, 1))()
        sols = {}
        xo = None
        if keep_solution:
            xo = (C.c_void_p * n)()
            for i, rows in keep_solution:
                sols[i] = np.zeros(rows, np.float64)
                xo[i] = sols[i].ctypes.data
        _check(lib().gph_server_run(self._h, n, _ptr(sysa), _ptr(rhsa), _ptr(arra),
                                    _ptr(seeda), C.cast(xo, _P) if xo is not None else None,
                                    res))
        out = {k: np.array([getattr(res[i], k) for i in range(n)])
               for k in ("latency", "wait", "service", "iterations", "inner_iterations",
                         "rel_residual", "converged")}
        return out, sols
