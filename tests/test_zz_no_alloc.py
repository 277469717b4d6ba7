"""gph_vector_norm_noalloc (reduction scratch in a static device array, no
allocation) against gph_vector_norm (scratch allocated with the matrix):
the same kernel and grid, so the same bits, on two calls in a row, the
second with a smaller input, in float32 and float64."""
import ctypes as C

import numpy as np
import pytest

import graphene_amd as g

pytestmark = pytest.mark.gpu


def norms(A, v):
    lib = g.lib()
    a, b = C.c_double(), C.c_double()
    g._check(lib.gph_vector_norm(A._A, v.type, v.ptr, C.byref(a)))
    g._check(lib.gph_vector_norm_noalloc(A._A, v.type, v.ptr, C.byref(b)))
    return a.value, b.value


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_noalloc_norm_equals_norm(dtype):
    big = g.generate3DPoissonMatrix(128, 128, 64).copyToTile()     # 1,048,576 rows
    small = g.generate3DPoissonMatrix(10, 9, 7).copyToTile()       # 630 rows
    for A, seed in ((big, 1), (small, 2)):
        x = g.uniform_pm1(seed, A.nrows).astype(dtype) * 3
        v = g.DeviceVector.from_numpy(x)
        ref, got = norms(A, v)
        assert got == ref, (A.nrows, got, ref)
        assert abs(got - np.linalg.norm(x.astype(np.float64))) <= 1e-5 * got
    # the big input again after the small one: the array's leftover contents
    # do not leak into the next call
    x = g.uniform_pm1(3, big.nrows).astype(dtype)
    ref, got = norms(big, g.DeviceVector.from_numpy(x))
    assert got == ref


def test_noalloc_norm_rejects_bad_input():
    A = g.generate3DPoissonMatrix(4, 4, 4).copyToTile()
    v = g.DeviceVector.from_numpy(np.ones(64, np.float16))
    out = C.c_double()
    assert g.lib().gph_vector_norm_noalloc(A._A, v.type, v.ptr, C.byref(out)) != 0
    assert g.lib().gph_vector_norm_noalloc(A._A, g.FLOAT32, None, C.byref(out)) != 0
