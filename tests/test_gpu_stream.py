"""Caller-provided streams (gph_matrix_set_stream): a matrix's work and its
solvers' on the caller's stream give the same bits as on the matrix's own
stream; with synchronous = 0, SpMV and residual return once enqueued and are
ordered on the stream."""
import numpy as np
import pytest

import graphene_amd as g
import oracle as o
from helpers import poisson_case, rhs_for

pytestmark = pytest.mark.gpu


def test_spmv_and_residual_on_caller_stream():
    h, L = poisson_case(16, 16, 16, block=(8, 8, 8))
    A = h.copyToTile()
    x = g.DeviceVector.from_numpy(o.uniform_pm1(3, h.nrows))
    ref = o.spmv(L, x.numpy())
    st = g.Stream()
    A.set_stream(st, synchronous=False)
    try:
        y = g.DeviceVector(h.nrows, np.float32)
        z = g.DeviceVector(h.nrows, np.float32)
        A.spmv(x, y)          # enqueued only
        A.spmv(y, z)          # ordered after the first on the same stream
        b = g.DeviceVector.from_numpy(o.uniform_pm1(4, h.nrows))
        r = A.residual(y, b)  # r = b - A y
        st.synchronize()
        assert np.array_equal(y.numpy(), ref)
        assert np.array_equal(z.numpy(), o.spmv(L, ref))
        assert np.array_equal(r.numpy(), o.residual_f32(L, ref, b.numpy()))
    finally:
        A.set_stream(None)
    # back on the matrix's own stream, synchronous again
    assert np.array_equal((A * x).numpy(), ref)


def test_solve_on_caller_stream_equals_own_stream():
    h, L = poisson_case(16, 16, 16, block=(8, 8, 8))
    A = h.copyToTile()
    _, b = rhs_for(h, L, seed=2)
    cfg = {"type": "PBiCGStab", "relResidual": 1e-6, "maxIterations": 200,
           "preconditioner": {"type": "ILU"}}
    outs = []
    for use in (False, True):
        st = g.Stream() if use else None
        A.set_stream(st, synchronous=True)
        x = g.DeviceVector.from_numpy(np.zeros(h.nrows, np.float32))
        s = g.Solver(A, cfg).solve(x, g.DeviceVector.from_numpy(b))
        outs.append((x.numpy(), s.iterations))
        A.set_stream(None)
    assert outs[0][1] == outs[1][1] and np.array_equal(outs[0][0], outs[1][0])
