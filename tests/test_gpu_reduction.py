"""Fused grid reductions at full size: the dot products and norms that the
SpMV and vector kernels reduce in their last block (two-level arrival
counting) against numpy, for grids of 1024 (vector kernels), 2048 (sliced
SpMV, resident grid) and about 8200 blocks (row-block SpMV), on repeated
launches (the arrival counters are reset by the launches themselves)."""
import numpy as np
import pytest

import graphene_amd as g

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def poisson128():
    n = 128
    h = g.generate3DPoissonMatrix(n, n, n)
    A = h.copyToTile(g.CRS)
    rng = np.random.default_rng(11)
    b = rng.uniform(-1.0, 1.0, h.nrows).astype(np.float32)
    return h, A, b


@pytest.mark.parametrize("sliced", [True, False])
def test_fused_norms_full_size(poisson128, sliced):
    h, A, b = poisson128
    A.set_sliced(sliced)
    try:
        cfg = {"type": "PBiCGStab", "relResidual": 1e-30, "maxIterations": 3,
               "workingType": "float32"}
        S = g.Solver(A, cfg)
        ref = float(np.linalg.norm(b.astype(np.float64)))
        runs = []
        for _ in range(3):
            x = g.DeviceVector.from_numpy(np.zeros(h.nrows, np.float32))
            st = S.solve(x, g.DeviceVector.from_numpy(b))
            # ||b - A 0|| from the residual kernel's fused norm (FP32 sums)
            assert abs(st.initialResidual - ref) <= 1e-5 * ref, (st.initialResidual, ref)
            assert st.iterations == 3
            runs.append((st.initialResidual, st.finalResidual, x.numpy()))
        for r in runs[1:]:
            assert r[0] == runs[0][0] and r[1] == runs[0][1]
            assert np.array_equal(r[2], runs[0][2])
        # the recursive residual after 3 iterations is the fused norm of
        # updateXR's r; check it against the true residual of the iterate
        xs = runs[0][2].astype(np.float64)
        y = A.spmv(g.DeviceVector.from_numpy(xs)).numpy()
        true = float(np.linalg.norm(b.astype(np.float64) - y))
        assert abs(runs[0][1] - true) <= 1e-3 * true, (runs[0][1], true)
    finally:
        A.set_sliced(True)
