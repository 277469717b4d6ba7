"""Several right-hand sides of one system solved as a block
(gph_batch_create_rhs): the matrix streams are read once per product for
all of them and each tile's ILU(0) substitution runs for all of them in one
launch, while every right-hand side keeps its own Krylov scalars, dot
products and convergence. A right-hand side's result in a block is the same,
bit for bit, as the same right-hand side solved alone, and agrees with the
single-system solver to the convergence tolerance."""
import numpy as np
import pytest

import graphene_amd as g
import oracle as o
from helpers import convdiff_case, powerlaw_case

pytestmark = pytest.mark.gpu

INNER = {"type": "PBiCGStab", "relResidual": 1e-3, "maxIterations": 100,
         "preconditioner": {"type": "ILU"}}
MPIR = {"type": "IterativeRefinement", "relResidual": 1e-10, "maxIterations": 30,
        "extendedPrecisionType": "float64", "workingType": "float32", "innerSolver": INNER}

CASES = {
    "convdiff16_crs": lambda: (convdiff_case(16, block=(16, 8, 8))[0], g.CRS),
    "convdiff16_ldu": lambda: (convdiff_case(16, block=(16, 8, 8))[0], g.LDU),
    "powerlaw20000": lambda: (powerlaw_case(20000, T=20)[0], g.CRS),
}


@pytest.fixture(scope="module", params=sorted(CASES))
def case(request):
    h, fmt = CASES[request.param]()
    return request.param, h, h.copyToTile(fmt)


def rhs_block(h, A, seeds):
    """b_j = fl32(A x_j), x_j ~ U[-1,1) from seed j (tile-major blocks)."""
    n = h.nrows
    bs = []
    for sd in seeds:
        xt = h.to_local(g.uniform_pm1(sd, n).astype(np.float64))
        bs.append(A.spmv(g.DeviceVector.from_numpy(xt)).numpy().astype(np.float32))
    return np.concatenate(bs)


def solve_block(A, h, b, k):
    bt = g.Batch(A, None, MPIR, n_rhs=k)
    assert bt.num_systems == k and bt.refinement
    x = g.DeviceVector(k * h.nrows, np.float64)
    stats, _ = bt.solve(x, g.DeviceVector.from_numpy(b))
    return x.numpy(), stats


def test_block_equals_alone_and_single_solver(case):
    name, h, A = case
    n, k = h.nrows, 5
    seeds = [101, 102, 103, 104, 105]
    b = rhs_block(h, A, seeds)
    x, stats = solve_block(A, h, b, k)
    assert all(s.converged for s in stats), stats
    for j in range(k):
        xa, sa = solve_block(A, h, b[j * n:(j + 1) * n], 1)
        assert np.array_equal(xa, x[j * n:(j + 1) * n]), (name, j)
        assert (sa[0].iterations, sa[0].innerIterations) == \
            (stats[j].iterations, stats[j].innerIterations)
        # true FP64 residual of the right-hand side
        r = A.residual(g.DeviceVector.from_numpy(xa), g.DeviceVector.from_numpy(b[j * n:(j + 1) * n]),
                       destType=g.FLOAT64, intermediateType=g.FLOAT64).numpy()
        assert np.linalg.norm(r) <= 1.0001e-10 * stats[j].initialResidual
        # the single-system solver: same algorithm, other reduction order
        xs = g.DeviceVector.from_numpy(np.zeros(n))
        ss = g.Solver(A, MPIR).solve(xs, g.DeviceVector.from_numpy(b[j * n:(j + 1) * n]))
        assert ss.converged and abs(ss.iterations - stats[j].iterations) <= 1
        err = np.linalg.norm(xs.numpy() - xa) / np.linalg.norm(xs.numpy())
        assert err < 1e-7, (name, j, err)


def test_block_of_eight_and_pbicgstab_only(case):
    name, h, A = case
    n = h.nrows
    b = rhs_block(h, A, range(200, 208))
    x, stats = solve_block(A, h, b, 8)
    assert all(s.converged for s in stats)
    cfg = dict(INNER, relResidual=1e-6, maxIterations=500)
    bt = g.Batch(A, None, cfg, n_rhs=3)
    assert not bt.refinement
    x32 = g.DeviceVector(3 * n, np.float32)
    st, _ = bt.solve(x32, g.DeviceVector.from_numpy(b[:3 * n]))
    assert all(s.converged and s.finalResidual < 1e-6 * s.initialResidual for s in st)


def test_rhs_count_is_checked(case):
    _, h, A = case
    for k in (0, 9):
        with pytest.raises(g.GrapheneError, match="right-hand sides"):
            g.Batch(A, None, MPIR, n_rhs=k)
