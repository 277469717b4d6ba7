"""Request server (continuous batching): every result is the same alone and
in a mix of overlapping requests, requests are admitted within the HBM
budget, and every right-hand side converges."""
import numpy as np
import pytest

import graphene_amd as g
from helpers import convdiff_case, poisson_case, powerlaw_case

pytestmark = pytest.mark.gpu

MPIR = {"type": "IterativeRefinement", "relResidual": 1e-10, "maxIterations": 30,
        "extendedPrecisionType": "float64", "workingType": "float32",
        "innerSolver": {"type": "PBiCGStab", "relResidual": 1e-3, "maxIterations": 100,
                        "workingType": "float32",
                        "preconditioner": {"type": "ILU", "factorType": "float32"}}}
PB32 = {"type": "PBiCGStab", "relResidual": 1e-6, "maxIterations": 1000,
        "workingType": "float32",
        "preconditioner": {"type": "GaussSeidel", "numFixedIterations": 2,
                           "symmetric": True}}


@pytest.fixture(scope="module")
def systems():
    mats = []
    h, _ = poisson_case(12, 12, 12)
    mats.append((h.copyToTile(g.CRS), MPIR, h.nrows))
    h, _ = convdiff_case(12, block=(4, 4, 4))
    mats.append((h.copyToTile(g.LDU), MPIR, h.nrows))
    h, _ = powerlaw_case(3000, T=6)
    mats.append((h.copyToTile(g.CRS), MPIR, h.nrows))
    mats.append((mats[2][0], PB32, h.nrows))
    return mats


def make_server(systems, workers=4, budget=64 << 30):
    S = g.Server(workers, budget)
    ids = [S.add_system(A, cfg) for A, cfg, _ in systems]
    return S, ids


def test_alone_equals_mix(systems):
    S, ids = make_server(systems[:3])
    rows = systems[1][2]
    alone, sol_a = S.run([ids[1]], [2], [0.0], [7], keep_solution=[(0, rows)])
    assert alone["converged"][0] == 1
    assert alone["rel_residual"][0] <= 1e-10
    rng = np.random.default_rng(3)
    n = 16
    sysv = rng.integers(0, 3, n)
    rhs = rng.integers(1, 4, n)
    arr = np.sort(rng.uniform(0, 0.05, n))
    seeds = rng.integers(1, 1 << 30, n)
    k = 5
    sysv[k], rhs[k], seeds[k] = ids[1], 2, 7
    mix, sol_m = S.run(sysv, rhs, arr, seeds, keep_solution=[(k, rows)])
    assert np.all(mix["converged"] == 1), mix
    assert np.all(mix["rel_residual"] <= 1e-10), mix["rel_residual"]
    assert np.array_equal(sol_a[0], sol_m[k])
    assert np.all(mix["latency"] >= mix["service"] - 1e-9)
    assert np.all(mix["wait"] >= -1e-9)


def test_budget_admits_one_at_a_time(systems):
    S, ids = make_server(systems[:3])
    fp = max(S.footprint(i) for i in ids)
    S1, ids1 = make_server(systems[:3], workers=4, budget=fp)
    n = 6
    sysv = [ids1[i % 3] for i in range(n)]
    res, _ = S1.run(sysv, [1] * n, [0.0] * n, list(range(1, n + 1)))
    assert np.all(res["converged"] == 1)
    # with room for one request at a time the services do not overlap
    fin = res["latency"]  # all arrived at 0: finish times
    start = fin - res["service"]
    order = np.argsort(start)
    assert np.all(start[order][1:] >= fin[order][:-1] - 1e-4)


def test_gauss_seidel_system(systems):
    S, ids = make_server(systems[3:])
    res, _ = S.run([ids[0]] * 3, [1, 2, 1], [0.0, 0.0, 0.001], [1, 2, 3])
    assert np.all(res["converged"] == 1)
    assert np.all(res["rel_residual"] <= 1e-6 * 1.01)


def test_request_errors(systems):
    S, ids = make_server(systems[:1])
    with pytest.raises(g.GrapheneError):
        S.run([5], [1], [0.0], [1])
    with pytest.raises(g.GrapheneError):
        S.run([ids[0]], [0], [0.0], [1])
    small = g.Server(1, 1024)
    sid = small.add_system(systems[0][0], MPIR)
    with pytest.raises(g.GrapheneError):
        small.run([sid], [1], [0.0], [1])


def test_requests_on_one_system_overlap_and_equal_alone(systems):
    """Concurrent requests on the same system run on slots of their own
    (block solves of their right-hand sides) and overlap on the device; each
    result equals the same request served alone."""
    S, ids = make_server(systems[1:2], workers=4)
    rows = systems[1][2]
    alone = {}
    for sd in (11, 12, 13, 14):
        _, sol = S.run([ids[0]], [3], [0.0], [sd], keep_solution=[(0, rows)])
        alone[sd] = sol[0]
    res, sol = S.run([ids[0]] * 4, [3] * 4, [0.0] * 4, [11, 12, 13, 14],
                     keep_solution=[(i, rows) for i in range(4)])
    assert np.all(res["converged"] == 1) and np.all(res["rel_residual"] <= 1e-10)
    for i, sd in enumerate((11, 12, 13, 14)):
        assert np.array_equal(sol[i], alone[sd])
    fin = res["latency"]
    start = fin - res["service"]
    # at least two services overlap in time
    order = np.argsort(start)
    assert np.any(start[order][1:] < fin[order][:-1])
