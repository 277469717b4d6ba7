"""Product host pipeline (CPU, through the C-ABI) against the oracle:
CRS structure, partitions, tile layouts, level sets and LDU addressing are
bit-exact; configuration parsing follows the reference's rules."""
import glob
import os

import numpy as np
import pytest

import graphene_amd as g
import oracle as o
from helpers import (convdiff_case, graph_partition_case, oracle_layout, poisson_case,
                     powerlaw_case)

HERE = os.path.dirname(os.path.abspath(__file__))

CASES = {
    "poisson32_T1": lambda: poisson_case(32, 32, 32),
    "poisson12_blocks": lambda: poisson_case(12, 12, 12, block=(4, 6, 3)),
    "convdiff10_T1": lambda: convdiff_case(10),
    "convdiff12_blocks": lambda: convdiff_case(12, block=(4, 4, 6)),
    "powerlaw2000_T7": lambda: powerlaw_case(2000, T=7),
    "convdiff11_graph9": lambda: graph_partition_case(11, 9),
}


@pytest.fixture(scope="module", params=sorted(CASES))
def case(request):
    return (request.param,) + CASES[request.param]()


def test_structure_bit_exact(case):
    name, h, L = case
    e = h.export()
    for k in ["tile_start", "new_to_old", "num_interior", "num_separator_regions",
              "row_ptr", "col_local", "off", "diag", "halo_start", "halo_global"]:
        assert np.array_equal(e[k], L[k]), (name, k)
    assert np.array_equal(e["col"].astype(np.int64), L["col"]), name


def test_levels_bit_exact(case):
    name, h, L = case
    lv = h.levels()
    ol = o.levels(h.nrows, h.num_tiles, L["tile_start"], L["row_ptr"], L["col"])
    for k in ["level", "color_sort_addr", "color_start", "num_colors"]:
        assert np.array_equal(lv[k], ol[k]), (name, k)


def test_ldu_bit_exact(case):
    name, h, L = case
    if name.startswith("powerlaw"):
        with pytest.raises(g.GrapheneError):
            h.ldu()
        return
    U = h.ldu()
    R = o.ldu_build(L)
    assert U["num_faces"] == R["num_faces"]
    for k in ["lower_addr", "upper_addr", "owner_start", "neighbour_start",
              "neighbour_col_ind", "upper", "lower"]:
        assert np.array_equal(U[k].astype(np.int64) if U[k].dtype == np.int32 else U[k],
                              R[k]), (name, k)


def test_permutation_roundtrip(case):
    name, h, L = case
    v = np.arange(h.nrows, dtype=np.float64)
    loc = h.to_local(v)
    assert np.array_equal(loc, v[L["new_to_old"]])
    assert np.array_equal(h.to_global(loc), v)


def test_generators():
    # Poisson identical to the oracle restatement of Poisson.cpp
    r, c, v = g.generate3DPoissonTriplets(6, 5, 4)
    ro, co, vo = o.poisson7(6, 5, 4)
    assert np.array_equal(r, ro) and np.array_equal(c, co) and np.array_equal(v, vo)
    # 27-point convection-diffusion: counts, M-matrix sign pattern, row sums
    m = 128
    r, c, v = g.generateConvectionDiffusion27Triplets(9, 9, 9, 10.0, 1)
    assert len(v) == (3 * 9 - 2) ** 3  # 27-point pattern count on 9^3
    nnz128 = (3 * m - 2) ** 3
    assert nnz128 - m ** 3 == 53645816  # off-diagonals at 128^3
    diag = v[r == c]
    off = v[r != c]
    assert np.all(off < 0) and np.all(diag > 0)
    interior = np.zeros(9 ** 3, bool)
    idx = np.arange(9 ** 3)
    i, j, k = idx % 9, (idx // 9) % 9, idx // 81
    interior[(i > 0) & (i < 8) & (j > 0) & (j < 8) & (k > 0) & (k < 8)] = True
    rs = np.bincount(r, weights=v.astype(np.float64), minlength=9 ** 3)
    assert np.allclose(rs[interior], 0.0, atol=5e-5)  # zero row sum inside
    assert np.all(rs[~interior] > 0)                  # Dirichlet rows dominant
    # power law: lengths in [5, 64], distinct columns, diag = 1.1 sum|off|
    N = 50000
    r, c, v = g.generatePowerLawTriplets(N, seed=2)
    lens = np.bincount(r, minlength=N)
    assert lens.min() >= 5 and lens.max() <= 64 and 12 < lens.mean() < 16
    key = r * N + c
    assert len(np.unique(key)) == len(key)
    d = v[r == c]
    sabs = np.bincount(r[r != c], weights=np.abs(v[r != c]).astype(np.float64), minlength=N)
    assert np.allclose(d, 1.1 * sabs, rtol=1e-6)
    near = np.abs(r[r != c] - c[r != c]) <= 2000
    # 90% drawn within +-2000, plus the uniform draws that land there (~0.8%)
    assert 0.89 < near.mean() < 0.92
    # RNG identical to the oracle's
    assert np.array_equal(g.uniform_pm1(7, 1000), o.uniform_pm1(7, 1000))


def test_mtx_loader_roundtrip(tmp_path):
    import scipy.io
    import scipy.sparse as sp
    A = sp.random(30, 30, density=0.2, random_state=1, format="coo") + sp.identity(30)
    S = (A + A.T).tocoo()
    p = tmp_path / "a.mtx"
    scipy.io.mmwrite(str(p), S, symmetry="symmetric")
    h = g.loadMatrixFromFile(str(p))
    e = h.export()
    dense = np.zeros((30, 30), np.float32)
    rows = np.repeat(np.arange(30), np.diff(e["row_ptr"]))
    dense[rows, e["col"]] = e["off"]
    dense[np.arange(30), np.arange(30)] = e["diag"]
    assert np.allclose(dense, S.toarray().astype(np.float32))


def test_empty_tile_compaction():
    # rowToTile with unused tile ids: empty tiles are dropped (CRSHostMatrix.cpp:238-264)
    r2t = np.repeat(np.array([0, 3, 5], np.int32), 64 // 3 + 1)[:64]
    h = g.generate3DPoissonMatrix(4, 4, 4, numTiles=6, rowToTile=r2t)
    assert h.num_tiles == 3


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(HERE, "golden", "configs", "*.jsonc"))))
def test_reference_configs_parse(path):
    import json
    text = open(path).read()
    cfg_name = None
    # the solver subtree of each benchmark config (benchmark.cpp:163-183)
    start = text.find('"solver"')
    if start < 0:
        pytest.skip("no solver section (spmv benchmark)")
    name = g.check_config(_solver_subtree(text))
    assert name


def _solver_subtree(text):
    i = text.find('"solver"')
    j = text.find("{", i)
    depth = 0
    for k in range(j, len(text)):
        if text[k] == "{":
            depth += 1
        elif text[k] == "}":
            depth -= 1
            if depth == 0:
                return text[j:k + 1]
    raise ValueError


def test_config_rules():
    assert g.check_config({"type": "PBiCGStab", "preconditioner": {"type": "ILU"}}) == \
        "PBiCGStab(ILU(0))"
    assert g.check_config({"type": "IterativeRefinement", "innerSolver": {"type": "GaussSeidel"}}) == \
        "IterativeRefinement"
    assert g.check_config('// comment\n{"type": "ILU", /* c */ "unknownKey": 3}') == "ILU(0)"
    for bad, msg in [({"type": "IterativeRefinement"}, "requires an innerSolver"),
                     ({"type": "restarter"}, "requires an innerSolver"),
                     ({}, "No solver type"),
                     ({"type": "PBiCGStab", "norm": "L3"}, "Invalid norm"),
                     ({"type": "PBiCGStab", "workingType": "int3"}, "Unknown type"),
                     ({"type": "ILU", "solveMulticolor": "maybe"}, "MultiColorMode")]:
        with pytest.raises(g.GrapheneError) as e:
            g.check_config(bad)
        assert msg in str(e.value)
    assert g.check_config({"type": "restarter", "innerSolver": {"type": "PBiCGStab"}}) == \
        "Restarter of BiCGStab"


def test_graph_partition_properties():
    """The graph partitioner (METIS stand-in): every tile non-empty, weights
    (row length + 1) balanced, a small edge cut on a grid, deterministic,
    input errors reported."""
    nx = 48
    r, c, v = g.generateConvectionDiffusion9Triplets(nx, nx, 1.0, 5.0, 1)
    n = nx * nx
    p = g.partition_graph(n, r, c, 12)
    assert p.min() == 0 and p.max() == 11
    w = np.bincount(r, minlength=n) + 1
    tw = np.bincount(p, weights=w, minlength=12)
    assert tw.max() <= 1.05 * tw.mean() and tw.min() >= 0.95 * tw.mean()
    offd = r != c
    assert np.mean(p[r[offd]] != p[c[offd]]) < 0.15
    assert np.array_equal(p, g.partition_graph(n, r, c, 12))
    # one tile per row, and disconnected pieces
    assert np.array_equal(np.sort(g.partition_graph(5, [0, 1, 2, 3, 4], [0, 1, 2, 3, 4], 5)),
                          np.arange(5))
    with pytest.raises(g.GrapheneError):
        g.partition_graph(4, [0], [0], 5)
    with pytest.raises(g.GrapheneError):
        g.partition_graph(4, [0], [7], 2)



def test_tile_layout_hand_derived_fixture():
    """The product's tile layout of the hand-derived 3-tile 4 x 3 grid
    (tests/test_oracle.py holds the derivation from CRSHostMatrix.cpp:273-354
    and DistributedTileLayout.cpp:91-126)."""
    from test_oracle import (HAND_HALO, HAND_NEW_TO_OLD, HAND_NUM_INTERIOR, HAND_NUM_REGIONS,
                             HAND_R2T, HAND_TILE_START)
    h = g.generate3DPoissonMatrix(4, 3, 1, numTiles=3, rowToTile=HAND_R2T)
    e = h.export()
    assert e["new_to_old"].tolist() == HAND_NEW_TO_OLD
    assert e["tile_start"].tolist() == HAND_TILE_START
    assert e["num_interior"].tolist() == HAND_NUM_INTERIOR
    assert e["num_separator_regions"].tolist() == HAND_NUM_REGIONS
    hs = e["halo_start"]
    assert [e["halo_global"][hs[t]:hs[t + 1]].tolist() for t in range(3)] == HAND_HALO


@pytest.mark.parametrize("reverse", [False, True])
def test_structure_beyond_one_work_chunk(reverse):
    """The host pipeline works in chunks on worker threads (1 Mi triplets,
    16 Ki rows, tiles and tile parts): a 100,000-row power-law matrix in 16
    tiles spans several chunks of each; bit-exact against the oracle layout
    with the triplets in row order and reversed (the bucketing path), levels
    too."""
    n, T = 100000, 16
    rows, cols, vals = g.generatePowerLawTriplets(n, seed=4)
    if reverse:
        rows, cols, vals = rows[::-1].copy(), cols[::-1].copy(), vals[::-1].copy()
    assert len(vals) > (1 << 20)
    h = g.HostMatrix.from_coo(n, rows, cols, vals, T)
    e = h.export()
    r2t = np.empty(n, np.int32)
    for t in range(T):
        r2t[e["new_to_old"][e["tile_start"][t]:e["tile_start"][t + 1]]] = t
    L = oracle_layout(n, rows, cols, vals, T, r2t)
    for k in ["tile_start", "new_to_old", "num_interior", "num_separator_regions",
              "row_ptr", "col_local", "off", "diag", "halo_start", "halo_global"]:
        assert np.array_equal(e[k], L[k]), k
    assert np.array_equal(e["col"].astype(np.int64), L["col"])
    lv = h.levels()
    ol = o.levels(n, T, L["tile_start"], L["row_ptr"], L["col"])
    for k in ["level", "color_sort_addr", "color_start", "num_colors"]:
        assert np.array_equal(lv[k], ol[k]), k


def test_errors_from_worker_threads():
    """Errors raised inside the threaded pipeline reach the caller: an
    out-of-range triplet in the last of several 1 Mi-triplet chunks, and the
    LDU faces of a pattern that is not structurally symmetric (rows handled
    by several workers)."""
    n = 100000
    rows, cols, vals = g.generatePowerLawTriplets(n, seed=4)
    bad = cols.copy()
    bad[-1] = n + 5
    with pytest.raises(g.GrapheneError) as e:
        g.HostMatrix.from_coo(n, rows, bad, vals, 16)
    assert "out of range" in str(e.value)
    h = g.HostMatrix.from_coo(n, rows, cols, vals, 16)
    with pytest.raises(g.GrapheneError) as e:
        h.ldu()
    assert "symmetric" in str(e.value)
