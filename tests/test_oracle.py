"""Pins of the CPU oracle (CPU only). Each check is independent of the
oracle's own formulas: closed forms, scipy/numpy library routines, textbook
dense algorithms on tiny inputs, and the reference's known-answer values."""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as o
from helpers import oracle_layout

HERE = os.path.dirname(os.path.abspath(__file__))


def to_scipy(L):
    """Oracle tile-major layout -> scipy CSR incl. diagonal (float64)."""
    n = len(L["diag"])
    rp = L["row_ptr"]
    rows = np.repeat(np.arange(n), np.diff(rp))
    A = sp.csr_matrix((L["off"].astype(np.float64), (rows, L["col"])), shape=(n, n))
    return (A + sp.diags(L["diag"].astype(np.float64))).tocsr()


def poisson_layout(nx, ny, nz, T=1, r2t=None):
    n = nx * ny * nz
    r, c, v = o.poisson7(nx, ny, nz)
    return oracle_layout(n, r, c, v, T, r2t)


# --- Poisson generator: closed form + Kronecker-sum construction ------------
@pytest.mark.parametrize("dims", [(32, 32, 32), (5, 3, 4), (1, 1, 7)])
def test_poisson_closed_form(dims):
    nx, ny, nz = dims
    r, c, v = o.poisson7(nx, ny, nz)
    nnz = nx * ny * nz + 2 * ((nx - 1) * ny * nz + nx * (ny - 1) * nz + nx * ny * (nz - 1))
    assert len(v) == nnz
    if dims == (32, 32, 32):
        assert nnz == 223232  # the 32^3 Poisson benchmark matrix

    def lap1(m):
        return sp.diags([-np.ones(m - 1), 2 * np.ones(m), -np.ones(m - 1)], [-1, 0, 1])

    I = sp.identity
    K = (sp.kron(I(nz), sp.kron(I(ny), lap1(nx))) + sp.kron(I(nz), sp.kron(lap1(ny), I(nx)))
         + sp.kron(lap1(nz), sp.kron(I(ny), I(nx))))
    A = sp.csr_matrix((v.astype(np.float64), (r, c)), shape=K.shape)
    # Dirichlet boundary rows of the reference keep the full diagonal 6
    Dfix = sp.diags(6.0 - K.diagonal())
    assert abs(A - (K + Dfix)).max() == 0


# --- COO -> CRS vs scipy (duplicates summed) ---------------------------------
def test_coo_to_crs_vs_scipy():
    rng = np.random.default_rng(0)
    n, m = 50, 600
    rows = rng.integers(0, n, m)
    cols = rng.integers(0, n, m)
    vals = rng.standard_normal(m).astype(np.float32)
    rp, col, off, dg = o.coo_to_crs(n, rows, cols, vals)
    S = sp.coo_matrix((vals.astype(np.float64), (rows, cols)), shape=(n, n)).tocsr()
    S.sum_duplicates()
    S.sort_indices()
    assert np.allclose(dg, S.diagonal(), atol=1e-5)
    offd = S - sp.diags(S.diagonal())
    offd.eliminate_zeros()
    # structure: every (row, col) off-diagonal pair, sorted by row then col
    pairs = set(zip(np.repeat(np.arange(n), np.diff(rp)).tolist(), col.tolist()))
    R, Cc = S.nonzero()
    want = {(a, b) for a, b in zip(R.tolist(), Cc.tolist()) if a != b}
    assert pairs == want
    for i in range(n):
        assert np.all(np.diff(col[rp[i]:rp[i + 1]]) > 0)
    dense = np.zeros((n, n))
    dense[np.repeat(np.arange(n), np.diff(rp)), col] = off
    assert np.allclose(dense, offd.toarray(), atol=1e-5)


# --- SpMV: FP64 scipy with a Higham bound; sequential float32 on tiny input --
def test_spmv_f32_higham_bound():
    L = poisson_layout(12, 10, 9)
    A = to_scipy(L)
    x = o.uniform_pm1(3, A.shape[0])
    y = o.spmv(L, x).astype(np.float64)
    exact = A @ x.astype(np.float64)
    k = np.diff(L["row_ptr"]) + 1
    gamma = k * 2.0 ** -24 / (1 - k * 2.0 ** -24)
    bound = gamma * (abs(A) @ abs(x.astype(np.float64)))
    assert np.all(np.abs(y - exact) <= bound + 1e-30)


def test_spmv_f32_sequential_order_tiny():
    rng = np.random.default_rng(1)
    n = 20
    rows, cols = np.nonzero(rng.random((n, n)) < 0.3)
    vals = rng.standard_normal(len(rows)).astype(np.float32)
    L = oracle_layout(n, rows.astype(np.int64), cols.astype(np.int64), vals)
    x = rng.standard_normal(n).astype(np.float32)
    y = o.spmv(L, x)
    A = np.zeros((n, n), np.float32)
    A[rows, cols] += vals
    for i in range(n):  # textbook: diagonal term first, then ascending columns
        s = np.float32(A[i, i]) * x[i]
        for j in range(n):
            if j != i and A[i, j] != 0:
                s = np.float32(s + np.float32(A[i, j] * x[j]))
        assert s == y[i]


# --- tile layout invariants ---------------------------------------------------
def test_tile_layout_invariants():
    nx = 8
    n = nx ** 3
    rng = np.random.default_rng(2)
    r2t = rng.integers(0, 5, n).astype(np.int32)
    L = poisson_layout(nx, nx, nx, 5, r2t)
    # every non-halo row exactly once, tiles contiguous and own their rows
    assert sorted(L["new_to_old"]) == list(range(n))
    ts = L["tile_start"]
    for t in range(5):
        assert np.all(r2t[L["new_to_old"][ts[t]:ts[t + 1]]] == t)
    # halo of a tile == its foreign columns
    A = to_scipy(poisson_layout(nx, nx, nx))
    for t in range(5):
        mine = np.nonzero(r2t == t)[0]
        foreign = set(A[mine].indices.tolist()) - set(mine.tolist())
        h = L["halo_global"][L["halo_start"][t]:L["halo_start"][t + 1]]
        assert set(h.tolist()) == foreign and len(h) == len(foreign)
    # same operator after the permutation
    P = L["new_to_old"]
    Ap = to_scipy(L)
    assert abs(Ap - A[P][:, P]).max() == 0


# --- level sets ---------------------------------------------------------------
@pytest.mark.parametrize("m", [8, 16, 32])
def test_levels_closed_form_poisson(m):
    L = poisson_layout(m, m, m)
    lv = o.levels(m ** 3, 1, L["tile_start"], L["row_ptr"], L["col"])
    assert lv["num_colors"][0] == 3 * (m - 1) + 1   # 94 at 32^3


def test_levels_closed_form_27pt():
    import graphene_amd as g
    m = 10
    r, c, v = g.generateConvectionDiffusion27Triplets(m, m, m)
    L = oracle_layout(m ** 3, r, c, v)
    lv = o.levels(m ** 3, 1, L["tile_start"], L["row_ptr"], L["col"])
    assert lv["num_colors"][0] == 7 * (m - 1) + 1


def test_levels_valid_layering():
    nx = 8
    rng = np.random.default_rng(3)
    r2t = rng.integers(0, 3, nx ** 3).astype(np.int32)
    L = poisson_layout(nx, nx, nx, 3, r2t)
    lv = o.levels(nx ** 3, 3, L["tile_start"], L["row_ptr"], L["col"])
    ts, rp, col, level = L["tile_start"], L["row_ptr"], L["col"], lv["level"]
    for t in range(3):
        for p in range(ts[t], ts[t + 1]):
            lower = [j for j in col[rp[p]:rp[p + 1]] if ts[t] <= j < p]
            if lower:
                assert level[p] == 1 + max(level[j] for j in lower)
            else:
                assert level[p] == 0


# --- LDU == CRS ---------------------------------------------------------------
def test_ldu_equals_crs():
    L = poisson_layout(6, 5, 4)
    U = o.ldu_build(L)
    x = np.arange(L["diag"].size, dtype=np.float32) % 7 - 3  # small integers: exact
    assert np.array_equal(o.ldu_spmv_f32(U, x), o.spmv(L, x))
    assert U["num_faces"] * 2 == len(L["off"])


# --- ILU(0) -------------------------------------------------------------------
def test_ilu0_pattern_identity():
    import graphene_amd as g
    m = 6
    r, c, v = g.generateConvectionDiffusion27Triplets(m, m, m, 10.0, 3)
    L = oracle_layout(m ** 3, r, c, v)
    lu, invD = o.ilu0_factor(L, np.float64)
    n = m ** 3
    A = to_scipy(L).toarray()
    Lm = np.eye(n)
    Um = np.diag(1.0 / invD)
    rows = np.repeat(np.arange(n), np.diff(L["row_ptr"]))
    for i, j, val in zip(rows, L["col"], lu):
        if j < i:
            Lm[i, j] = val
        else:
            Um[i, j] = val
    LU = Lm @ Um
    mask = A != 0
    assert np.allclose(LU[mask], A[mask], rtol=1e-12, atol=1e-12)


def test_ilu0_dense_brute_force():
    rng = np.random.default_rng(4)
    n = 12
    M = (rng.random((n, n)) < 0.35) * rng.standard_normal((n, n))
    np.fill_diagonal(M, 5 + rng.random(n))
    pattern = M != 0
    # textbook IKJ ILU(0) restricted to the pattern (Saad Alg. 10.4)
    W = M.copy()
    for i in range(1, n):
        for k in range(i):
            if not pattern[i, k]:
                continue
            W[i, k] = W[i, k] / W[k, k]
            for j in range(k + 1, n):
                if pattern[i, j]:
                    W[i, j] = W[i, j] - W[i, k] * W[k, j]
    rows, cols = np.nonzero(pattern)
    L = oracle_layout(n, rows.astype(np.int64), cols.astype(np.int64),
                      M[rows, cols].astype(np.float32))
    lu, invD = o.ilu0_factor(L, np.float64)
    Mf = M.astype(np.float32).astype(np.float64)
    W = Mf.copy()
    for i in range(1, n):
        for k in range(i):
            if not pattern[i, k]:
                continue
            W[i, k] = W[i, k] / W[k, k]
            for j in range(k + 1, n):
                if pattern[i, j]:
                    W[i, j] = W[i, j] - W[i, k] * W[k, j]
    rr = np.repeat(np.arange(n), np.diff(L["row_ptr"]))
    assert np.allclose(lu, W[rr, L["col"]], rtol=1e-13)
    assert np.allclose(1.0 / invD, np.diag(W), rtol=1e-13)


def test_ilu0_apply_vs_triangular_solves():
    L = poisson_layout(7, 6, 5)
    lu, invD = o.ilu0_factor(L, np.float64)
    n = len(invD)
    rows = np.repeat(np.arange(n), np.diff(L["row_ptr"]))
    lo = L["col"] < rows
    Lm = sp.csr_matrix((lu[lo], (rows[lo], L["col"][lo])), shape=(n, n)) + sp.identity(n)
    Um = sp.csr_matrix((lu[~lo], (rows[~lo], L["col"][~lo])), shape=(n, n)) + sp.diags(1 / invD)
    b = o.uniform_pm1(5, n).astype(np.float64)
    y = spla.spsolve_triangular(Lm.tocsr(), b, lower=True)
    x = spla.spsolve_triangular(Um.tocsr(), y, lower=False)
    assert np.allclose(o.ilu0_apply(L, lu, invD, b), x, rtol=1e-12, atol=1e-13)


def _dilu_dense_blocks(L, invD, tiled_off):
    """Dense (D+L) D^-1 (D+U) restricted to each tile's block (FP64)."""
    n = len(invD)
    ts = L["tile_start"]
    rows = np.repeat(np.arange(n), np.diff(L["row_ptr"]))
    M = np.zeros((n, n))
    for t in range(len(ts) - 1):
        a, b = ts[t], ts[t + 1]
        sel = (rows >= a) & (rows < b) & (L["col"] >= a) & (L["col"] < b)
        r, c, v = rows[sel], L["col"][sel], tiled_off[sel].astype(np.float64)
        D = np.diag(1.0 / invD[a:b])
        Lo = np.zeros((b - a, b - a))
        Up = np.zeros((b - a, b - a))
        lo = c < r
        Lo[r[lo] - a, c[lo] - a] = v[lo]
        Up[r[~lo] - a, c[~lo] - a] = v[~lo]
        M[a:b, a:b] = (D + Lo) @ np.diag(invD[a:b]) @ (D + Up)
    return M


@pytest.mark.parametrize("T,block", [(1, None), (8, (3, 3, 3))])
def test_dilu_diagonal_identity_and_apply(T, block):
    """DILU (ilu/Solver.cpp:191-216, :405-441): M = (D+L) D^-1 (D+U) with
    diag(M) = diag(A) per tile block; the apply solves M y = b."""
    import graphene_amd as g
    m = 6
    r, c, v = g.generateConvectionDiffusion27Triplets(m, m, m, 10.0, 3)
    n = m ** 3
    r2t = None
    if block is not None:
        r2t, T = g.partition_grid3d(m, m, m, *block)
    L = oracle_layout(n, r, c, v, T, r2t)
    lu, invD = o.ilu0_factor(L, np.float64, diagonalBased=True)
    assert np.array_equal(lu, L["off"].astype(np.float64))  # off-diagonals untouched
    M = _dilu_dense_blocks(L, invD, L["off"])
    assert np.allclose(np.diag(M), L["diag"].astype(np.float64), rtol=1e-12, atol=1e-12)
    b = o.uniform_pm1(9, n).astype(np.float64)
    y = o.ilu0_apply(L, lu, invD, b, diagonalBased=True)
    assert np.allclose(M @ y, b, rtol=1e-11, atol=1e-11)
    # FP32 factors/apply stay close to the FP64 ones
    lu32, invD32 = o.ilu0_factor(L, np.float32, diagonalBased=True)
    assert np.allclose(invD32, invD, rtol=1e-5)
    y32 = o.ilu0_apply(L, lu32, invD32, b.astype(np.float32), diagonalBased=True)
    assert np.allclose(y32, y, rtol=1e-4, atol=1e-5)


def test_dilu_dense_brute_force():
    """Textbook DILU recursion d_i = a_ii - sum_{k<i} a_ik a_ki / d_k."""
    rng = np.random.default_rng(11)
    n = 14
    M = (rng.random((n, n)) < 0.3) * rng.standard_normal((n, n))
    np.fill_diagonal(M, 6 + rng.random(n))
    pattern = M != 0
    rows, cols = np.nonzero(pattern)
    L = oracle_layout(n, rows.astype(np.int64), cols.astype(np.int64),
                      M[rows, cols].astype(np.float32))
    Mf = M.astype(np.float32).astype(np.float64)
    d = np.diag(Mf).copy()
    for i in range(n):
        for k in range(i):
            if pattern[i, k]:
                d[i] -= Mf[i, k] * Mf[k, i] / d[k]
    _, invD = o.ilu0_factor(L, np.float64, diagonalBased=True)
    assert np.allclose(1.0 / invD, d, rtol=1e-13)


# --- Gauss-Seidel ---------------------------------------------------------------
def test_gs_dense_textbook_and_fixed_point():
    L = poisson_layout(5, 4, 3)
    A = to_scipy(L).toarray()
    n = A.shape[0]
    b = o.uniform_pm1(6, n).astype(np.float64)
    x0 = o.uniform_pm1(7, n).astype(np.float64)
    # forward sweep, then backward sweep
    x = x0.copy()
    for i in range(n):
        x[i] = (b[i] - A[i, :i] @ x[:i] - A[i, i + 1:] @ x[i + 1:]) / A[i, i]
    assert np.allclose(o.gs_sweep(L, b, x0), x, rtol=1e-13, atol=1e-14)
    xb = x.copy()
    for i in reversed(range(n)):
        xb[i] = (b[i] - A[i, :i] @ xb[:i] - A[i, i + 1:] @ xb[i + 1:]) / A[i, i]
    assert np.allclose(o.gs_sweep(L, b, x, backward=True), xb, rtol=1e-13, atol=1e-14)
    y = x0.copy()
    for _ in range(400):
        y = o.gs_sweep(L, b, y)
    assert np.allclose(A @ y, b, atol=1e-10)


# --- reductions: reference known answers + the worker chunking ----------------
def test_reduction_kats():
    k = json.load(open(os.path.join(HERE, "golden", "tensor_kats.json")))
    v = np.arange(16, dtype=np.float32).reshape(8, 2)
    ts = np.array([0, 2, 4, 6, 8])
    got = [o.reduce_add(np.ascontiguousarray(v[:, c]), ts) for c in range(2)]
    assert got == k["reduce_dim0_4tiles"]["expected"]
    w = np.arange(64, dtype=np.float32).reshape(8, 8)
    got = [o.reduce_add(w[r], np.array([0, 8])) for r in range(8)]
    assert got == k["reduce_dim1"]["expected"]
    a = np.array(k["dot"]["a"], np.float32)
    b = np.array(k["dot"]["b"], np.float32)
    got = [o.reduce_add(a[i] * b[i], np.array([0, 3])) for i in range(2)]
    assert got == k["dot"]["expected"]
    q = np.array(k["square_norm_rows"]["v"], np.float32)
    got = [o.reduce_add(q[i], np.array([0, 3]), square=True) for i in range(4)]
    assert got == k["square_norm_rows"]["expected"]


def test_reduction_worker_chunks():
    # 48 elements on one tile -> 6 workers of 8 (Expression.cpp:472-485).
    # v = [2^24, 1 x 47]: worker 0 loses its 7 ones (2^24 + 1 rounds to 2^24),
    # workers 1..5 hold 8 each -> 2^24 + 40. A single sequential sum gives 2^24.
    v = np.ones(48, np.float32)
    v[0] = 2.0 ** 24
    assert o.reduce_add(v, np.array([0, 48])) == 2.0 ** 24 + 40
    # 6 or fewer elements: no worker split
    assert o.reduce_add(np.array([2.0 ** 24, 1, 1, 1], np.float32), np.array([0, 4])) == 2.0 ** 24


# --- Krylov solvers -------------------------------------------------------------
def test_bicgstab_matches_scipy():
    L = poisson_layout(10, 10, 10)
    A = to_scipy(L)
    b = o.uniform_pm1(8, A.shape[0]).astype(np.float64)
    S = o.System(L)
    x, st = S.pbicgstab(np.zeros_like(b), b, S.pbicgstab_cfg(1e-8, 500, "none",
                                                             working64=True))
    its = []
    xs, info = spla.bicgstab(A, b, rtol=1e-8, atol=0.0, maxiter=500,
                             callback=lambda xk: its.append(1))
    assert st.converged and info == 0
    assert abs(st.iterations - len(its)) <= 1
    assert np.allclose(x, xs, rtol=1e-6, atol=1e-8)


def test_pbicgstab_ilu_true_residual_and_mpir():
    L = poisson_layout(16, 16, 16)
    A = to_scipy(L)
    xt = o.uniform_pm1(9, A.shape[0])
    b = (A @ xt.astype(np.float64)).astype(np.float32)
    S = o.System(L)
    x, st = S.pbicgstab(np.zeros_like(b), b, S.pbicgstab_cfg(1e-6, 500))
    assert st.converged and not st.singular
    assert np.linalg.norm(b - A @ x.astype(np.float64)) <= 2e-6 * np.linalg.norm(b)
    phi, st, inner = S.mpir(b, 1e-10, 30)
    assert st.converged
    assert np.linalg.norm(b - A @ phi) <= 1.0001e-10 * np.linalg.norm(b)


def test_singular_rhs_sets_flag():
    L = poisson_layout(4, 4, 4)
    S = o.System(L)
    b = np.zeros(64, np.float32)
    x, st = S.pbicgstab(np.zeros_like(b), b, S.pbicgstab_cfg(1e-6, 50))
    assert st.singular and st.iterations == 0


# --- block (T > 1) branches -------------------------------------------------------
# 2D 5-point Poisson on a 4 x 3 grid (row = 4 j + i), three tiles:
#   j=2:  8  9 10 11      tiles  1 1 2 2
#   j=1:  4  5  6  7             0 1 1 2
#   j=0:  0  1  2  3             0 0 0 2
# Derived by hand from CRSHostMatrix.cpp:273-354 (a row with no foreign
# neighbour is interior; otherwise it joins the separator region keyed by
# its set of neighbour tiles, regions in first-encounter order while rows are
# visited ascending) and DistributedTileLayout.cpp:91-126 (interior rows,
# then each separator region's cells):
#   tile 0 {0,1,2,4}:  0 interior; 1 -> {1}; 2 -> {1,2}; 4 -> {1}
#                      local order 0 | 1 4 | 2
#   tile 1 {5,6,8,9}:  5 -> {0}; 6 -> {0,2}; 8 -> {0}; 9 -> {2}
#                      local order 5 8 | 6 | 9
#   tile 2 {3,7,10,11}: 3 -> {0}; 7 -> {1}; 10 -> {1}; 11 interior
#                      local order 11 | 3 | 7 10
# Halo cells: the foreign columns, taken from each source tile's separator
# regions in order (source tile ascending; the reference's halo region order
# follows thread timing, this order is the oracle's documented choice).
HAND_R2T = np.array([0, 0, 0, 2, 0, 1, 1, 2, 1, 1, 2, 2], np.int32)
HAND_NEW_TO_OLD = [0, 1, 4, 2, 5, 8, 6, 9, 11, 3, 7, 10]
HAND_TILE_START = [0, 4, 8, 12]
HAND_NUM_INTERIOR = [1, 0, 1]
HAND_NUM_REGIONS = [2, 3, 2]
HAND_HALO = [[5, 8, 6, 3], [1, 4, 2, 7, 10], [2, 6, 9]]


def hand_layout():
    r, c, v = o.poisson7(4, 3, 1)
    return oracle_layout(12, r, c, v, 3, HAND_R2T)


def test_tile_layout_hand_derived():
    L = hand_layout()
    assert L["new_to_old"].tolist() == HAND_NEW_TO_OLD
    assert L["tile_start"].tolist() == HAND_TILE_START
    assert L["num_interior"].tolist() == HAND_NUM_INTERIOR
    assert L["num_separator_regions"].tolist() == HAND_NUM_REGIONS
    hs = L["halo_start"]
    assert [L["halo_global"][hs[t]:hs[t + 1]].tolist() for t in range(3)] == HAND_HALO
    # local columns: in-tile local ids, then halo ids (n_t + position in the
    # halo list), ascending per row; e.g. global row 2 (tile 0, local 3) has
    # neighbours 1 (local 1), 6 (halo 2 -> local 6), 3 (halo 3 -> local 7)
    pos = {g: i for i, g in enumerate(HAND_NEW_TO_OLD)}
    p = pos[2]
    rp = L["row_ptr"]
    assert L["col_local"][rp[p]:rp[p + 1]].tolist() == [1, 6, 7]
    assert L["col"][rp[p]:rp[p + 1]].tolist() == [pos[1], pos[6], pos[3]]
    for i in range(12):
        assert np.all(np.diff(L["col_local"][rp[i]:rp[i + 1]]) > 0)


def _tiled_case(T_block=(3, 3, 3), m=6, seed=3):
    import graphene_amd as g
    r, c, v = g.generateConvectionDiffusion27Triplets(m, m, m, 10.0, seed)
    r2t, T = g.partition_grid3d(m, m, m, *T_block)
    return oracle_layout(m ** 3, r, c, v, T, r2t), T


def _block_factors(L, lu, invD):
    """Block-diagonal unit-lower L and upper U (diagonal 1/invD) from the
    in-tile factor entries (FP64 dense)."""
    n = len(invD)
    ts = L["tile_start"]
    rows = np.repeat(np.arange(n), np.diff(L["row_ptr"]))
    tile = np.searchsorted(ts, rows, side="right") - 1
    intile = (L["col"] >= ts[tile]) & (L["col"] < ts[tile + 1])
    Lm = np.eye(n)
    Um = np.diag(1.0 / invD)
    lo = intile & (L["col"] < rows)
    up = intile & (L["col"] > rows)
    Lm[rows[lo], L["col"][lo]] = lu[lo]
    Um[rows[up], L["col"][up]] = lu[up]
    return Lm, Um, intile


def test_ilu0_block_identity_T8():
    """Per tile, (L U)_ij = A_ij on the tile block's pattern; entries that
    couple to another tile (halo columns) are left as they are."""
    L, T = _tiled_case()
    assert T == 8
    lu, invD = o.ilu0_factor(L, np.float64)
    Lm, Um, intile = _block_factors(L, lu, invD)
    assert np.array_equal(lu[~intile], L["off"][~intile].astype(np.float64))
    A = to_scipy(L).toarray()
    ts = L["tile_start"]
    LU = Lm @ Um
    for t in range(T):
        a, b = ts[t], ts[t + 1]
        blk = A[a:b, a:b]
        mask = blk != 0
        assert np.allclose(LU[a:b, a:b][mask], blk[mask], rtol=1e-12, atol=1e-12), t
        # block-diagonal: no coupling between tiles in the factors
        assert not LU[a:b, :a].any() and not LU[a:b, b:].any()


def test_ilu0_block_apply_vs_triangular_solves_T8():
    L, T = _tiled_case(seed=5)
    lu, invD = o.ilu0_factor(L, np.float64)
    Lm, Um, _ = _block_factors(L, lu, invD)
    n = len(invD)
    b = o.uniform_pm1(21, n).astype(np.float64)
    y = spla.spsolve_triangular(sp.csr_matrix(Lm), b, lower=True)
    x = spla.spsolve_triangular(sp.csr_matrix(Um), y, lower=False)
    assert np.allclose(o.ilu0_apply(L, lu, invD, b), x, rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("random_partition", [False, True])
def test_gs_block_sweeps_read_halo_before_sweep(random_partition):
    """Block Gauss-Seidel (gauss-seidel/Solver.cpp:45-160): the halo is
    exchanged before the sweep, so each tile sweeps its rows in local order
    with the newest values of its own rows and the pre-sweep values of every
    other tile's rows."""
    m = 6
    n = m ** 3
    if random_partition:
        r, c, v = o.poisson7(m, m, m)
        r2t = np.random.default_rng(8).integers(0, 5, n).astype(np.int32)
        L = oracle_layout(n, r, c, v, 5, r2t)
    else:
        L, _ = _tiled_case(seed=7)
    A = to_scipy(L).toarray()
    ts = L["tile_start"]
    b = o.uniform_pm1(22, n).astype(np.float64)
    x0 = o.uniform_pm1(23, n).astype(np.float64)

    def sweep(xin, backward):
        x = xin.copy()
        for t in range(len(ts) - 1):
            a, e = ts[t], ts[t + 1]
            rng_ = range(e - 1, a - 1, -1) if backward else range(a, e)
            for i in rng_:
                s = b[i]
                for j in np.nonzero(A[i])[0]:
                    if j != i:
                        s -= A[i, j] * (x[j] if a <= j < e else xin[j])
                x[i] = s / A[i, i]
        return x

    xf = sweep(x0, False)
    assert np.allclose(o.gs_sweep(L, b, x0), xf, rtol=1e-13, atol=1e-14)
    xb = sweep(xf, True)
    assert np.allclose(o.gs_sweep(L, b, xf, backward=True), xb, rtol=1e-13, atol=1e-14)
    # a sweep that read the other tiles' updated values would differ
    assert not np.allclose(xf, _gs_global(A, b, x0))


def _gs_global(A, b, x0):
    x = x0.copy()
    for i in range(len(b)):
        x[i] = (b[i] - A[i, :i] @ x[:i] - A[i, i + 1:] @ x[i + 1:]) / A[i, i]
    return x


# --- vector norms -------------------------------------------------------------
def test_vector_norms_vs_numpy():
    """L2 and L1 (sum of magnitudes) in the reference's reduction order,
    LINF (largest magnitude) exactly; a plausible slip (signed sum, missing
    root) fails these."""
    rng = np.random.default_rng(12)
    v = rng.standard_normal(5000)
    ts = np.array([0, 1200, 3100, 5000])
    assert o.vec_norm(v, ts, "LINF") == np.max(np.abs(v))
    assert np.isclose(o.vec_norm(v, ts, "L1"), np.sum(np.abs(v)), rtol=1e-13)
    assert np.isclose(o.vec_norm(v, ts, "L2"), np.linalg.norm(v), rtol=1e-13)
    v32 = v.astype(np.float32)
    assert o.vec_norm(v32, ts, "LINF") == np.max(np.abs(v32))
    assert np.isclose(o.vec_norm(v32, ts, "L1"), np.sum(np.abs(v.astype(np.float32)),
                                                          dtype=np.float64), rtol=1e-5)


# --- double-word ("twofloat32") arithmetic: exact rational pins ---------------
from fractions import Fraction as Fr  # noqa: E402

U = Fr(1, 2 ** 24)  # unit roundoff of float32


def _fr(v):
    return Fr(float(v))


def _dw_of(d):
    """Normalized double words (hi, lo) of float64 values."""
    hi = d.astype(np.float32)
    return np.stack([hi, (d - hi.astype(np.float64)).astype(np.float32)], 1)


def _dw_samples(rng, m, cancel=False):
    """Pairs of double-word operands; with cancel, x + y nearly cancels."""
    d = rng.standard_normal(m) * 10.0 ** rng.integers(-8, 8, m)
    if cancel:
        e = -d * (1 + rng.standard_normal(m) * 10.0 ** rng.integers(-12, -2, m))
    else:
        e = rng.standard_normal(m) * 10.0 ** rng.integers(-8, 8, m)
    return _dw_of(d), _dw_of(e)


def test_dw_error_free_transforms():
    """2Sum and the FMA product are exact: hi + lo == a + b (a * b) in
    rational arithmetic, hi == fl32 of the operation."""
    rng = np.random.default_rng(5)
    a = (rng.standard_normal(400) * 10.0 ** rng.integers(-20, 20, 400)).astype(np.float32)
    b = (rng.standard_normal(400) * 10.0 ** rng.integers(-20, 20, 400)).astype(np.float32)
    for x, y in list(zip(a, b)) + [(1.0, 2.0 ** -30), (3.0, -3.0), (0.0, 0.0), (-1e30, 1e-30)]:
        x, y = np.float32(x), np.float32(y)
        s, e = o.dw_two_sum(x, y)
        assert _fr(s) + _fr(e) == _fr(x) + _fr(y)
        assert s == np.float32(np.float32(x) + np.float32(y))
        p, q = o.dw_two_prod(x, y)
        if np.isfinite(p) and abs(float(p)) > 1e-30:
            assert _fr(p) + _fr(q) == _fr(x) * _fr(y)


@pytest.mark.parametrize("cancel", [False, True])
def test_dw_add_within_bound(cancel):
    """Accurate double-word sum: |z - (x + y)| <= (3u^2 + 13u^3) |x + y| (also
    under cancellation, where the sloppy variant fails), result normalized."""
    rng = np.random.default_rng(7 + cancel)
    x, y = _dw_samples(rng, 500, cancel)
    z = o.dw_add(x, y)
    bound = 3 * U ** 2 + 13 * U ** 3
    for xi, yi, zi in zip(x, y, z):
        exact = _fr(xi[0]) + _fr(xi[1]) + _fr(yi[0]) + _fr(yi[1])
        got = _fr(zi[0]) + _fr(zi[1])
        assert abs(got - exact) <= bound * abs(exact)
        assert np.float32(zi[0] + zi[1]) == zi[0]  # |lo| <= ulp(hi) / 2


def test_dw_mul_within_bound():
    """Double-word product: |z - x y| <= 5u^2 |x y| (dropping a cross term
    gives errors of order u)."""
    rng = np.random.default_rng(9)
    x, y = _dw_samples(rng, 500)
    z = o.dw_mul(x, y)
    for xi, yi, zi in zip(x, y, z):
        exact = (_fr(xi[0]) + _fr(xi[1])) * (_fr(yi[0]) + _fr(yi[1]))
        got = _fr(zi[0]) + _fr(zi[1])
        assert abs(got - exact) <= 5 * U ** 2 * abs(exact)


def _dw_vec(rng, n):
    d = rng.uniform(-1, 1, n) / 3.0
    hi = d.astype(np.float32)
    return np.stack([hi, (d - hi.astype(np.float64)).astype(np.float32)], 1)


def test_dw_residual_vs_exact():
    """Double-word residual of a small convection-diffusion-like system in
    several tiles: within the double-word error bound of the exact rational
    residual (a float64-accumulated residual is u-accurate only when there is
    no cancellation; here b ~ A x so the residual cancels to ~1e-7 |terms|)."""
    rng = np.random.default_rng(3)
    L = poisson_layout(4, 3, 3, 2, np.repeat([0, 1], 18).astype(np.int32))
    n = len(L["diag"])
    x = _dw_vec(rng, n)
    xf = np.array([_fr(a) + _fr(b) for a, b in x], dtype=object)
    # b = fl32(A x) so that the residual is small relative to its terms
    b = o.spmv(L, x[:, 0].astype(np.float64) + x[:, 1]).astype(np.float32)
    r32, r64 = o.residual_dw(L, x, b)
    rp, col, off, dg = L["row_ptr"], L["col"], L["off"], L["diag"]
    for p in range(n):
        terms = [_fr(dg[p]) * xf[p]] + [_fr(off[a]) * xf[col[a]] for a in range(rp[p], rp[p + 1])]
        exact = _fr(b[p]) - sum(terms)
        scale = abs(_fr(b[p])) + sum(abs(t) for t in terms)
        k = len(terms) + 1
        assert abs(Fr(float(r64[p])) - exact) <= 2 * k * 5 * U ** 2 * scale + abs(exact) * 2 ** -53
        assert r32[p] == np.float32(r64[p]) or abs(_fr(r32[p]) - exact) <= abs(exact) * U * 2


def test_dw_ldu_equals_crs_on_exact_inputs():
    """LDU and CRS double-word residuals agree exactly when every operation is
    exact (small integers), so they differ only in summation order."""
    L = poisson_layout(4, 4, 3)
    n = len(L["diag"])
    rng = np.random.default_rng(1)
    x = np.stack([rng.integers(-8, 8, n).astype(np.float32), np.zeros(n, np.float32)], 1)
    b = rng.integers(-64, 64, n).astype(np.float32)
    U_ = o.ldu_build(L)
    assert np.array_equal(o.residual_dw(L, x, b)[1], o.ldu_residual_dw(U_, x, b)[1])
    ref = b.astype(np.float64) - o.spmv(L, x[:, 0].astype(np.float64))
    assert np.array_equal(o.residual_dw(L, x, b)[1], ref)


def test_mpir_twofloat_converges():
    """MPIR with twofloat32 extended precision reaches the tolerance in a
    float64 residual of its solution, with the outer count of the float64
    variant (+-1)."""
    L = poisson_layout(10, 10, 10)
    n = len(L["diag"])
    b = o.spmv(L, o.uniform_pm1(4, n).astype(np.float64)).astype(np.float32)
    S = o.System(L)
    phi, st, inner = S.mpir_dw(b, 1e-10, 30)
    assert st.converged
    x = phi[:, 0].astype(np.float64) + phi[:, 1]
    r = b.astype(np.float64) - o.spmv(L, x)
    assert np.linalg.norm(r) <= 1.01e-10 * st.initialResidual
    _, st64, _ = S.mpir(b, 1e-10, 30)
    assert abs(st.iterations - st64.iterations) <= 1
