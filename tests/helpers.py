"""Shared test inputs: product host matrices paired with the oracle layout of
the same matrix (built independently by the oracle from the same triplets)."""
import numpy as np

import graphene_amd as g
import oracle as o


def oracle_layout(n, rows, cols, vals, T=1, rowToTile=None):
    rp, col, off, dg = o.coo_to_crs(n, rows, cols, vals)
    r2t = np.zeros(n, np.int32) if rowToTile is None else np.asarray(rowToTile, np.int32)
    return o.tile_layout(n, rp, col, off, dg, T, r2t)


def poisson_case(nx, ny, nz, block=None):
    """Poisson-7 matrix: (product host matrix, oracle layout)."""
    n = nx * ny * nz
    rows, cols, vals = o.poisson7(nx, ny, nz)
    if block is None:
        h = g.generate3DPoissonMatrix(nx, ny, nz)
        L = oracle_layout(n, rows, cols, vals)
    else:
        r2t, T = g.partition_grid3d(nx, ny, nz, *block)
        h = g.generate3DPoissonMatrix(nx, ny, nz, numTiles=T, rowToTile=r2t)
        L = oracle_layout(n, rows, cols, vals, T, r2t)
    return h, L


def convdiff_case(nx, block=None, peclet=10.0, seed=1, ny=None, nz=None):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    n = nx * ny * nz
    rows, cols, vals = g.generateConvectionDiffusion27Triplets(nx, ny, nz, peclet, seed)
    if block is None:
        h = g.HostMatrix.from_coo(n, rows, cols, vals)
        L = oracle_layout(n, rows, cols, vals)
    else:
        r2t, T = g.partition_grid3d(nx, ny, nz, *block)
        h = g.HostMatrix.from_coo(n, rows, cols, vals, T, r2t)
        L = oracle_layout(n, rows, cols, vals, T, r2t)
    return h, L


def graph_partition_case(nx, T):
    """27-point convection-diffusion on an nx^3 grid split by the library's
    graph partitioner (the METIS stand-in) into T tiles."""
    n = nx ** 3
    rows, cols, vals = g.generateConvectionDiffusion27Triplets(nx, nx, nx, 10.0, 3)
    r2t = g.partition_graph(n, rows, cols, T)
    h = g.HostMatrix.from_coo(n, rows, cols, vals, T, r2t)
    L = oracle_layout(n, rows, cols, vals, T, r2t)
    return h, L


def powerlaw_case(n, T=1, seed=2):
    rows, cols, vals = g.generatePowerLawTriplets(n, seed=seed)
    h = g.HostMatrix.from_coo(n, rows, cols, vals, T)
    e = h.export()
    # the contiguous partition of the product is an input of the oracle layout
    r2t = np.empty(n, np.int32)
    ts = e["tile_start"]
    for t in range(T):
        r2t[e["new_to_old"][ts[t]:ts[t + 1]]] = t
    L = oracle_layout(n, rows, cols, vals, T, r2t)
    return h, L


def rhs_for(h, L, seed=0):
    """x_true ~ U[-1,1] (natural order) and b = fl32(A x_true in FP64), both
    returned in tile-major order."""
    n = h.nrows
    xt = o.uniform_pm1(seed, n)
    xl = xt[L["new_to_old"]]
    b = o.spmv(L, xl.astype(np.float64)).astype(np.float32)
    return xl, b
