"""GPU parity: the gfx950 kernels (through the C-ABI) against the CPU oracle on
the same seeded inputs. Bit-exact where the reference's per-row order is
kept (SpMV, residual, ILU(0) factor/apply, Gauss-Seidel); iteration counts
within 5% and true residuals for the Krylov solvers."""
import numpy as np
import pytest

import graphene_amd as g
import oracle as o
from helpers import convdiff_case, poisson_case, powerlaw_case, rhs_for

pytestmark = pytest.mark.gpu

CASES = {
    "poisson16_T1": lambda: poisson_case(16, 16, 16),
    "poisson16_T8": lambda: poisson_case(16, 16, 16, block=(8, 8, 8)),
    "convdiff12_T27": lambda: convdiff_case(12, block=(4, 4, 4)),
    "powerlaw3000_T6": lambda: powerlaw_case(3000, T=6),
}


@pytest.fixture(scope="module", params=sorted(CASES))
def case(request):
    h, L = CASES[request.param]()
    A = h.copyToTile()
    return request.param, h, L, A


def dev(a):
    return g.DeviceVector.from_numpy(np.ascontiguousarray(a))


@pytest.mark.parametrize("sliced", [True, False])
def test_spmv_f32_bit_exact(case, sliced):
    name, h, L, A = case
    A.set_sliced(sliced)
    try:
        x = o.uniform_pm1(11, h.nrows)
        y = (A * dev(x)).numpy()
        ref = o.spmv(L, x)
        assert np.array_equal(y.view(np.uint32), ref.view(np.uint32)), name
        r64 = g.DeviceVector(h.nrows, np.float64)
        x64 = x.astype(np.float64) / 3.0
        b = o.uniform_pm1(14, h.nrows)
        r = A.residual(dev(x64), dev(b), destType=g.FLOAT32, intermediateType=g.FLOAT64,
                       r64=r64)
        ref32, ref64 = o.residual_mixed(L, x64, b)
        assert np.array_equal(r.numpy(), ref32) and np.array_equal(r64.numpy(), ref64)
    finally:
        A.set_sliced(True)


def test_spmv_fp16_values(case):
    """FP16-stored values: equal to the oracle run on the FP16-rounded values."""
    name, h, L, A = case
    x = o.uniform_pm1(21, h.nrows)
    y = A.spmv(dev(x), value_type=g.FLOAT16).numpy()
    L16 = dict(L, off=L["off"].astype(np.float16).astype(np.float32),
               diag=L["diag"].astype(np.float16).astype(np.float32))
    assert np.array_equal(y, o.spmv(L16, x)), name


def slice_column_spans(L):
    """Column span (max - min) of every 64-row slice of the tile-major CRS
    rows (slices of the unsorted sliced layout)."""
    rp, col, n = L["row_ptr"], L["col"], L["n"]
    spans = []
    for r0 in range(0, n, 64):
        c = col[rp[r0]:rp[min(n, r0 + 64)]]
        spans.append(int(c.max() - c.min()) if len(c) else 0)
    return np.array(spans)


WIDE_CASES = {
    # z-neighbour tiles 128 tiles (131072 rows) apart: slices on tile z-faces
    # need 32-bit columns, the others take 16-bit offsets
    "convdiff128x128x16_T256": lambda: convdiff_case(128, block=(16, 8, 8), ny=128, nz=16),
    # long-range columns: every slice 32-bit
    "powerlaw150000_T8": lambda: powerlaw_case(150000, T=8),
}


@pytest.mark.parametrize("name", sorted(WIDE_CASES))
def test_spmv_wide_and_narrow_slices(name):
    """Sliced SpMV with per-slice 16-bit column offsets and 32-bit columns:
    bit-exact against the oracle (CRS and, for the stencil, LDU; FP32 and
    FP64 values)."""
    h, L = WIDE_CASES[name]()
    if name.startswith("convdiff"):
        spans = slice_column_spans(L)
        assert (spans < 65536).any() and (spans >= 65536).any()
    A = h.copyToTile()
    x = o.uniform_pm1(31, h.nrows)
    y = (A * dev(x)).numpy()
    assert np.array_equal(y.view(np.uint32), o.spmv(L, x).view(np.uint32))
    x64 = x.astype(np.float64) * 1.0000001
    assert np.array_equal(A.spmv(dev(x64)).numpy(), o.spmv(L, x64))
    if name.startswith("convdiff"):
        A_ldu = h.copyToTile(g.LDU)
        U = o.ldu_build(L)
        assert np.array_equal((A_ldu * dev(x)).numpy(), o.ldu_spmv_f32(U, x))


def test_spmv_column_ordered_row_blocks():
    """An irregular matrix (power-law, 600,000 rows, 10 % uniformly random
    columns) runs the row-block kernel with its chunks in column order:
    bit-exact against the oracle with FP32, FP64 and FP16 values, FP64 x and
    the mixed residual, on the matrix's own stream and on a caller stream."""
    h, L = powerlaw_case(600000, T=600)
    A = h.copyToTile(g.CRS)
    x = o.uniform_pm1(41, h.nrows)
    x64 = x.astype(np.float64) * 1.0000001
    b = o.uniform_pm1(42, h.nrows)
    L16 = dict(L, off=L["off"].astype(np.float16).astype(np.float32),
               diag=L["diag"].astype(np.float16).astype(np.float32))
    ref32, ref64 = o.spmv(L, x), o.spmv(L, x64)
    refr32, refr64 = o.residual_mixed(L, x64 / 3.0, b)
    ref16 = o.spmv(L16, x)
    st = g.Stream()
    for stream in (None, st):
        A.set_stream(stream)
        assert np.array_equal((A * dev(x)).numpy().view(np.uint32), ref32.view(np.uint32))
        assert np.array_equal(A.spmv(dev(x64)).numpy(), ref64)
        assert np.array_equal(A.spmv(dev(x), value_type=g.FLOAT16).numpy(), ref16)
        r64 = g.DeviceVector(h.nrows, np.float64)
        r = A.residual(dev(x64 / 3.0), dev(b), destType=g.FLOAT32,
                       intermediateType=g.FLOAT64, r64=r64)
        assert np.array_equal(r.numpy(), refr32) and np.array_equal(r64.numpy(), refr64)
    A.set_stream(None)


def test_spmv_f64(case):
    name, h, L, A = case
    x = o.uniform_pm1(12, h.nrows).astype(np.float64) * 1.0000001
    y = A.spmv(dev(x)).numpy()
    ref = o.spmv(L, x)
    assert np.array_equal(y, ref), name


def test_residual_mixed_bit_exact(case):
    name, h, L, A = case
    x = o.uniform_pm1(13, h.nrows).astype(np.float64) / 3.0
    b = o.uniform_pm1(14, h.nrows)
    r64 = g.DeviceVector(h.nrows, np.float64)
    r = A.residual(dev(x), dev(b), destType=g.FLOAT32, intermediateType=g.FLOAT64, r64=r64)
    ref32, ref64 = o.residual_mixed(L, x, b)
    assert np.array_equal(r.numpy(), ref32), name
    assert np.array_equal(r64.numpy(), ref64), name


@pytest.mark.parametrize("sliced", [True, False])
def test_ldu_spmv_bit_exact(case, sliced):
    name, h, L, A = case
    if name.startswith("powerlaw"):
        pytest.skip("LDU needs a structurally symmetric pattern")
    A_ldu = h.copyToTile(g.LDU)
    A_ldu.set_sliced(sliced)
    U = o.ldu_build(L)
    x = o.uniform_pm1(15, h.nrows)
    y = (A_ldu * dev(x)).numpy()
    assert np.array_equal(y, o.ldu_spmv_f32(U, x)), name
    # same operator as CRS up to summation order
    yc = o.spmv(L, x.astype(np.float64))
    assert np.allclose(y, yc, rtol=1e-5, atol=1e-5)
    x64 = x.astype(np.float64) / 7.0
    b = o.uniform_pm1(16, h.nrows)
    r = A_ldu.residual(dev(x64), dev(b), destType=g.FLOAT32, intermediateType=g.FLOAT64)
    assert np.array_equal(r.numpy(), o.ldu_residual_mixed(U, x64, b)[0])


def _sub(sub):
    """ILU substitution settings."""
    return {"substitution": sub}


@pytest.mark.parametrize("sub", ["packed", "rows", "level", "block", "tile"])
@pytest.mark.parametrize("ftype", ["float32", "float64"])
def test_ilu0_apply_bit_exact(case, ftype, sub):
    """ILU(0) factorization + substitution equal the oracle's, with each
    substitution kernel (product slabs, lane-per-row level records,
    workgroup per tile over the level stream, row blocks)."""
    name, h, L, A = case
    dt = np.float32 if ftype == "float32" else np.float64
    S = g.Solver(A, {"type": "ILU", "factorType": ftype, **_sub(sub)})
    b = o.uniform_pm1(17, h.nrows).astype(dt)
    x = dev(np.zeros(h.nrows, dt))
    st = S.solve(x, dev(b))
    assert st.iterations == 1
    lu, d = o.ilu0_factor(L, dt)
    ref = o.ilu0_apply(L, lu, d, b)
    assert np.array_equal(x.numpy(), ref), name


@pytest.mark.parametrize("sub", ["packed", "rows", "level", "block"])
@pytest.mark.parametrize("ftype", ["float32", "float64"])
def test_dilu_apply_bit_exact(case, ftype, sub):
    """DILU (diagonalBased) factorization + substitution equal the oracle's."""
    name, h, L, A = case
    dt = np.float32 if ftype == "float32" else np.float64
    S = g.Solver(A, {"type": "ILU", "factorType": ftype, "diagonalBased": True,
                     **_sub(sub)})
    assert S.name() == "DILU"
    b = o.uniform_pm1(27, h.nrows).astype(dt)
    x = dev(np.zeros(h.nrows, dt))
    S.solve(x, dev(b))
    lu, d = o.ilu0_factor(L, dt, diagonalBased=True)
    ref = o.ilu0_apply(L, lu, d, b, diagonalBased=True)
    assert np.array_equal(x.numpy(), ref), name


@pytest.mark.parametrize("sub", ["packed", "rows", "level", "block"])
def test_ilu0_fp16_factors_bit_exact(case, sub):
    """FP16-stored factors: the FP32 off-diagonal factors rounded to FP16,
    the FP32 inverse diagonal, applied in FP32 (the level-record kernel has
    no FP16 form and runs the packed kernel)."""
    name, h, L, A = case
    S = g.Solver(A, {"type": "ILU", "factorType": "float16", **_sub(sub)})
    b = o.uniform_pm1(28, h.nrows).astype(np.float32)
    x = dev(np.zeros(h.nrows, np.float32))
    S.solve(x, dev(b))
    lu, d = o.ilu0_factor(L, np.float32)
    lu16 = lu.astype(np.float16).astype(np.float32)
    ref = o.ilu0_apply(L, lu16, d, b)
    assert np.array_equal(x.numpy(), ref), name


@pytest.fixture(scope="module", params=[32, 48])
def one_tile(request):
    """Poisson-7 on an m^3 grid as one tile (32^3: 32,768 rows, x fits in
    LDS in float32; 48^3: 110,592 rows, levels of more than 1,024 rows and x
    in global memory)."""
    m = request.param
    h, L = poisson_case(m, m, m)
    assert h.num_tiles == 1
    return m, h, L, h.copyToTile()


@pytest.mark.parametrize("variant", ["float32", "float64", "dilu", "float16"])
def test_ilu0_one_large_tile_bit_exact(one_tile, variant):
    """A single tile too large for the one-wave kernels runs the
    workgroup-per-tile kernel by default; factors and substitution equal
    the oracle's."""
    m, h, L, A = one_tile
    ft = "float32" if variant in ("dilu", "float16") else variant
    dt = np.float64 if ft == "float64" else np.float32
    cfg = {"type": "ILU", "factorType": "float16" if variant == "float16" else ft}
    if variant == "dilu":
        cfg["diagonalBased"] = True
    S = g.Solver(A, cfg)
    b = o.uniform_pm1(31, h.nrows).astype(dt)
    x = dev(np.zeros(h.nrows, dt))
    S.solve(x, dev(b))
    lu, d = o.ilu0_factor(L, dt, diagonalBased=variant == "dilu")
    if variant == "float16":
        lu = lu.astype(np.float16).astype(np.float32)
    ref = o.ilu0_apply(L, lu, d, b, diagonalBased=variant == "dilu")
    assert np.array_equal(x.numpy(), ref), (m, variant)


@pytest.mark.parametrize("symmetric", [False, True])
def test_gauss_seidel_bit_exact(case, symmetric):
    name, h, L, A = case
    S = g.Solver(A, {"type": "GaussSeidel", "numFixedIterations": 2,
                     "symmetric": symmetric})
    b = o.uniform_pm1(18, h.nrows)
    x0 = o.uniform_pm1(19, h.nrows) * 0.5
    x = dev(x0)
    S.solve(x, dev(b))
    ref = x0.copy()
    for _ in range(2):
        ref = o.gs_sweep(L, b, ref)
        if symmetric:
            ref = o.gs_sweep(L, b, ref, backward=True)
    assert np.array_equal(x.numpy(), ref), name


def test_gauss_seidel_split_levels():
    """One 3000-row power-law tile: its largest forward levels hold more
    entries than the sweep kernel stages at once (4864 float32 slots less
    the tile's 3000 rows), so they are streamed as several consecutive
    levels; the sweeps stay bit-exact against the oracle."""
    h, L = powerlaw_case(3000, T=1)
    rp, cl = L["row_ptr"], L["col_local"]
    lvl = np.zeros(h.nrows, np.int64)
    for i in range(h.nrows):
        lo = cl[rp[i]:rp[i + 1]]
        lo = lo[lo < i]
        lvl[i] = 1 + lvl[lo].max() if len(lo) else 0
    per_level = np.bincount(lvl, weights=np.diff(rp))
    assert per_level.max() > 4864 - h.nrows
    A = h.copyToTile()
    b = o.uniform_pm1(18, h.nrows)
    x0 = o.uniform_pm1(19, h.nrows) * 0.5
    for symmetric in (False, True):
        S = g.Solver(A, {"type": "GaussSeidel", "numFixedIterations": 2,
                         "symmetric": symmetric})
        x = dev(x0)
        S.solve(x, dev(b))
        ref = x0.copy()
        for _ in range(2):
            ref = o.gs_sweep(L, b, ref)
            if symmetric:
                ref = o.gs_sweep(L, b, ref, backward=True)
        assert np.array_equal(x.numpy(), ref), symmetric


def _pbicgstab_cfg(rel, precond="ILU", working="float32", maxit=1000):
    c = {"type": "PBiCGStab", "relResidual": rel, "maxIterations": maxit,
         "workingType": working}
    if precond == "ILU":
        c["preconditioner"] = {"type": "ILU", "factorType": working}
    elif precond == "GS":
        c["preconditioner"] = {"type": "GaussSeidel", "numFixedIterations": 1,
                               "symmetric": True}
    elif precond == "DILU":
        c["preconditioner"] = {"type": "ILU", "factorType": working,
                               "diagonalBased": True}
    return c


@pytest.mark.parametrize("precond,working", [("ILU", "float32"), ("ILU", "float64"),
                                              ("none", "float64"), ("GS", "float64"),
                                              ("none", "float32"), ("GS", "float32"),
                                              ("DILU", "float32"), ("DILU", "float64")])
def test_pbicgstab_iterations_match_oracle(case, precond, working):
    """Iteration counts within 5% of the oracle (the reduction order of the
    dot products differs, so the iterates are not bit-identical). FP32 runs
    without ILU amplify that rounding difference; they check convergence and
    the true residual only."""
    name, h, L, A = case
    dt = np.float32 if working == "float32" else np.float64
    xt, b = rhs_for(h, L)
    rel = 1e-6 if working == "float32" else 1e-9
    x = dev(np.zeros(h.nrows, dt))
    st = g.Solver(A, _pbicgstab_cfg(rel, precond, working)).solve(x, dev(b.astype(dt)))
    S = o.System(L)
    ocfg = S.pbicgstab_cfg(rel, 1000, precond={"ILU": "ilu", "none": "none",
                                                 "GS": "gs", "DILU": "dilu"}[precond],
                           working64=(working == "float64"), gsSweeps=1, gsSymmetric=True)
    xo, sto = S.pbicgstab(np.zeros(h.nrows, dt), b.astype(dt), ocfg)
    assert st.converged and sto.converged, (name, st)
    if precond != "none" and (precond in ("ILU", "DILU") or working == "float64"):
        assert abs(st.iterations - sto.iterations) <= max(1, 0.05 * sto.iterations), \
            (name, st.iterations, sto.iterations)
    elif working == "float64":
        # plain BiCGStab on a symmetric Poisson matrix has an irregular
        # residual history; the rounding of the dot products moves the
        # stopping iteration by more than 5%
        assert abs(st.iterations - sto.iterations) <= max(2, 0.2 * sto.iterations), \
            (name, st.iterations, sto.iterations)
    r = o.residual_mixed(L, x.numpy().astype(np.float64), b)[1]
    assert np.linalg.norm(r) <= 10 * rel * np.linalg.norm(b)


def test_pbicgstab_f64_and_mpir(case):
    name, h, L, A = case
    xt, b = rhs_for(h, L, seed=3)
    # plain FP64
    x = dev(np.zeros(h.nrows, np.float64))
    st = g.Solver(A, _pbicgstab_cfg(1e-10, "ILU", "float64")).solve(x, dev(b.astype(np.float64)))
    assert st.converged, st
    r = o.residual_mixed(L, x.numpy(), b)[1]
    assert np.linalg.norm(r) <= 2e-10 * np.linalg.norm(b)
    # MPIR: FP32 inner, FP64 outer
    cfg = {"type": "IterativeRefinement", "relResidual": 1e-10, "maxIterations": 30,
           "extendedPrecisionType": "float64", "workingType": "float32",
           "innerSolver": _pbicgstab_cfg(1e-3, "ILU", "float32", 100)}
    phi = dev(np.zeros(h.nrows, np.float64))
    st = g.Solver(A, cfg).solve(phi, dev(b))
    assert st.converged, st
    r = o.residual_mixed(L, phi.numpy(), b)[1]
    assert np.linalg.norm(r) <= 1e-10 * np.linalg.norm(b) * 1.0001
    S = o.System(L)
    _, sto, inner = S.mpir(b, 1e-10, 30, inner=S.pbicgstab_cfg(1e-3, 100))
    # the inner solves stop on rounding-sensitive thresholds: outer counts may
    # differ by one, total inner iterations by 10%
    assert abs(st.iterations - sto.iterations) <= 1, (st.iterations, sto.iterations)
    if st.iterations == sto.iterations:
        assert abs(st.innerIterations - inner) <= max(3, 0.1 * inner), \
            (st.innerIterations, inner)


@pytest.mark.parametrize("ftype", ["float32", "float16"])
def test_pbicgstab_level_substitution_identical(case, ftype):
    """PBiCGStab with the lane-per-row ILU(0) kernels (their fused p- and
    s-updates included) gives the same iterates, bit for bit, as with the
    product-slab kernel."""
    name, h, L, A = case
    xt, b = rhs_for(h, L, seed=5)
    outs = []
    for sub in ("packed", "level", "rows"):
        cfg = _pbicgstab_cfg(1e-6)
        cfg["preconditioner"] = {"type": "ILU", "factorType": ftype, **_sub(sub)}
        x = dev(np.zeros(h.nrows, np.float32))
        st = g.Solver(A, cfg).solve(x, dev(b))
        outs.append((st.iterations, x.numpy()))
    for it, xv in outs[1:]:
        assert it == outs[0][0]
        assert np.array_equal(xv, outs[0][1])


def test_solver_deterministic(case):
    name, h, L, A = case
    xt, b = rhs_for(h, L, seed=4)
    S = g.Solver(A, _pbicgstab_cfg(1e-6))
    outs = []
    for _ in range(2):
        x = dev(np.zeros(h.nrows, np.float32))
        S.solve(x, dev(b))
        outs.append(x.numpy())
    assert np.array_equal(outs[0], outs[1])
