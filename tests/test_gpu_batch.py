"""Batched solves (gph_batch_*): many 2D 9-point convection-diffusion
systems in one block-diagonal matrix, solved together. Checked against the
true FP64 residual of every system and against each system solved alone
with the single-system solver (same tiles, so the same ILU(0) operator; the
dot products are summed in a different order, so iterates agree to the
convergence tolerance rather than bit for bit)."""
import numpy as np
import pytest
import scipy.sparse as sp

import graphene_amd as g

pytestmark = pytest.mark.gpu

SIZES = [(12, 9), (40, 31), (64, 64), (23, 77), (8, 8), (90, 50)]
BLOCK = (16, 16)


def make_systems(seed=7):
    rng = np.random.default_rng(seed)
    out = []
    for k, (nx, ny) in enumerate(SIZES):
        D, Pe = rng.uniform(0.5, 2.0), rng.uniform(0.0, 20.0)
        r, c, v = g.generateConvectionDiffusion9Triplets(nx, ny, D, Pe, 100 + k)
        out.append((nx, ny, r, c, v))
    return out


def mpir_cfg(inner_rel=1e-3, sub=None):
    pre = {"type": "ILU"}
    if sub:
        pre["substitution"] = sub
    return {"type": "IterativeRefinement", "relResidual": 1e-10, "maxIterations": 20,
            "extendedPrecisionType": "float64", "workingType": "float32",
            "innerSolver": {"type": "PBiCGStab", "relResidual": inner_rel,
                            "maxIterations": 200, "preconditioner": pre}}


@pytest.fixture(scope="module")
def batch():
    systems = make_systems()
    h, ts = g.batch_union(systems, BLOCK)
    A = h.copyToTile(g.CRS)
    n = h.nrows
    # b = fl32(A x_true), A x_true in FP64 (natural order per system)
    xt = g.uniform_pm1(5, n).astype(np.float64)
    b64 = A.spmv(g.DeviceVector.from_numpy(h.to_local(xt)))
    b = b64.numpy().astype(np.float32)
    off = np.concatenate([[0], np.cumsum([nx * ny for nx, ny, *_ in systems])])
    return dict(systems=systems, h=h, ts=ts, A=A, b=b, off=off)


def true_rel_residuals(B, x_local):
    """Per system ||b - A x|| / ||b|| in FP64 (natural order)."""
    h = B["h"]
    xg = h.to_global(np.asarray(x_local, np.float64))
    bg = h.to_global(B["b"].astype(np.float64))
    out = []
    for (nx, ny, r, c, v), o in zip(B["systems"], B["off"]):
        n = nx * ny
        M = sp.csr_matrix((np.asarray(v, np.float64), (r, c)), shape=(n, n))
        rs = bg[o:o + n] - M @ xg[o:o + n]
        out.append(np.linalg.norm(rs) / np.linalg.norm(bg[o:o + n]))
    return np.array(out)


@pytest.mark.parametrize("sub", [None, "rows", "packed"])
def test_batch_mpir_converges_and_matches_alone(batch, sub):
    """With the substitution kernel picked by timing, and with each kernel
    (finished systems' tiles or wave groups skipped)."""
    B = batch
    bt = g.Batch(B["A"], B["ts"], mpir_cfg(sub=sub))
    assert bt.num_systems == len(SIZES) and bt.refinement
    assert bt.rows_per_wave in ((0,) if sub == "rows" else (512, 1024, 2048, 4096)
                                if sub == "packed" else (0, 512, 1024, 2048, 4096))
    x = g.DeviceVector(B["h"].nrows, np.float64)
    stats, sec = bt.solve(x, g.DeviceVector.from_numpy(B["b"]))
    assert all(s.converged for s in stats), stats
    assert all(s.finalResidual < 1e-10 * s.initialResidual for s in stats)
    rel = true_rel_residuals(B, x.numpy())
    assert (rel < 2e-10).all(), rel
    # each system alone, with the same tiling and configuration
    xb = B["h"].to_global(x.numpy())
    bg = B["h"].to_global(B["b"].astype(np.float64)).astype(np.float32)
    for k, ((nx, ny, r, c, v), o) in enumerate(zip(B["systems"], B["off"])):
        n = nx * ny
        r2t, T = g.partition_grid3d(nx, ny, 1, *BLOCK, 1)
        hk = g.HostMatrix.from_coo(n, r, c, v, T, r2t)
        Ak = hk.copyToTile(g.CRS)
        bk = g.DeviceVector.from_numpy(hk.to_local(bg[o:o + n]))
        xk = g.DeviceVector.from_numpy(np.zeros(n))  # the batch starts from 0 too
        st = g.Solver(Ak, mpir_cfg()).solve(xk, bk)
        assert st.converged
        assert abs(st.iterations - stats[k].iterations) <= 1, (k, st, stats[k])
        if st.iterations == stats[k].iterations:  # else one more inner solve
            assert abs(st.innerIterations - stats[k].innerIterations) <= \
                max(2, 0.1 * st.innerIterations), (k, st, stats[k])
        xa = hk.to_global(xk.numpy())
        err = np.linalg.norm(xa - xb[o:o + n]) / np.linalg.norm(xa)
        assert err < 1e-7, (k, err)


def test_batch_repeat_identical(batch):
    """Two solves with one batch (the captured graph replayed) give the same
    bits."""
    B = batch
    bt = g.Batch(B["A"], B["ts"], mpir_cfg())
    outs = []
    for _ in range(2):
        x = g.DeviceVector(B["h"].nrows, np.float64)
        stats, _ = bt.solve(x, g.DeviceVector.from_numpy(B["b"]))
        outs.append((x.numpy(), [s.innerIterations for s in stats]))
    assert np.array_equal(outs[0][0], outs[1][0]) and outs[0][1] == outs[1][1]


@pytest.mark.parametrize("sub", ["rows", "packed"])
def test_batch_outer_step_device_equals_host(batch, sub):
    """The refinement's outer step on the device (each system restarts its
    inner solve as soon as its outer step is done) and driven by the host
    (every system's inner solve ends first) give every system the same
    operations, hence the same bits and iteration counts."""
    B = batch
    outs = []
    for outer in ("device", "host"):
        bt = g.Batch(B["A"], B["ts"], dict(mpir_cfg(sub=sub), outerStep=outer))
        x = g.DeviceVector(B["h"].nrows, np.float64)
        stats, _ = bt.solve(x, g.DeviceVector.from_numpy(B["b"]))
        outs.append((x.numpy(), [(s.iterations, s.innerIterations, s.finalResidual)
                                 for s in stats]))
    assert outs[0][1] == outs[1][1]
    assert np.array_equal(outs[0][0].view(np.uint64), outs[1][0].view(np.uint64))


def test_batch_rejects_bad_outer_step(batch):
    B = batch
    with pytest.raises(g.GrapheneError, match="outerStep"):
        g.Batch(B["A"], B["ts"], dict(mpir_cfg(), outerStep="gpu"))


def test_batch_pbicgstab_only(batch):
    B = batch
    cfg = {"type": "PBiCGStab", "relResidual": 1e-6, "maxIterations": 500,
           "preconditioner": {"type": "ILU", "substitution": "packed", "rowsPerWave": 1024}}
    bt = g.Batch(B["A"], B["ts"], cfg)
    assert not bt.refinement and bt.rows_per_wave == 1024
    x = g.DeviceVector(B["h"].nrows, np.float32)
    stats, _ = bt.solve(x, g.DeviceVector.from_numpy(B["b"]))
    assert all(s.converged for s in stats), stats
    rel = true_rel_residuals(B, x.numpy())
    assert (rel < 1e-5).all(), rel


def test_batch_rejects_coupled_systems(batch):
    B = batch
    ts = B["ts"]
    # splitting the third system's tiles into two "systems" couples them
    bad = np.concatenate([ts[:3], [ts[2] + 1], ts[3:]])
    with pytest.raises(g.GrapheneError, match="couples two systems"):
        g.Batch(B["A"], bad, mpir_cfg())
    with pytest.raises(g.GrapheneError, match="cover"):
        g.Batch(B["A"], ts[:-1], mpir_cfg())
