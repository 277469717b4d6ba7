"""Stopping rules of the solvers on the GPU against the oracle.

SolverStats::checkConvergence / checkSingularity (SolverStats.cpp:33-63) as
PBiCGStab uses them (pbicgstab/Solver.cpp:66-90, :95, :159, :170-180):
  * zero right-hand side: rho = 0 is singular before the first iteration;
  * NaN in b: nothing converges and nothing is singular (|NaN| < 1e-37 is
    false), so the loop ends at maxIterations;
  * relTolerance compares with ||x|| of the initial guess (the reference's
    bNorm quirk, :69-70), or with ||b|| when relToleranceFromB is set;
  * absTolerance; minIterations keeps iterating after convergence.
The restarter (restarter/Solver.cpp:40-56) re-runs its inner solver from the
current x until it converges or maxRestarts is reached; the oracle side is
the same loop around the oracle's PBiCGStab.
"""
import numpy as np
import pytest

import graphene_amd as g
import oracle as o
from helpers import convdiff_case, rhs_for

pytestmark = pytest.mark.gpu


def dev(a):
    return g.DeviceVector.from_numpy(np.ascontiguousarray(a))


@pytest.fixture(scope="module")
def sysm():
    h, L = convdiff_case(12, block=(4, 4, 4))
    return h, L, h.copyToTile(), o.System(L)


def cfg(working="float32", **kw):
    c = {"type": "PBiCGStab", "maxIterations": 200, "workingType": working,
         "preconditioner": {"type": "ILU", "factorType": working}}
    c.update(kw)
    return c


def ocfg(S, working="float32", **kw):
    return S.pbicgstab_cfg(kw.pop("relResidual", 0.0), kw.pop("maxIterations", 200),
                           precond="ilu", working64=(working == "float64"), **kw)


def both(sysm, gcfg, oc, b, x0, working="float32"):
    h, L, A, S = sysm
    dt = np.float32 if working == "float32" else np.float64
    x = dev(np.asarray(x0, dt))
    st = g.Solver(A, gcfg).solve(x, dev(np.asarray(b, dt)))
    xo, sto = S.pbicgstab(np.asarray(x0, dt), np.asarray(b, dt), oc)
    return st, sto, x.numpy(), xo


def close(a, b):
    return abs(a - b) <= max(1, 0.05 * b)


@pytest.mark.parametrize("working", ["float32", "float64"])
def test_zero_rhs_is_singular(sysm, working):
    n = sysm[0].nrows
    st, sto, x, _ = both(sysm, cfg(working, relResidual=1e-6),
                         ocfg(sysm[3], working, relResidual=1e-6),
                         np.zeros(n), np.zeros(n), working)
    assert sto.singular and sto.iterations == 0 and not sto.converged
    assert st.singular and st.iterations == 0 and not st.converged
    assert not x.any()


@pytest.mark.parametrize("working", ["float32", "float64"])
def test_nan_rhs_runs_to_max_iterations(sysm, working):
    h, L, _, _ = sysm
    _, b = rhs_for(h, L, seed=31)
    b = b.astype(np.float64)
    b[17] = np.nan
    st, sto, _, _ = both(sysm, cfg(working, relResidual=1e-6, maxIterations=7),
                         ocfg(sysm[3], working, relResidual=1e-6, maxIterations=7),
                         b, np.zeros(h.nrows), working)
    assert (sto.iterations, sto.converged, sto.singular) == (7, 0, 0)
    assert (st.iterations, st.converged, st.singular) == (7, False, False)
    assert np.isnan(st.finalResidual)


def test_rel_tolerance_uses_initial_guess_norm(sysm):
    """relTolerance alone: with x0 = 0 the threshold ||x0|| * relTol is 0 and
    the solve runs to maxIterations; with a nonzero x0 it stops once
    ||r|| < ||x0|| relTol, as the oracle does; relToleranceFromB uses ||b||."""
    h, L, _, S = sysm
    _, b = rhs_for(h, L, seed=32)
    n = h.nrows
    st, sto, _, _ = both(sysm, cfg(relTolerance=1e-4, maxIterations=25),
                         ocfg(S, relTolerance=1e-4, maxIterations=25), b, np.zeros(n))
    assert sto.iterations == 25 and not sto.converged
    assert st.iterations == 25 and not st.converged
    x0 = 0.25 * o.uniform_pm1(33, n)
    st, sto, x, xo = both(sysm, cfg(relTolerance=1e-4), ocfg(S, relTolerance=1e-4), b, x0)
    assert sto.converged and st.converged and close(st.iterations, sto.iterations)
    assert st.finalResidual < 1e-4 * np.linalg.norm(x0.astype(np.float64)) * (1 + 1e-6)
    st, sto, _, _ = both(sysm, cfg(relTolerance=1e-5, relToleranceFromB=True),
                         ocfg(S, relTolerance=1e-5, bNormFromX=False), b, x0)
    assert sto.converged and st.converged and close(st.iterations, sto.iterations)
    assert st.finalResidual < 1e-5 * np.linalg.norm(b.astype(np.float64)) * (1 + 1e-6)


@pytest.mark.parametrize("working", ["float32", "float64"])
def test_abs_tolerance(sysm, working):
    h, L, _, S = sysm
    _, b = rhs_for(h, L, seed=34)
    tol = float(1e-5 * np.linalg.norm(b.astype(np.float64)))
    st, sto, _, _ = both(sysm, cfg(working, absTolerance=tol),
                         ocfg(S, working, absTolerance=tol), b, np.zeros(h.nrows), working)
    assert sto.converged and st.converged and close(st.iterations, sto.iterations)
    assert st.finalResidual < tol


def test_min_iterations(sysm):
    h, L, _, S = sysm
    _, b = rhs_for(h, L, seed=35)
    st, sto, _, _ = both(sysm, cfg(relResidual=1e-1), ocfg(S, relResidual=1e-1), b,
                         np.zeros(h.nrows))
    few = st.iterations
    assert st.converged and few < 12
    st, sto, _, _ = both(sysm, cfg(relResidual=1e-1, minIterations=12),
                         ocfg(S, relResidual=1e-1, minIterations=12), b, np.zeros(h.nrows))
    assert sto.iterations == 12 and st.iterations == 12 and st.converged


def test_restarter_against_oracle_loop(sysm):
    """maxRestarts re-runs of a 4-iteration PBiCGStab: the GPU's restarts and
    final stats follow the oracle's restart loop."""
    h, L, A, S = sysm
    _, b = rhs_for(h, L, seed=36)
    inner = cfg("float64", relResidual=1e-9, maxIterations=4)
    rcfg = {"type": "restarter", "maxRestarts": 3, "innerSolver": inner}
    x = dev(np.zeros(h.nrows))
    R = g.Solver(A, rcfg)
    assert R.name().startswith("Restarter")
    st = R.solve(x, dev(b.astype(np.float64)))
    # oracle: solve, then while restarts < max and not converged: solve again
    oc = ocfg(S, "float64", relResidual=1e-9, maxIterations=4)
    xo, sto = S.pbicgstab(np.zeros(h.nrows), b.astype(np.float64), oc)
    total, restarts = sto.iterations, 0
    while restarts < 3 and not sto.converged:
        restarts += 1
        xo, sto = S.pbicgstab(xo, b.astype(np.float64), oc)
        total += sto.iterations
    assert restarts == 3 and total == 16  # 4 solves of 4 iterations
    assert st.iterations == sto.iterations and st.converged == bool(sto.converged)
    assert st.innerIterations == total
    r = o.residual_mixed(L, x.numpy(), b)[1]
    ro = o.residual_mixed(L, xo, b)[1]
    assert abs(np.linalg.norm(r) - np.linalg.norm(ro)) <= 1e-3 * np.linalg.norm(ro)
    # a restarter whose inner solve converges at once does not restart
    x2 = dev(np.zeros(h.nrows))
    st2 = g.Solver(A, {"type": "restarter", "maxRestarts": 3,
                       "innerSolver": cfg("float64", relResidual=1e-6)}).solve(
        x2, dev(b.astype(np.float64)))
    assert st2.converged and st2.innerIterations == st2.iterations


# --- norms (lg/matrix/Norm.hpp): L1 = sum of magnitudes, LINF = largest
# magnitude (the reference's L1 sums signed values and its LINF reduction is
# unavailable: documented deviations) ------------------------------------------
@pytest.mark.parametrize("norm", ["L1", "LINF"])
def test_vector_norm_kinds(sysm, norm):
    h, L, A, _ = sysm
    for dt in (np.float32, np.float64):
        x = (o.uniform_pm1(40, h.nrows) * 3).astype(dt)
        got = A.vectorNorm(dev(x), norm)
        ref = o.vec_norm(x, L["tile_start"], norm)
        if norm == "LINF":
            assert got == float(ref)
        else:
            assert abs(got - ref) <= (1e-5 if dt == np.float32 else 1e-12) * ref


@pytest.mark.parametrize("working", ["float32", "float64"])
@pytest.mark.parametrize("norm", ["L1", "LINF"])
def test_pbicgstab_norms_against_oracle(sysm, norm, working):
    h, L, _, S = sysm
    _, b = rhs_for(h, L, seed=41)
    st, sto, x, xo = both(sysm, cfg(working, relResidual=1e-6, norm=norm),
                          ocfg(S, working, relResidual=1e-6, norm=norm), b,
                          np.zeros(h.nrows), working)
    assert sto.converged and st.converged and close(st.iterations, sto.iterations)
    assert abs(st.initialResidual - sto.initialResidual) <= 1e-5 * sto.initialResidual
    assert st.finalResidual < 1e-6 * st.initialResidual
    # the norm of the returned residual, recomputed on the host
    r = o.residual_mixed(L, x.astype(np.float64), b)[1]
    rn = np.sum(np.abs(r)) if norm == "L1" else np.max(np.abs(r))
    assert rn <= 10e-6 * sto.initialResidual


def test_mpir_and_gauss_seidel_norms(sysm):
    h, L, A, S = sysm
    _, b = rhs_for(h, L, seed=42)
    inner = cfg("float32", relResidual=1e-3, maxIterations=100, norm="L1")
    mcfg = {"type": "IterativeRefinement", "relResidual": 1e-10, "maxIterations": 30,
            "extendedPrecisionType": "float64", "workingType": "float32", "norm": "LINF",
            "innerSolver": inner}
    phi = dev(np.zeros(h.nrows))
    st = g.Solver(A, mcfg).solve(phi, dev(b))
    _, sto, inn = S.mpir(b, 1e-10, 30, inner=ocfg(S, "float32", relResidual=1e-3,
                                                    maxIterations=100, norm="L1"),
                         norm="LINF")
    assert st.converged and sto.converged and abs(st.iterations - sto.iterations) <= 1
    r = o.residual_mixed(L, phi.numpy(), b)[1]
    assert np.max(np.abs(r)) <= 1e-10 * st.initialResidual * 1.0001
    assert abs(st.initialResidual - np.max(np.abs(o.residual_mixed(
        L, np.zeros(h.nrows), b)[1]))) <= 1e-12 * st.initialResidual
    # Gauss-Seidel iterated to a relative L1 residual
    gcfg = {"type": "GaussSeidel", "relResidual": 1e-4, "maxIterations": 500,
            "numSweeps": 1, "norm": "L1"}
    x = dev(np.zeros(h.nrows, np.float32))
    sg = g.Solver(A, gcfg).solve(x, dev(b))
    assert sg.converged and sg.finalResidual < 1e-4 * sg.initialResidual
    rr = o.residual_mixed(L, x.numpy().astype(np.float64), b)[1]
    assert abs(np.sum(np.abs(rr)) - sg.finalResidual) <= 1e-3 * sg.finalResidual
