"""GPU parity of the twofloat32 (double-word) extended precision: the
double-word residual kernel bit-exact against the oracle (CRS in both kernel
forms, LDU), the casts, and refinement with "extendedPrecisionType":
"twofloat32" against the oracle's double-word refinement."""
import numpy as np
import pytest

import graphene_amd as g
import oracle as o
from helpers import convdiff_case, poisson_case, powerlaw_case, rhs_for

pytestmark = pytest.mark.gpu

CASES = {
    "poisson16_T8": lambda: poisson_case(16, 16, 16, block=(8, 8, 8)),
    "convdiff12_T27": lambda: convdiff_case(12, block=(4, 4, 4)),
    "powerlaw3000_T6": lambda: powerlaw_case(3000, T=6),
}


def dw_vector(seed, n):
    """Double words of x ~ U[-1,1] / 3 (hi = fl32, lo = fl32 of the rest)."""
    d = o.uniform_pm1(seed, n).astype(np.float64) / 3.0
    a = np.zeros(n, g.TWOFLOAT)
    a["hi"] = d.astype(np.float32)
    a["lo"] = (d - a["hi"].astype(np.float64)).astype(np.float32)
    return a


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("fmt", ["crs-sliced", "crs-rows", "ldu"])
def test_residual_dw_bit_exact(name, fmt):
    h, L = CASES[name]()
    if fmt == "ldu" and name.startswith("powerlaw"):
        pytest.skip("LDU needs a structurally symmetric pattern")
    A = h.copyToTile(g.LDU if fmt == "ldu" else g.CRS)
    A.set_sliced(fmt != "crs-rows")
    n = h.nrows
    x = dw_vector(31, n)
    _, b = rhs_for(h, L, seed=5)
    r64 = g.DeviceVector(n, np.float64)
    r = A.residual(g.DeviceVector.from_numpy(x), g.DeviceVector.from_numpy(b),
                   destType=g.FLOAT32, intermediateType=g.TWOFLOAT32, r64=r64)
    xp = np.stack([x["hi"], x["lo"]], 1)
    if fmt == "ldu":
        ref32, ref64 = o.ldu_residual_dw(o.ldu_build(L), xp, b)
    else:
        ref32, ref64 = o.residual_dw(L, xp, b)
    assert np.array_equal(r.numpy().view(np.uint32), ref32.view(np.uint32)), (name, fmt)
    assert np.array_equal(r64.numpy().view(np.uint64), ref64.view(np.uint64)), (name, fmt)


def test_casts():
    n = 5000
    f = o.uniform_pm1(2, n) * np.float32(1e3)
    d = o.uniform_pm1(3, n).astype(np.float64) / 7.0
    t1 = g.DeviceVector.from_numpy(f).cast(g.TWOFLOAT).numpy()
    assert np.array_equal(t1["hi"], f) and not t1["lo"].any()
    assert np.array_equal(g.DeviceVector.from_numpy(t1).cast(np.float32).numpy(), f)
    t2 = g.DeviceVector.from_numpy(d).cast(g.TWOFLOAT).numpy()
    hi = d.astype(np.float32)
    assert np.array_equal(t2["hi"], hi)
    assert np.array_equal(t2["lo"], (d - hi.astype(np.float64)).astype(np.float32))
    back = g.DeviceVector.from_numpy(t2).cast(np.float64).numpy()
    assert np.array_equal(back, t2["hi"].astype(np.float64) + t2["lo"])


@pytest.mark.parametrize("name", ["poisson16_T8", "convdiff12_T27"])
def test_mpir_twofloat_vs_oracle(name):
    """Refinement in twofloat32: converged, outer count of the oracle's
    double-word refinement (+-1), float64 residual of the solution within the
    tolerance."""
    h, L = CASES[name]()
    A = h.copyToTile()
    n = h.nrows
    _, b = rhs_for(h, L, seed=7)
    cfg = {"type": "IterativeRefinement", "relResidual": 1e-10, "maxIterations": 30,
           "extendedPrecisionType": "twofloat32",
           "innerSolver": {"type": "PBiCGStab", "relResidual": 1e-3, "maxIterations": 100,
                           "preconditioner": {"type": "ILU"}}}
    x = g.DeviceVector(n, np.float32)
    x.fill(0.0)
    st = A.solve(x, g.DeviceVector.from_numpy(b), cfg)
    _, sto, _ = o.System(L).mpir_dw(b, 1e-10, 30)
    assert st.converged and sto.converged
    assert abs(st.iterations - sto.iterations) <= 1, (st.iterations, sto.iterations)
    # the solution handed back is fl32(phi): its float64 residual is at the
    # float32 level; the refinement's own residual is the one below 1e-10
    assert st.finalResidual < 1e-10 * st.initialResidual


def test_twofloat_only_as_extended_precision():
    h, L = poisson_case(8, 8, 8)
    A = h.copyToTile()
    x = g.DeviceVector(h.nrows, np.float32)
    b = g.DeviceVector.from_numpy(np.ones(h.nrows, np.float32))
    with pytest.raises(g.GrapheneError):
        A.solve(x, b, {"type": "PBiCGStab", "workingType": "twofloat32"})
