"""jm_block_bcgstab: the reference's solveCvVectorR5Bcgstab on a block system
the caller holds on the device, against the reference's own first linear
solve of runBackwardEuler (tests/golden/rhs_*.npz: A, RHS, LOCAL_DT and the
correction DQ that the reference's solver returned).

Bar: 1e-9 relative, as for the context's own solve (the block-reduced dot
products sum in another order than the reference's serial loops)."""
import numpy as np
import pytest

from myjoe_amd import joemi
from tests.helpers import compare
from tests.test_oracle_golden import load

pytestmark = pytest.mark.gpu

FIXTURES = ["box4_hllc", "box4_roe", "bump24_hllc", "box4_hllc_visc_su_sa"]


def reference_system(fixture):
    """The block system runBackwardEuler hands its linear solver
    (JoeWithModels.cpp:1011-1026): rhs = UNDER_RELAXATION (0.3) x RHS and
    A with V / dt added to each diagonal block, reference numbering."""
    g = load("rhs_" + fixture)
    m = load("mesh_" + fixture.split("_")[0])
    ncv = int(m["COUNTS"][3])
    ni = m["NBOCV_I"].astype(np.int32)
    nv = m["NBOCV_V"].astype(np.int32)
    A = g["A"].reshape(-1, 5, 5).copy()
    vol = m["CV_VOLUME"][:ncv]
    for c in range(ncv):
        A[ni[c]] += np.eye(5) * (vol[c] / g["LOCAL_DT"][c])
    rhs = 0.3 * np.concatenate([g["RHS_RHO"][:, None], g["RHS_RHOU"].reshape(-1, 3), g["RHS_RHOE"][:, None]], 1)
    return ni, nv, A, rhs, g["DQ"].reshape(-1, 5)


@pytest.mark.parametrize("fixture", FIXTURES)
def test_block_bcgstab_matches_reference_solve(fixture):
    print(f"block bcgstab {fixture}", flush=True)
    ni, nv, A, rhs, dq = reference_system(fixture)
    t = joemi.DeviceArray
    phi = t(np.zeros(rhs.shape))
    it, st = joemi.block_bcgstab(t(ni), t(nv), t(A), t(rhs), phi)
    assert it >= 1 and st in (1, 2)
    compare(phi.numpy(), dq, 1e-9, f"{fixture}:DQ")


def test_block_bcgstab_initial_guess_is_used():
    # phi comes in as the initial guess: started from the converged answer of
    # a tight solve, the first stop test after one iteration is met and phi
    # stays the solution of the dense system
    ni, nv, A, rhs, _ = reference_system("box4_hllc")
    t = joemi.DeviceArray
    dni, dnv, dA, drhs = t(ni), t(nv), t(A), t(rhs)
    phi = t(np.zeros(rhs.shape))
    joemi.block_bcgstab(dni, dnv, dA, drhs, phi, zero_abs=1e-15, zero_rel=1e-15, max_iter=200)
    it, st = joemi.block_bcgstab(dni, dnv, dA, drhs, phi, zero_abs=1e-6, zero_rel=1e-30, max_iter=30)
    assert st == 1 and it <= 1
    exact = np.linalg.solve(dense(ni, nv, A), rhs.reshape(-1))
    assert np.max(np.abs(phi.numpy().reshape(-1) - exact)) <= 1e-9 * np.max(np.abs(exact))


def dense(ni, nv, A):
    ncv = len(ni) - 1
    M = np.zeros((5 * ncv, 5 * ncv))
    for c in range(ncv):
        for noc in range(ni[c], ni[c + 1]):
            n = nv[noc]
            M[5 * c:5 * c + 5, 5 * n:5 * n + 5] += A[noc]
    return M
