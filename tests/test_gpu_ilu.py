"""Block ILU(0) on subdomains (GPU_PRECONDITIONER = ILU0), this build's
alternative to the reference's block-Jacobi preconditioner.

Pinned by what ILU(0) is: on every subdomain (GPU_ILU_SUBDOMAIN consecutive
cells of the device order), the factors reproduce the matrix on its pattern,
(L U)_cj = A_cj with L unit lower and U upper, U_kk the inverse of the stored
pivot inverse. The preconditioner changes the Krylov path only: at a tight
tolerance the corrections equal block Jacobi's, and a steady solve reaches
the same steady state.
"""
import numpy as np
import pytest

from myjoe_amd import cases, joemi

pytestmark = pytest.mark.gpu

# tight enough to compare corrections, loose enough that BiCGStab does not
# stagnate (at REL 1e-10 both preconditioners hit a 300-iteration cap)
TIGHT = "MAX_ITER=300 ABS_RESID=1.0e-14 REL_RESID=1.0e-6"


def ilu_text(joe, sub=16):
    return joe + f"GPU_PRECONDITIONER = ILU0\nGPU_ILU_SUBDOMAIN = {sub}\n"


def bump(nsteps=5, ls=TIGHT, sa=False):
    mm, joe = cases.bump_case(24, 8, nsteps=nsteps, ls=ls)
    if sa:
        visc = "MU_REF = 0.002\nMU_MODE = POWERLAW\nMU_POWER_LAW = 0.76\nPR = 0.72"
        joe = cases.coupled(cases.with_sa(joe.replace("MU_REF = 0.0", visc), 0.006, cases.SA_BUMP))
    return joemi.Mesh.from_faces(mm), joe


def check_factors(mesh, s, b, sub):
    """(L U) = A on each subdomain's pattern, in the reference's numbering."""
    A = s.get("A").reshape(-1, b, b)
    F = s.get("ILU_F").reshape(-1, b, b)
    Dinv = s.get("ILU_DINV").reshape(-1, b, b)
    pos = s.get("CELL_ORDER").astype(np.int64)
    ni = mesh.array("NBOCV_I").astype(np.int64)
    nv = mesh.array("NBOCV_V").astype(np.int64)
    n = mesh.ncv
    dom = pos // sub
    entry = {}
    for c in range(n):
        for e in range(ni[c], ni[c + 1]):
            entry[(c, int(nv[e]))] = e
    U_diag = np.linalg.inv(Dinv)

    def Lblk(c, k):  # unit lower factor
        if k == c:
            return np.eye(b)
        return F[entry[(c, k)]] if pos[k] < pos[c] else np.zeros((b, b))

    def Ublk(k, j):  # upper factor
        if j == k:
            return U_diag[k]
        return F[entry[(k, j)]] if pos[j] > pos[k] else np.zeros((b, b))

    worst, checked = 0.0, 0
    scale = np.max(np.abs(A))
    for (c, j), e in entry.items():
        if dom[c] != dom[j]:
            continue
        acc = np.zeros((b, b))
        for e2 in range(ni[c], ni[c + 1]):  # k runs over row c's pattern with k <= c and k <= j
            k = int(nv[e2])
            if dom[k] != dom[c] or pos[k] > min(pos[c], pos[j]):
                continue
            if k != j and (k, j) not in entry:
                continue
            acc += Lblk(c, k) @ Ublk(k, j)
        worst = max(worst, float(np.max(np.abs(acc - A[e]))) / scale)
        checked += 1
    return worst, checked


@pytest.mark.parametrize("sa", [False, True])
def test_ilu_factors_reproduce_the_matrix_on_each_subdomain(sa):
    print(f"ilu factors sa={sa}", flush=True)
    mesh, joe = bump(sa=sa)
    s = joemi.Solver(mesh, ilu_text(joe, 16))
    s.init_uniform()
    for _ in range(3):
        s.step()
    worst, checked = check_factors(mesh, s, 6 if sa else 5, 16)
    print("ilu factors", sa, checked, f"{worst:.2e}", flush=True)
    assert checked > mesh.ncv and worst <= 1e-12


@pytest.mark.parametrize("sa", [False, True])
def test_ilu_correction_equals_block_jacobi_at_a_tight_tolerance(sa):
    print(f"ilu solve sa={sa}", flush=True)
    mesh, joe = bump(sa=sa)
    out = {}
    for name, text in (("bj", joe), ("ilu", ilu_text(joe, 32))):
        s = joemi.Solver(mesh, text)
        s.init_uniform()
        for _ in range(3):
            info = s.step()
        out[name] = (s.get("DQ"), info.lin_iters, info.lin_status)
    (dq_bj, it_bj, st_bj), (dq_ilu, it_ilu, st_ilu) = out["bj"], out["ilu"]
    print("iterations bj", it_bj, "ilu", it_ilu, flush=True)
    assert st_bj == 1 and st_ilu == 1
    assert it_ilu < it_bj
    assert np.max(np.abs(dq_ilu - dq_bj)) <= 1e-4 * np.max(np.abs(dq_bj))


@pytest.mark.parametrize("precision", [joemi.JM_FP64, joemi.JM_MIXED])
def test_ilu_steady_state_equals_block_jacobi(precision):
    print(f"ilu steady precision={precision}", flush=True)
    mesh, joe = bump(nsteps=1200, ls="MAX_ITER=30 ABS_RESID=1.0e-8 REL_RESID=1.0e-2")
    states = []
    for text in (joe, ilu_text(joe, 64)):
        s = joemi.Solver(mesh, text, precision)
        s.init_uniform()
        n, hist = s.run(1200)
        assert hist[-1, 4] < 1e-6 * hist[0, 4]
        states.append(s.state())
    for a, b, k in zip(states[0], states[1], ("rho", "rhou", "rhoE")):
        d = float(np.max(np.abs(a - b)) / np.max(np.abs(a)))
        print("ilu steady", precision, k, f"{d:.1e}", flush=True)
        assert d <= 1e-8, (k, d)


def test_batched_server_rejects_ilu():
    mesh, joe = bump()
    srv = joemi.Server(1 << 30, 2, batched=True)
    with pytest.raises(joemi.JoemiError) as e:
        srv.submit(mesh, ilu_text(joe))
    assert e.value.status == 5 and "block-Jacobi" in str(e.value)
    srv.close()
