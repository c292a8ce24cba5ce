"""The CPU restatement (oracle/) against the reference's own outputs.

Every fixture under tests/golden/ was written by the reference's code
(oracle/_ref/joe_ref_dump, built from the reference sources by
tests/golden/make_golden.py). Both are serial, FP64 and compiled without
floating-point contraction, so the restatement is held to bit-exact equality
on every field: primitive and boundary states, local time steps, limited
gradients, residuals, Jacobian blocks, one BiCGStab correction and the
state trajectory of JoeWithModels::run.
"""
import glob
import os

import numpy as np
import pytest

from myjoe_amd import cases
from oracle.oracle import Oracle, parse_joe_in

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
RHS_FIXTURES = sorted(os.path.basename(p)[4:-4] for p in glob.glob(os.path.join(GOLD, "rhs_*.npz")))
FIELDS = ["PRESS", "VEL", "RHO_BFA", "VEL_BFA", "P_BFA", "T_BFA", "H_BFA", "LOCAL_DT", "GRAD_RHO", "GRAD_P",
          "GRAD_U", "ALPHA_RHO", "RHS_RHO", "RHS_RHOU", "RHS_RHOE", "A"]
VISC_FIELDS = ["MUL_FA", "GRAD_ENTHALPY"]
# Spalart-Allmaras: model fields, the nuSA boundary values, gradient,
# diffusivity, eddy viscosity, residual and matrix rows AScal[6][nnz]
SA_FIELDS = ["WALL_DIST", "STR_MAG", "VORT_MAG", "MUT_FA", "NUSA_BFA", "GRAD_NUSA", "DIFF_NUSA", "RHS_NUSA", "ASCAL"]
# Menter SST: the model fields (divergence, cross diffusion, blending functions
# F1 and F2), kine and omega boundary values (omega's wall hook included),
# gradients, diffusivities, eddy viscosity, residuals and AScal[2][6][nnz]
SST_FIELDS = ["WALL_DIST", "STR_MAG", "VORT_MAG", "DIVERG", "CROSS_DIFF", "F1", "F2", "MUT_FA", "KINE_BFA",
              "OMEGA_BFA", "GRAD_KINE", "GRAD_OMEGA", "DIFF_KINE", "DIFF_OMEGA", "RHS_KINE", "RHS_OMEGA", "ASCAL"]


def load(name):
    with np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def mesh_for(fixture):
    return load("mesh_" + fixture.split("_")[0])


def joe_text(d):
    """The fixture's Joe.in; with a PROFILES zone, its profile table written
    to a file (the format the reference reads) that the text then names."""
    text = bytes(d["JOE_IN"]).decode()
    if "PROFILE" in d:
        import hashlib
        import tempfile

        tab = np.asarray(d["PROFILE"], np.float64)
        path = os.path.join(tempfile.gettempdir(), f"joemi_prof_{hashlib.sha1(tab.tobytes()).hexdigest()[:16]}.dat")
        if not os.path.exists(path):
            tmp = path + f".{os.getpid()}"
            with open(tmp, "w") as f:
                f.write(f"n={tab.shape[0]}\td={tab.shape[1]}\n")
                for r in tab:
                    f.write(" ".join(repr(float(v)) for v in r) + "\n")
            os.replace(tmp, path)
        text = text.replace("PROFILES prof.dat", "PROFILES " + path)
    return text


def test_fixture_set_complete():
    # three schemes, both gradient schemes, first order, every BC kind and SYMM_BC mode
    assert len(RHS_FIXTURES) >= 11
    kinds = set()
    symm = set()
    for f in RHS_FIXTURES:
        P = parse_joe_in(joe_text(load("rhs_" + f)))
        for z in ("inlet", "outlet", "wall", "farfield", "symm", "bump_wall", "top", "zsym"):
            if z in P:
                kinds.add(P[z][0][1])
        symm.add(P["SYMM_BC"][0][1])
    assert kinds >= {"CBC", "CBC_SUBSONIC_INLET", "CBC_SUBSONIC_OUTLET", "WALL", "FARFIELD", "NEUMANN", "SYMMETRY",
                     "PROFILES"}
    assert symm == {"ORIGINAL_NUMERIC", "NEW_NUMERIC", "ANALYTICAL"}
    # viscous terms: both viscosity laws, with and without the gradient limiter, first order
    visc = [f for f in RHS_FIXTURES if "visc" in f]
    assert any("_pl" in f for f in visc) and any("_su" in f for f in visc) and any("_nl" in f for f in visc)
    assert any("first" in f for f in visc)
    # Spalart-Allmaras on both geometries, LS and GG, first order, every BC kind
    sa = [f for f in RHS_FIXTURES if "_sa" in f]
    assert len(sa) >= 4 and any("gg" in f for f in sa) and any("first" in f for f in sa)
    assert any(f.startswith("bump") for f in sa) and any("alt" in f for f in sa)
    sst = [f for f in RHS_FIXTURES if "_sst" in f]
    assert len(sst) >= 4 and any("gg" in f for f in sst) and any("first" in f for f in sst)
    assert any(f.startswith("bump") for f in sst) and any("alt" in f for f in sst) and any("dtsum" in f for f in sst)


@pytest.mark.parametrize("fixture", RHS_FIXTURES)
def test_residual_bit_exact(fixture):
    g = load("rhs_" + fixture)
    o = Oracle(mesh_for(fixture), joe_text(g))
    o.set_state(g["RHO"], g["RHOU"], g["RHOE"])
    sa, sst = "_sa" in fixture, "_sst" in fixture
    if sa:
        o.set_scalar(g["NUSA"])
    if sst:
        o.set_scalar(np.concatenate([g["KINE"], g["OMEGA"]]))
    o.residual(True)
    first = "first" in fixture
    for k in (FIELDS + (VISC_FIELDS if "visc" in fixture else []) + (SA_FIELDS if sa else [])
              + (SST_FIELDS if sst else [])):
        if k not in g:
            assert first and k.startswith(("GRAD", "ALPHA")), k
            continue
        got = o.get(k)
        assert got.shape == g[k].shape, k
        nbad = int(np.sum(got != g[k]))
        assert nbad == 0, f"{fixture}:{k}: {nbad} of {got.size} differ; max |d| {np.max(np.abs(got - g[k])):.3e}"
    o.linear_solve()
    assert np.array_equal(o.get("DQ"), g["DQ"]), fixture
    if sa or sst:  # the scalar half of the step: relaxation, coupling columns, Jacobi BiCGStab
        assert np.array_equal(o.get("DSCAL"), g["DSCAL"]), fixture


def test_sa_fields_physical():
    # what the SA closure fixes independently of the reference: wall distance
    # is positive and bounded by the nearest wall-face centroid distance, the
    # eddy viscosity is zero on walls and non-negative elsewhere, strain and
    # vorticity magnitudes are non-negative
    g = load("rhs_bump24_hllc_visc_pl_sa")
    m = mesh_for("bump24_hllc_visc_pl_sa")
    o = Oracle(m, joe_text(g))
    o.set_state(g["RHO"], g["RHOU"], g["RHOE"])
    o.set_scalar(g["NUSA"])
    o.residual(True)
    from oracle import jmdump
    walls = [(f, l) for name, k, f, l in jmdump.zones(m) if k == 1 and name == "bump_wall"]
    wf = np.concatenate([np.arange(f, l + 1) for f, l in walls])
    x_cv = m["X_CV"].reshape(-1, 3)[: o.ncv]
    x_fa = m["X_FA"].reshape(-1, 3)
    dmin = np.min(np.linalg.norm(x_cv[:, None, :] - x_fa[wf][None, :, :], axis=2), axis=1)
    wd = o.get("WALL_DIST")
    assert np.all(wd > 0.0) and np.all(wd <= dmin * (1 + 1e-12))
    mut = o.get("MUT_FA")
    assert np.all(mut[wf] == 0.0) and np.all(mut >= 0.0)
    assert np.all(o.get("STR_MAG") >= 0.0) and np.all(o.get("VORT_MAG") >= 0.0)


@pytest.mark.parametrize("name", ["bump24", "bump96", "bump24visc", "bump24sa", "bump24sst"])
def test_run_trajectory_bit_exact(name):
    g = load("solve_" + name)
    joe = joe_text(g)
    P = parse_joe_in(joe)
    m = load("mesh_" + name.replace("visc", "").replace("sst", "").replace("sa", ""))
    o = Oracle(m, joe)
    ncv = int(m["COUNTS"][3])
    # JoeWithModels::initialHook uniform state (JoeWithModels.cpp:180-205)
    u = [float(x) for x in P["U_INITIAL"][0][1:4]]
    rho0, p0, gam = 1.0, 1.0, 1.4
    uu = 0.0
    for x in u:
        uu += x * x
    o.set_state(np.full(ncv, rho0), np.tile(np.array(u) * rho0, (ncv, 1)),
                np.full(ncv, p0 / (gam - 1.0) + 0.5 * rho0 * uu))
    for k in range(1, 6):
        o.step()
        rho, rhou, rhoE = o.state()
        assert np.array_equal(rho, g[f"RHO@{k}"]), (name, k)
        assert np.array_equal(rhou.reshape(-1), g[f"RHOU@{k}"].reshape(-1)), (name, k)
        assert np.array_equal(rhoE, g[f"RHOE@{k}"]), (name, k)
        if name.endswith("sa"):
            assert np.array_equal(o.get("NUSA"), g[f"NUSA@{k}"]), (name, k)
        if name.endswith("sst"):
            for s in ("KINE", "OMEGA"):
                assert np.array_equal(o.get(s), g[f"{s}@{k}"]), (name, k, s)
    nsteps = int(g["STEP"][0])
    n, _ = o.run(nsteps - 5)
    assert n == nsteps - 5
    rho, rhou, rhoE = o.state()
    assert np.array_equal(rho, g["RHO"]) and np.array_equal(rhoE, g["RHOE"]), name
    assert np.array_equal(rhou.reshape(-1), g["RHOU"].reshape(-1)), name
    if name.endswith("sa"):
        assert np.array_equal(o.get("NUSA"), g["NUSA"]), name
    if name.endswith("sst"):
        assert np.array_equal(o.get("KINE"), g["KINE"]) and np.array_equal(o.get("OMEGA"), g["OMEGA"]), name


def test_uniform_state_matches_case_definition():
    # the initial state used above is the one the case text asks for
    _, joe = cases.bump_case(24, 8)
    P = parse_joe_in(joe)
    assert float(P["U_INITIAL"][0][1]) == cases.U_BUMP
    assert "RHO_INITIAL" not in P and "P_INITIAL" not in P


def _inviscid(fixture):
    P = parse_joe_in(joe_text(load("rhs_" + fixture)))
    return float(P["MU_REF"][0][1]) == 0.0 and "TURB_MODEL" not in P and "GRAVITY" not in P


@pytest.mark.parametrize("fixture", [f for f in RHS_FIXTURES if _inviscid(f)])
def test_oracle_face_fluxes_sum_to_reference_residual(fixture):
    # FLUX_FA, the oracle's per-face inviscid fluxes, summed per cell in the
    # reference's face order (internal faces ascending: icv0 -F, icv1 +F; then
    # the boundary zones in list order) give the reference's own residual bit
    # for bit: the per-face values are the ones the reference accumulated
    from oracle import jmdump

    g = load("rhs_" + fixture)
    mesh = mesh_for(fixture)
    o = Oracle(mesh, joe_text(g))
    o.set_state(g["RHO"], g["RHOU"], g["RHOE"])
    o.residual(False)
    F = o.get("FLUX_FA").reshape(-1, 5)
    counts = mesh["COUNTS"]
    nfa, nfa_b, ncv = int(counts[1]), int(counts[2]), int(counts[3])
    cvofa = mesh["CVOFA"].reshape(-1, 2)
    R = np.zeros((ncv, 5))
    for f in range(nfa_b, nfa):
        R[cvofa[f, 0]] -= F[f]
        R[cvofa[f, 1]] += F[f]
    for name, kind, f0, f1 in jmdump.zones(mesh):
        if kind != 1:
            continue
        for f in range(f0, f1 + 1):
            R[cvofa[f, 0]] -= F[f]
    # the reference then adds the gravity source with g = 0: vol rho g and
    # vol (g . rhou), whose zeros carry the signs of the products
    vol = mesh["CV_VOLUME"][:ncv]
    ru = g["RHOU"].reshape(-1, 3)
    R[:, 1:4] += (vol * g["RHO"])[:, None] * 0.0
    R[:, 4] += vol * (0.0 * ru[:, 0] + 0.0 * ru[:, 1] + 0.0 * ru[:, 2])
    assert np.array_equal(R[:, 0], g["RHS_RHO"])
    assert np.array_equal(R[:, 1:4].reshape(-1), g["RHS_RHOU"])
    assert np.array_equal(R[:, 4], g["RHS_RHOE"])
