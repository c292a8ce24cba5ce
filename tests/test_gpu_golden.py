"""HIP path against the reference's own outputs (tests/golden/).

Same bars as against the oracle: residual-path fields to 1e-12 relative
(bit-exact except where the reference's libm pow enters: FARFIELD and
subsonic-inlet boundary states, the limiter's eps^2), the BiCGStab
correction to 1e-9 (block-reduced dot products), and the first steps of
JoeWithModels::run to 1e-9.
"""
import numpy as np
import pytest

from myjoe_amd import joemi, meshgen
from tests.helpers import compare
from tests.test_oracle_golden import FIELDS, RHS_FIXTURES, joe_text, load

# the Spalart-Allmaras fields the device exposes; the eddy viscosity is
# formed per face inside the flux kernel, and MUT_FA reads back what it used
SA_GPU_FIELDS = ["NUSA_BFA", "GRAD_NUSA", "WALL_DIST", "STR_MAG", "VORT_MAG", "RHS_NUSA", "ASCAL", "MUT_FA"]
# the SST fields the device exposes (the diffusivities are formed per face
# inside the scalar kernels, as the eddy viscosity inside the flux kernel)
SST_GPU_FIELDS = ["KINE_BFA", "OMEGA_BFA", "GRAD_KINE", "GRAD_OMEGA", "WALL_DIST", "STR_MAG", "VORT_MAG", "DIVERG",
                  "CROSS_DIFF", "F1", "F2", "RHS_KINE", "RHS_OMEGA", "ASCAL", "MUT_FA"]
from tests.test_mesh_golden import MESHES

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fixture", RHS_FIXTURES)
def test_residual_matches_reference(fixture):
    print(f"golden {fixture}", flush=True)
    g = load("rhs_" + fixture)
    mesh = joemi.Mesh.from_faces(MESHES[fixture.split("_")[0]]())
    s = joemi.Solver(mesh, joe_text(g))
    s.set_state(g["RHO"], g["RHOU"], g["RHOE"])
    sa, sst = "_sa" in fixture, "_sst" in fixture
    if sa:
        s.set_scalar(g["NUSA"])
    if sst:
        s.set_scalar(g["KINE"], "kine")
        s.set_scalar(g["OMEGA"], "omega")
    s.calc_rhs(True)
    exact = total = 0
    for k in (FIELDS + (["GRAD_ENTHALPY", "MUL_FA"] if "visc" in fixture else []) + (SA_GPU_FIELDS if sa else [])
              + (SST_GPU_FIELDS if sst else [])):
        if k not in g:
            continue
        e, n = compare(s.get(k), g[k], 1e-12, f"{fixture}:{k}")
        exact += e
        total += n
    print(f"{fixture}: {exact}/{total} bit-exact")
    info = s.step()
    assert info.lin_iters >= 1
    compare(s.get("DQ"), g["DQ"], 1e-9, f"{fixture}:DQ")
    if sa:  # the scalar half of the step (Jacobi BiCGStab, block-reduced dots)
        assert info.scal_iters >= 1
        compare(s.get("DSCAL"), g["DSCAL"], 1e-9, f"{fixture}:DSCAL")
    if sst:  # kine's and omega's solves
        assert info.scal_iters >= 2
        compare(s.get("DSCAL"), g["DSCAL"], 1e-9, f"{fixture}:DSCAL")


@pytest.mark.parametrize("name", ["bump24", "bump96", "bump24visc", "bump24sa", "bump24sst"])
def test_run_matches_reference(name):
    print(f"golden solve {name}", flush=True)
    g = load("solve_" + name)
    nx, ny = (96, 32) if name == "bump96" else (24, 8)
    mesh = joemi.Mesh.from_faces(meshgen.bump_channel(nx, ny))
    s = joemi.Solver(mesh, joe_text(g))
    s.init_uniform()
    for k in range(1, 6):
        s.step()
        rho, rhou, rhoE = s.state()
        compare(rho, g[f"RHO@{k}"], 1e-9, f"{name} rho@{k}")
        compare(rhou.reshape(-1), g[f"RHOU@{k}"].reshape(-1), 1e-9, f"{name} rhou@{k}")
        compare(rhoE, g[f"RHOE@{k}"], 1e-9, f"{name} rhoE@{k}")
        if name.endswith("sa"):
            compare(s.get("NUSA"), g[f"NUSA@{k}"], 1e-9, f"{name} nuSA@{k}")
        if name.endswith("sst"):
            compare(s.get("KINE"), g[f"KINE@{k}"], 1e-9, f"{name} kine@{k}")
            compare(s.get("OMEGA"), g[f"OMEGA@{k}"], 1e-9, f"{name} omega@{k}")


def test_device_numbering_is_invisible():
    # the context renumbers cells and faces for locality; every result leaves
    # the ABI in the reference's numbering
    g = load("rhs_box4_hllc")
    mesh = joemi.Mesh.from_faces(MESHES["box4"]())
    s = joemi.Solver(mesh, joe_text(g))
    order = s.get("CELL_ORDER").astype(np.int64)
    assert np.array_equal(np.sort(order), np.arange(mesh.ncv))
    s.set_state(g["RHO"], g["RHOU"], g["RHOE"])
    rho, rhou, rhoE = s.state()
    assert np.array_equal(rho, g["RHO"]) and np.array_equal(rhoE, g["RHOE"])
    assert np.array_equal(rhou.reshape(-1), g["RHOU"].reshape(-1))
    # the device state buffer (SoA, device order) holds the same values
    import ctypes

    ptr = s.device_state_ptr()
    assert ptr != 0
    s.sync()
    host = np.empty(5 * mesh.ncv)
    libhip = ctypes.CDLL("libamdhip64.so")
    libhip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    assert libhip.hipMemcpy(host.ctypes.data, ptr, host.nbytes, 2) == 0  # hipMemcpyDeviceToHost
    assert np.array_equal(host[order], g["RHO"])
    assert np.array_equal(host[4 * mesh.ncv + order], g["RHOE"])


@pytest.mark.parametrize("order", ["NATIVE", "GROUPED", "MORTON", "MORTON_GROUPED"])
def test_cell_order_does_not_change_results(order):
    print(f"cell order {order}", flush=True)
    g = load("rhs_box4_roe_alt_ana")
    mesh = joemi.Mesh.from_faces(MESHES["box4"]())
    s = joemi.Solver(mesh, joe_text(g) + f"GPU_CELL_ORDER = {order}\n")
    s.set_state(g["RHO"], g["RHOU"], g["RHOE"])
    s.calc_rhs(True)
    for k in FIELDS:
        compare(s.get(k), g[k], 1e-12, f"{order}:{k}")
    s.step()
    compare(s.get("DQ"), g["DQ"], 1e-9, f"{order}:DQ")


@pytest.mark.parametrize("fixture", ["bump24_hllc", "box4_hllc_gg", "box4_roe_first", "box4_ausm_alt_ana_gg",
                                     "box4_hllc_dtsum"])
def test_cell_kernels_match_reference(fixture):
    # the other form of each residual kernel gives the same bits
    print(f"cell kernels {fixture}", flush=True)
    g = load("rhs_" + fixture)
    mesh = joemi.Mesh.from_faces(MESHES[fixture.split("_")[0]]())
    s = joemi.Solver(mesh, joe_text(g) + "GPU_GRAD_KERNEL = ENTRY\nGPU_FLUX_KERNEL = CELL\n")
    s.set_state(g["RHO"], g["RHOU"], g["RHOE"])
    s.calc_rhs(True)
    for k in FIELDS:
        if k in g:
            compare(s.get(k), g[k], 1e-12, f"cell:{fixture}:{k}")


@pytest.mark.parametrize("fixture", ["bump24_hllc", "box4_hllc_gg", "box4_roe_first", "box4_ausm_alt_ana_gg",
                                     "box4_roe_alt_ana", "box4_hllc_alt_new", "box4_hllc_dtsum"])
def test_face_flux_form_matches_reference(fixture):
    # the inviscid flux evaluated once per face and gathered per cell gives
    # the same bits as the per-entry form; the first-order fallback count too
    print(f"face flux {fixture}", flush=True)
    g = load("rhs_" + fixture)
    mesh = joemi.Mesh.from_faces(MESHES[fixture.split("_")[0]]())
    s = joemi.Solver(mesh, joe_text(g) + "GPU_FLUX_KERNEL = FACE\n")
    s.set_state(g["RHO"], g["RHOU"], g["RHOE"])
    s.calc_rhs(True)
    for k in FIELDS:
        if k in g:
            compare(s.get(k), g[k], 1e-12, f"face:{fixture}:{k}")
    e = joemi.Solver(mesh, joe_text(g))
    e.set_state(g["RHO"], g["RHOU"], g["RHOE"])
    e.calc_rhs(True)
    for k in ("RHS_RHO", "RHS_RHOU", "RHS_RHOE", "LOCAL_DT"):
        assert np.array_equal(s.get(k), e.get(k)), (fixture, k)
    assert np.array_equal(s.get("COUNTERS")[:2], e.get("COUNTERS")[:2]), fixture
    s.step()
    compare(s.get("DQ"), g["DQ"], 1e-9, f"face:{fixture}:DQ")


@pytest.mark.parametrize("name", ["bump24", "bump96", "bump24visc"])
def test_graph_step_matches_reference(name):
    # the device-driven step (CFL ramp and step counter on the device, the
    # step replayed from captured hipGraphs) against the reference's own first
    # five steps of JoeWithModels::run, not against the host-polled step
    print(f"golden graph solve {name}", flush=True)
    g = load("solve_" + name)
    nx, ny = (96, 32) if name == "bump96" else (24, 8)
    mesh = joemi.Mesh.from_faces(meshgen.bump_channel(nx, ny))
    s = joemi.Solver(mesh, joe_text(g))
    s.init_uniform()
    for k in range(1, 6):
        info = s.step_graph()
        assert info.step == k and info.lin_iters >= 1
        rho, rhou, rhoE = s.state()
        compare(rho, g[f"RHO@{k}"], 1e-9, f"{name} graph rho@{k}")
        compare(rhou.reshape(-1), g[f"RHOU@{k}"].reshape(-1), 1e-9, f"{name} graph rhou@{k}")
        compare(rhoE, g[f"RHOE@{k}"], 1e-9, f"{name} graph rhoE@{k}")


@pytest.mark.parametrize("fixture", ["box4_hllc", "box4_roe_alt_ana", "bump24_hllc_visc_pl_sa"])
def test_boundary_state_read_matches_reference(fixture):
    # jm_get_boundary_state gives each zone's setBC values as the reference
    # hands them to boundaryHook (T_bfa, vel_bfa, p_bfa)
    print(f"golden boundary state {fixture}", flush=True)
    g = load("rhs_" + fixture)
    mesh = joemi.Mesh.from_faces(MESHES[fixture.split("_")[0]]())
    s = joemi.Solver(mesh, joe_text(g))
    s.set_state(g["RHO"], g["RHOU"], g["RHOE"])
    if "_sa" in fixture:
        s.set_scalar(g["NUSA"])
    s.calc_rhs(False)
    vb = g["VEL_BFA"].reshape(-1, 3)
    nb = 0
    for iz, (zname, kind, f0, f1) in enumerate(mesh.zones()):
        if kind != 1:
            continue
        T, vel, p = s.boundary_state(iz, f1 - f0 + 1)
        compare(T, g["T_BFA"][f0:f1 + 1], 1e-12, f"{fixture}:{zname}:T")
        compare(vel, vb[f0:f1 + 1], 1e-12, f"{fixture}:{zname}:vel")
        compare(p, g["P_BFA"][f0:f1 + 1], 1e-12, f"{fixture}:{zname}:p")
        nb += f1 - f0 + 1
    assert nb == mesh.nfa_b


@pytest.mark.parametrize("fixture", ["box4_hllc", "box4_hllc_visc_su_sa", "bump24_hllc_visc_pl_sa"])
def test_residual_on_reference_arrays(fixture):
    # jm_mesh_from_ugp, the entry a host holding the reference's own Ugp
    # arrays binds: the mesh is built from the reference's dumped arrays
    # (tests/golden/mesh_*.npz) without renumbering, and the residual matches
    # the reference's on it
    print(f"golden ugp mesh {fixture}", flush=True)
    g = load("rhs_" + fixture)
    mesh = joemi.Mesh.from_ugp(load("mesh_" + fixture.split("_")[0]))
    s = joemi.Solver(mesh, joe_text(g))
    s.set_state(g["RHO"], g["RHOU"], g["RHOE"])
    sa = "_sa" in fixture
    if sa:
        s.set_scalar(g["NUSA"])
    s.calc_rhs(True)
    for k in FIELDS + (["GRAD_ENTHALPY", "MUL_FA"] if "visc" in fixture else []) + (SA_GPU_FIELDS if sa else []):
        if k in g:
            compare(s.get(k), g[k], 1e-12, f"ugp {fixture}:{k}")
