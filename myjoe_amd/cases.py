"""Workload definitions: meshes, Joe.in text and synthetic states.

* ``bump_case``: steady 2D inviscid Euler channel with a 10 %-thick
  circular-arc bump, gamma 1.4, inlet Mach 0.5, p = rho = 1 (R = 1), HLLC,
  2nd order with least-squares gradients and the smooth Barth-Jespersen
  limiter, backward Euler with a CFL ramp 1 -> 50 (x1.2 every 5 steps from
  step 20; a 10 %-per-step ramp makes the reference itself diverge).
* ``box_case``: the split/jittered box of the single-residual throughput test
  with inlet/outlet CBC, WALL, FARFIELD and SYMMETRY zones.
* ``airfoil_case``: inviscid flow past a NACA 4-digit section on an O-mesh
  (the serving workload's requests), free stream p = rho = 1.
* ``random_state``: rho ~ U(0.5, 1.5), p ~ lognormal(0, 0.2),
  u_i ~ N(0, 0.3) * c, as conservative variables.
* ``plate_case``: zero-pressure-gradient flat plate, Mach 0.2, Reynolds
  number per unit length Re, Spalart-Allmaras in the coupled loop (6x6
  blocks), free-stream nuSA = 3 nu.
* ``with_sa``: adds the Spalart-Allmaras scalar nuSA to a viscous case.
* ``with_sst``: adds the SST k-omega scalars kine, omega to a viscous case.
* ``coupled``: the same case under runBackwardEulerCoupled.
"""
from __future__ import annotations

import math

import numpy as np

from . import meshgen

GAMMA = 1.4
MACH_BUMP = 0.5
U_BUMP = MACH_BUMP * math.sqrt(GAMMA)  # c = sqrt(gamma p / rho) with p = rho = 1


def _common(scheme: str, second_order: bool, grad: str, eps: float, extra: str,
            symm_bc: str = "ORIGINAL_NUMERIC") -> str:
    lines = [
        "GAMMA = 1.4", "P_REF = 1.0", "RHO_REF = 1.0", "T_REF = 1.0", "MU_REF = 0.0",
        f"SPACE_INTEGRATION = {scheme}",
        f"GRAD_RECONSTRUCTION SCHEME={grad} LIMITER=YES EPS={eps!r}",
        "TIME_INTEGRATION = BACKWARD_EULER", "LINEAR_SOLVER_NS = BCGSTAB",
        f"SYMM_BC = {symm_bc}", "INITIAL_CONDITION_TURB = 0.0", "CHECK_INTERVAL = 1",
    ]
    if second_order:
        lines.append("SPATIAL_SECOND_ORDER")
    return "\n".join(lines) + "\n" + extra


def bump_case(nx: int = 96, ny: int = 32, scheme: str = "HLLC", second_order: bool = True, nsteps: int = 2000,
              cfl_max: float = 50.0, restart: str = "bump.msh", resid_th: float = 0.0,
              ls: str = "MAX_ITER=30 ABS_RESID=1.0e-8 REL_RESID=1.0e-2", thickness: float = 0.1,
              mach: float = MACH_BUMP, resid_rel: float = 0.0):
    mesh = meshgen.bump_channel(nx, ny, thickness)
    u = repr(U_BUMP if mach == MACH_BUMP else mach * math.sqrt(GAMMA))
    extra = "\n".join([
        f"RESTART = {restart}",
        f"LINEAR_SOLVER_NS_TRESHOLDS {ls}",
        "CFL = 1.0",
        f"CFL_RAMP AFTER_ITER=20 INTERVAL_ITER=5 FACTOR_CFL=1.2 MAX_CFL={cfl_max!r}",
        f"NSTEPS = {nsteps}",
        f"RESID_ENERG_TH = {resid_th!r}",
    ] + ([f"RESID_ENERG_REL = {resid_rel!r}"] if resid_rel > 0 else []) + [
        f"U_INITIAL = {u} 0.0 0.0",
        f"inlet = CBC {u} 0.0 0.0 1.0 1.0",
        f"outlet = CBC {u} 0.0 0.0 1.0 1.0",
        "bump_wall = WALL",
        "top = SYMMETRY",
        "zsym = SYMMETRY",
    ]) + "\n"
    return mesh, _common(scheme, second_order, "LEASTSQUARES", 0.01, extra)


BOX_BCS = {
    "inlet": "CBC 0.5 0.1 0.0 1.0 1.0",
    "outlet": "CBC 0.5 0.1 0.0 1.0 1.0",
    "wall": "WALL",
    "farfield": "FARFIELD 0.4 0.2 0.1 1.0 1.0 1.4",
    "symm": "SYMMETRY",
}
# the other boundary kinds of setBC: total-condition inlet (flow angle, T_tot,
# p_tot), static-pressure outlet, zero-gradient
BOX_BCS_ALT = {
    "inlet": "CBC_SUBSONIC_INLET 0.8 0.6 0.0 1.3 1.4",
    "outlet": "CBC_SUBSONIC_OUTLET 0.9",
    "wall": "WALL",
    "farfield": "NEUMANN",
    "symm": "SYMMETRY",
}


def box_joe_in(scheme: str = "HLLC", second_order: bool = True, grad: str = "LEASTSQUARES",
               restart: str = "box.msh", bcs: dict | None = None, symm_bc: str = "ORIGINAL_NUMERIC") -> str:
    bcs = BOX_BCS if bcs is None else bcs
    extra = "\n".join([
        f"RESTART = {restart}",
        "CFL = 1.0",
        "NSTEPS = 10",
        "U_INITIAL = 0.3 0.1 0.0",
    ] + [f"{z} = {bcs[z]}" for z in ("inlet", "outlet", "wall", "farfield", "symm")]) + "\n"
    return _common(scheme, second_order, grad, 0.01, extra, symm_bc)


def box_case(n: int, frac: float = 0.5, jitter: float = 0.2, seed: int = 0, scheme: str = "HLLC",
             second_order: bool = True, grad: str = "LEASTSQUARES", restart: str = "box.msh",
             bcs: dict | None = None, symm_bc: str = "ORIGINAL_NUMERIC"):
    mesh = meshgen.split_box(n, frac, jitter, seed)
    return mesh, box_joe_in(scheme, second_order, grad, restart, bcs, symm_bc)


def coupled(joe: str, keep_symm_bc: bool = False) -> str:
    """The same case run by runBackwardEulerCoupled: TIME_INTEGRATION =
    BACKWARD_EULER_COUPLED with the reference's coupled linear solver keyword
    (PETSC_GMRES; solved here by the block-Jacobi BiCGStab). Without
    keep_symm_bc the SYMM_BC line is dropped, so the default applies (the
    reference's init registers ORIGINAL_NUMERIC before calcFluxCoupled reads
    the key with its own default)."""
    joe = joe.replace("TIME_INTEGRATION = BACKWARD_EULER\n", "TIME_INTEGRATION = BACKWARD_EULER_COUPLED\n")
    joe = joe.replace("LINEAR_SOLVER_NS = BCGSTAB\n", "LINEAR_SOLVER_NS = PETSC_GMRES\n")
    if not keep_symm_bc:
        joe = "\n".join(ln for ln in joe.split("\n") if not ln.startswith("SYMM_BC"))
    return joe


def random_state(ncv: int, seed: int = 1):
    """(rho, rhou[ncv,3], rhoE) drawn as in the single-residual throughput test."""
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.5, 1.5, ncv)
    p = np.exp(rng.normal(0.0, 0.2, ncv))
    c = np.sqrt(GAMMA * p / rho)
    u = rng.normal(0.0, 0.3, (ncv, 3)) * c[:, None]
    rhou = rho[:, None] * u
    rhoE = p / (GAMMA - 1.0) + 0.5 * rho * np.sum(u * u, axis=1)
    return rho, rhou, rhoE


def airfoil_case(ni: int = 200, nj: int = 100, thickness: float = 0.12, camber: float = 0.0, mach: float = 0.5,
                 aoa_deg: float = 2.0, scheme: str = "HLLC", nsteps: int = 2000, resid_rel: float = 1e-6,
                 cfl_max: float = 50.0, restart: str = "airfoil.msh"):
    """Steady inviscid flow past a NACA 4-digit section: FARFIELD at Mach
    mach and angle of attack aoa_deg, WALL on the section, SYMMETRY in z.
    Stops at nsteps or when the energy L1 residual has dropped by resid_rel
    (RESID_ENERG_REL)."""
    mesh = meshgen.naca_omesh(ni, nj, thickness, camber)
    a = math.radians(aoa_deg)
    c = math.sqrt(GAMMA)
    u, v = mach * c * math.cos(a), mach * c * math.sin(a)
    extra = "\n".join([
        f"RESTART = {restart}",
        "LINEAR_SOLVER_NS_TRESHOLDS MAX_ITER=30 ABS_RESID=1.0e-8 REL_RESID=1.0e-2",
        "CFL = 1.0",
        f"CFL_RAMP AFTER_ITER=20 INTERVAL_ITER=5 FACTOR_CFL=1.2 MAX_CFL={cfl_max!r}",
        f"NSTEPS = {nsteps}",
        f"RESID_ENERG_REL = {resid_rel!r}",
        f"U_INITIAL = {u!r} {v!r} 0.0",
        "airfoil = WALL",
        f"farfield = FARFIELD {u!r} {v!r} 0.0 1.0 1.0 1.4",
        "zsym = SYMMETRY",
    ]) + "\n"
    return mesh, _common(scheme, True, "LEASTSQUARES", 0.01, extra)


def serving_requests(n: int, seed: int = 0, min_cells: int = 3_000, max_cells: int = 12_000,
                     rate_per_s: float = 1.0):
    """n steady bump-channel requests (the bump case family): size U(min,
    max) cells as nx = 3 ny, bump thickness U(6, 12) %, inlet Mach U(0.3, 0.5);
    Poisson arrival times at rate_per_s (all at 0 when rate_per_s is 0).

    Inviscid NACA O-mesh requests were tried first: with this scheme (the
    CPU restatement, bit-identical to the reference, shows the same) they
    diverge at the trailing edge once the CFL ramp passes ~25, open or closed
    trailing edge, FARFIELD or CBC far field; the bump family converges."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate_per_s, n)) if rate_per_s > 0 else np.zeros(n)
    out = []
    for k in range(n):
        cells = int(rng.uniform(min_cells, max_cells))
        ny = max(8, int(round(math.sqrt(cells / 3.0))))
        out.append(dict(arrival=float(t[k]), nx=3 * ny, ny=ny, thickness=float(rng.uniform(0.06, 0.12)),
                        mach=float(rng.uniform(0.3, 0.5))))
    return out


def with_sa(joe_in: str, nu_init: float, zone_values: dict) -> str:
    """The case with the Spalart-Allmaras model: ``TURB_MODEL = SA`` (the
    reference picks its RansTurbSA class by run number instead), the initial
    nuSA (INITIAL_CONDITION_TURB) and ``<zone>.nuSA`` Dirichlet values; zones
    without a value get zero-gradient nuSA. The case must be viscous."""
    joe_in = joe_in.replace("INITIAL_CONDITION_TURB = 0.0", f"INITIAL_CONDITION_TURB = {nu_init!r}")
    return joe_in + "TURB_MODEL = SA\n" + "".join(f"{z}.nuSA = {v!r}\n" for z, v in zone_values.items())


# nuSA boundary values of the SA variants of the two test geometries
SA_BUMP = {"inlet": 0.006, "bump_wall": 0.0}
SA_BOX = {"inlet": 0.006, "wall": 0.0, "farfield": 0.004}


def with_sst(joe_in: str, kine_init: float, omega_init: float, zone_values: dict) -> str:
    """The case with the Menter SST k-omega model: ``TURB_MODEL = SST`` (the
    reference picks its RansTurbKOmSST class by run number instead), the
    initial kine and omega (INITIAL_CONDITION_TURB) and per zone a pair
    (kine, omega) of Dirichlet values, omega ``"HOOK"`` on walls (the model's
    wall value 60 mu / (rho beta_1 d^2)); zones without values get
    zero-gradient scalars. The case must be viscous."""
    joe_in = joe_in.replace("INITIAL_CONDITION_TURB = 0.0", f"INITIAL_CONDITION_TURB = {kine_init!r} {omega_init!r}")
    lines = ["TURB_MODEL = SST"]
    for z, (k, w) in zone_values.items():
        lines.append(f"{z}.kine = {k!r}")
        lines.append(f"{z}.omega = {w if isinstance(w, str) else repr(w)}")
    return joe_in + "\n".join(lines) + "\n"


# kine, omega boundary values of the SST variants of the two test geometries
SST_BUMP = {"inlet": (0.001, 5.0), "bump_wall": (0.0, "HOOK")}
SST_BOX = {"inlet": (0.001, 5.0), "wall": (0.0, "HOOK"), "farfield": (0.0008, 4.0)}


MACH_PLATE = 0.2


def plate_case(nx_up: int = 40, nx: int = 324, ny: int = 110, nz: int = 50, re: float = 5.0e6,
               ratio: float = 1.1, nsteps: int = 20000, cfl_max: float = 1.0e3, resid_th: float = 0.0,
               ls: str = "MAX_ITER=30 ABS_RESID=1.0e-8 REL_RESID=1.0e-2",
               ratio_x: float = 1.05, cfl_ramp: str = "AFTER_ITER=10 INTERVAL_ITER=5 FACTOR_CFL=1.2",
               restart: str = "plate.msh", turb: str = "SA"):
    """Implicit RANS on a zero-pressure-gradient flat plate: Mach 0.2,
    p = rho = 1 (R = 1, so T = 1), Re per unit length from a constant
    viscosity MU_REF = U / Re (power law with T = T_REF at the free stream),
    Spalart-Allmaras in the coupled loop (runBackwardEulerCoupled, 6x6
    blocks), free-stream and inlet nuSA = 3 nu, nuSA = 0 on the plate.
    Characteristic inlet and top, static-pressure outlet, symmetry ahead of
    the plate and in z. The defaults give the 2M-cell plate (40 + 324 cells
    along x, 110 across the boundary layer with the first cell at y+ < 1, 50
    in z). With turb = "SST" the Menter SST model in the segregated loop
    (runBackwardEuler): free-stream kine from 1 % turbulence intensity,
    omega from an eddy viscosity equal to the laminar one, kine = 0 and the
    model's omega hook on the plate."""
    mesh = meshgen.flat_plate(nx_up, nx, ny, nz, ratio=ratio, ratio_x=ratio_x)
    u = MACH_PLATE * math.sqrt(GAMMA)
    mu = u / re
    nu_sa = 3.0 * mu
    lines = [
        "GAMMA = 1.4", "P_REF = 1.0", "RHO_REF = 1.0", "T_REF = 1.0",
        f"MU_REF = {mu!r}", "MU_MODE = POWERLAW", "MU_POWER_LAW = 0.76", "PR = 0.72", "PR_TURB = 0.9",
        "SPACE_INTEGRATION = HLLC", "GRAD_RECONSTRUCTION SCHEME=LEASTSQUARES LIMITER=YES EPS=0.01",
        "SPATIAL_SECOND_ORDER",
        "TIME_INTEGRATION = BACKWARD_EULER_COUPLED", "LINEAR_SOLVER_NS = PETSC_GMRES",
        f"LINEAR_SOLVER_NS_TRESHOLDS {ls}",
        "TURB_MODEL = SA", f"INITIAL_CONDITION_TURB = {nu_sa!r}", "CHECK_INTERVAL = 1",
        f"RESTART = {restart}", "CFL = 1.0", f"CFL_RAMP {cfl_ramp} MAX_CFL={cfl_max!r}",
        f"NSTEPS = {nsteps}", f"RESID_ENERG_TH = {resid_th!r}",
        f"U_INITIAL = {u!r} 0.0 0.0",
        f"inlet = CBC {u!r} 0.0 0.0 1.0 1.0",
        "outlet = CBC_SUBSONIC_OUTLET 1.0",
        "symm_up = SYMMETRY",
        "plate = WALL",
        f"top = CBC {u!r} 0.0 0.0 1.0 1.0",
        "zsym = SYMMETRY",
        f"inlet.nuSA = {nu_sa!r}", f"top.nuSA = {nu_sa!r}", "plate.nuSA = 0.0",
    ]
    if turb == "SST":
        k = 1.5 * (0.01 * u) ** 2
        w = k / mu
        lines = [ln for ln in lines if ".nuSA" not in ln and not ln.startswith(("TURB_MODEL", "INITIAL_CONDITION"))]
        lines = [ln.replace("BACKWARD_EULER_COUPLED", "BACKWARD_EULER").replace("PETSC_GMRES", "BCGSTAB")
                 for ln in lines]
        lines += ["TURB_MODEL = SST", f"INITIAL_CONDITION_TURB = {k!r} {w!r}",
                  f"inlet.kine = {k!r}", f"inlet.omega = {w!r}", f"top.kine = {k!r}", f"top.omega = {w!r}",
                  "plate.kine = 0.0", "plate.omega = HOOK"]
    elif turb != "SA":
        raise ValueError(turb)
    return mesh, "\n".join(lines) + "\n"


def plate_uniform(ncv: int):
    """The plate's free stream as the initial state (rho, rhou, rhoE)."""
    u = MACH_PLATE * math.sqrt(GAMMA)
    rhou = np.zeros((ncv, 3))
    rhou[:, 0] = u
    return np.ones(ncv), rhou, np.full(ncv, 1.0 / (GAMMA - 1.0) + 0.5 * u * u)
