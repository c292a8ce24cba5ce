"""The C++ JoeWithModels mirror (myjoe_amd/host) driving the GPU solve through
the C ABI, with case classes written as a reference joe.cpp writes them
(myjoe_amd/host/cases/joe_cases.cpp).

Each hook is checked against a run that needs no hook and must give the same
bits: a HOOK inlet that imposes the CBC inflow state, a sourceHook override
that calls the default gravity hook (host) against the device's own gravity,
and an initialHook override against the same initial state set through the
ABI. The split step that hooks use must reproduce the fused step exactly.
"""
import os
import subprocess

import numpy as np
import pytest

from myjoe_amd import cases, joemi, meshgen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "myjoe_amd", "bin", "joe_cases")
NSTEPS = 30


def write_case(d, joe):
    mm, _ = cases.bump_case(24, 8)
    meshgen.write_msh(mm, os.path.join(d, "bump.msh"))
    with open(os.path.join(d, "Joe.in"), "w") as f:
        f.write(joe)
    return mm


def run_case(d, case, expect_rc=0, expect_out=None):
    out = os.path.join(d, case + ".bin")
    r = subprocess.run([EXE, case, "Joe.in", out], cwd=d, capture_output=True, text=True, timeout=120)
    assert r.returncode == expect_rc, r.stdout + r.stderr
    if expect_out:
        assert expect_out in r.stdout, r.stdout
    raw = np.fromfile(out, np.float64)
    n = int(raw[0])
    hist = raw[1:1 + 5 * n].reshape(n, 5)
    rest = raw[1 + 5 * n:]
    ncv = rest.size // 5
    return n, hist, rest[:ncv], rest[ncv:4 * ncv].reshape(ncv, 3), rest[4 * ncv:]


def joe_text(**kw):
    _, joe = cases.bump_case(24, 8, nsteps=NSTEPS)
    for k, v in kw.items():
        joe = joe.replace(k, v)
    return joe


def test_case_program_reports_missing_device(tmp_path):
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    write_case(tmp_path, joe_text())
    r = subprocess.run([EXE, "plain", "Joe.in", "o.bin"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 1 and "status 4" in r.stderr


def test_case_program_reports_missing_restart_file(tmp_path):
    # RESTART may name a restart file of the reference's format; a missing one is an error
    write_case(tmp_path, joe_text(**{"RESTART = bump.msh": "RESTART = restart.out"}))
    r = subprocess.run([EXE, "plain", "Joe.in", "o.bin"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 1 and "restart file" in r.stderr, r.stderr


@pytest.mark.gpu
def test_plain_case_matches_abi_run(tmp_path):
    print("host mirror: plain", flush=True)
    joe = joe_text()
    mm = write_case(tmp_path, joe)
    n, hist, rho, rhou, rhoE = run_case(tmp_path, "plain")
    assert n == NSTEPS
    mesh = joemi.Mesh.from_faces(mm)
    s = joemi.Solver(mesh, joe)
    s.init_uniform()
    n2, hist2 = s.run(NSTEPS)
    assert n2 == NSTEPS
    assert np.array_equal(hist, hist2)
    r2, ru2, rE2 = s.state()
    assert np.array_equal(rho, r2) and np.array_equal(rhou, ru2) and np.array_equal(rhoE, rE2)


@pytest.mark.gpu
def test_boundary_hook_equals_cbc_inflow(tmp_path):
    print("host mirror: boundaryHook", flush=True)
    u = repr(cases.U_BUMP)
    write_case(tmp_path, joe_text())
    ref = run_case(tmp_path, "plain")
    hooked = joe_text(**{f"inlet = CBC {u} 0.0 0.0 1.0 1.0": "inlet = HOOK"}) + \
        f"INLET_HOOK_T = 1.0\nINLET_HOOK_U = {u}\nINLET_HOOK_P = 1.0\n"
    assert "inlet = HOOK" in hooked
    write_case(tmp_path, hooked)
    got = run_case(tmp_path, "inlethook")
    assert got[0] == ref[0]
    for a, b in zip(got[1:], ref[1:]):
        assert np.array_equal(a, b)


@pytest.mark.gpu
def test_source_hook_override_equals_device_gravity(tmp_path):
    print("host mirror: sourceHook", flush=True)
    joe = joe_text() + "GX = 0.05\nGY = -0.2\nGZ = 0.0\n"
    write_case(tmp_path, joe)
    dev = run_case(tmp_path, "plain")
    host = run_case(tmp_path, "gravity")
    nog = joe_text()
    write_case(tmp_path, nog)
    none = run_case(tmp_path, "plain")
    assert not np.array_equal(dev[2], none[2])  # gravity acts
    assert host[0] == dev[0]
    for a, b in zip(host[1:], dev[1:]):
        assert np.array_equal(a, b)


@pytest.mark.gpu
def test_coupled_case_matches_abi_run_and_ignores_source_hook(tmp_path):
    # TIME_INTEGRATION = BACKWARD_EULER_COUPLED through the mirror: the ABI's
    # coupled run bit for bit; the reference's coupled loop calls
    # sourceHookCoupled (default: gravity), not an overridden sourceHook, so
    # the case with a host sourceHook equals the plain one
    print("host mirror: coupled", flush=True)
    joe = cases.coupled(joe_text() + "GX = 0.05\nGY = -0.2\nGZ = 0.0\n")
    mm = write_case(tmp_path, joe)
    plain = run_case(tmp_path, "plain")
    # the case program checks that its sourceHook ran every step and exits
    # with 3 otherwise, after writing its state: here it must not have run
    host = run_case(tmp_path, "gravity", expect_rc=3, expect_out="hook_calls=0")
    assert host[0] == plain[0] == NSTEPS
    for a, b in zip(host[1:], plain[1:]):
        assert np.array_equal(a, b)
    s = joemi.Solver(joemi.Mesh.from_faces(mm), joe)
    s.init_uniform()
    n2, hist2 = s.run(NSTEPS)
    assert n2 == NSTEPS and np.array_equal(plain[1], hist2)
    for a, b in zip(plain[2:], s.state()):
        assert np.array_equal(a, b)


@pytest.mark.gpu
def test_initial_hook_override(tmp_path):
    print("host mirror: initialHook", flush=True)
    joe = joe_text()
    mm = write_case(tmp_path, joe)
    n, hist, rho, rhou, rhoE = run_case(tmp_path, "perturb")
    mesh = joemi.Mesh.from_faces(mm)
    x = mesh.array("X_CV").reshape(-1, 3)
    s = joemi.Solver(mesh, joe)
    s.init_uniform()
    r0, ru0, rE0 = s.state()
    f = 1.0 + 0.01 * np.sin(3.0 * x[:, 0]) * np.cos(2.0 * x[:, 1])
    s.set_state(r0 * f, ru0 * f[:, None], rE0 * f)
    n2, hist2 = s.run(NSTEPS)
    assert n == n2
    # libm sin/cos of the case program and numpy may differ in the last bit
    np.testing.assert_allclose(hist, hist2, rtol=1e-9, atol=0)
    np.testing.assert_allclose(rho, s.state()[0], rtol=1e-9)


@pytest.mark.gpu
def test_sa_mixin_case_matches_abi_run(tmp_path):
    # a case class deriving from RansTurbSA (joe.cpp's MyJoeSA) solves nuSA:
    # the Joe.in carries no model keyword, the class does
    print("host mirror: sa", flush=True)
    joe = joe_text(**{"MU_REF = 0.0": "MU_REF = 0.002\nMU_MODE = POWERLAW\nMU_POWER_LAW = 0.76\nPR = 0.72"})
    joe_sa = cases.with_sa(joe, 0.006, cases.SA_BUMP)
    joe_case = joe_sa.replace("TURB_MODEL = SA\n", "")
    mm = write_case(tmp_path, joe_case)
    out = os.path.join(tmp_path, "sa.bin")
    r = subprocess.run([EXE, "sa", "Joe.in", out], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.fromfile(out, np.float64)
    n = int(raw[0])
    ncv = mm.ncv
    hist = raw[1:1 + 5 * n].reshape(n, 5)
    rest = raw[1 + 5 * n:]
    rho, rhou, rhoE, nu = rest[:ncv], rest[ncv:4 * ncv], rest[4 * ncv:5 * ncv], rest[5 * ncv:6 * ncv]
    s = joemi.Solver(joemi.Mesh.from_faces(mm), joe_sa)
    s.init_uniform()
    n2, hist2 = s.run(NSTEPS)
    assert n == n2 == NSTEPS
    assert np.array_equal(hist, hist2)
    r2, ru2, rE2 = s.state()
    assert np.array_equal(rho, r2) and np.array_equal(rhou, ru2.reshape(-1)) and np.array_equal(rhoE, rE2)
    assert np.array_equal(nu, s.get("NUSA"))


@pytest.mark.gpu
def test_sst_mixin_case_matches_abi_run(tmp_path):
    # a case class deriving from RansTurbKOmSST (joe.cpp's MyJoeSST) solves
    # kine and omega: the Joe.in carries no model keyword, the class does
    print("host mirror: sst", flush=True)
    joe = joe_text(**{"MU_REF = 0.0": "MU_REF = 0.002\nMU_MODE = POWERLAW\nMU_POWER_LAW = 0.76\nPR = 0.72"})
    joe_sst = cases.with_sst(joe, 0.001, 5.0, cases.SST_BUMP)
    joe_case = joe_sst.replace("TURB_MODEL = SST\n", "")
    mm = write_case(tmp_path, joe_case)
    out = os.path.join(tmp_path, "sst.bin")
    r = subprocess.run([EXE, "sst", "Joe.in", out], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.fromfile(out, np.float64)
    n = int(raw[0])
    ncv = mm.ncv
    hist = raw[1:1 + 5 * n].reshape(n, 5)
    rest = raw[1 + 5 * n:]
    rho, rhou, rhoE = rest[:ncv], rest[ncv:4 * ncv], rest[4 * ncv:5 * ncv]
    kine, omega = rest[5 * ncv:6 * ncv], rest[6 * ncv:7 * ncv]
    s = joemi.Solver(joemi.Mesh.from_faces(mm), joe_sst)
    s.init_uniform()
    n2, hist2 = s.run(NSTEPS)
    assert n == n2 == NSTEPS
    assert np.array_equal(hist, hist2)
    r2, ru2, rE2 = s.state()
    assert np.array_equal(rho, r2) and np.array_equal(rhou, ru2.reshape(-1)) and np.array_equal(rhoE, rE2)
    assert np.array_equal(kine, s.get("KINE")) and np.array_equal(omega, s.get("OMEGA"))


@pytest.mark.gpu
def test_case_resumes_from_its_restart_file(tmp_path):
    # run() writes restart.out at its end (runBackwardEuler); a second run with
    # RESTART = restart.out reads mesh, state and STEP from it and carries on.
    # The C ABI resumed from the same file gives the same bits.
    print("host mirror: restart", flush=True)
    write_case(tmp_path, joe_text(**{"NSTEPS = 30": "NSTEPS = 20"}))
    n1, _, _, _, _ = run_case(tmp_path, "plain")
    assert n1 == 20 and (tmp_path / "restart.out").exists()
    first = joemi.Mesh.read_restart(str(tmp_path / "restart.out"))
    assert first.restart_step()[0] == 20
    (tmp_path / "restart20.out").write_bytes((tmp_path / "restart.out").read_bytes())
    joe2 = joe_text(**{"NSTEPS = 30": "NSTEPS = 25", "RESTART = bump.msh": "RESTART = restart.out"})
    (tmp_path / "Joe.in").write_text(joe2)
    n2, _, rho, rhou, rhoE = run_case(tmp_path, "plain")
    assert n2 == 5
    mesh = joemi.Mesh.read_restart(str(tmp_path / "restart20.out"))
    s = joemi.Solver(mesh, joe2)
    s.set_state(mesh.restart_field("RHO"), mesh.restart_field("RHOU"), mesh.restart_field("RHOE"))
    s.set_step(20)
    for _ in range(5):
        s.step()
    r2, ru2, rE2 = s.state()
    assert np.array_equal(rho, r2) and np.array_equal(rhou, ru2) and np.array_equal(rhoE, rE2)
    assert joemi.Mesh.read_restart(str(tmp_path / "restart.out")).restart_step()[0] == 25


@pytest.mark.gpu
def test_case_class_reads_wall_values_from_host_copies(tmp_path):
    # a case class constructed as joe.cpp's MyJoe is (naming the virtual base
    # UgpWithCvCompFlow) forms each wall face's shear, y+, heat flux and
    # viscosity in finalHook from the mirror's host copies (mul_fa, the state,
    # R_gas, gamma, Pr, faZoneList, getParam); they equal the device's
    # jm_wall_values of the same solve, and its temporalHook saw a fresh host
    # state after every step
    print("host mirror: wall values", flush=True)
    visc = "MU_REF = 0.002\nMU_MODE = POWERLAW\nMU_POWER_LAW = 0.76\nPR = 0.72"
    joe = joe_text(**{"MU_REF = 0.0": visc})
    mm = write_case(tmp_path, joe)
    r = subprocess.run([EXE, "walls", "Joe.in", "w.bin"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"fresh={NSTEPS}" in r.stdout, r.stdout
    raw = np.fromfile(tmp_path / "w.bin", np.float64)
    n = int(raw[0])
    assert n == NSTEPS
    rest = raw[1 + 5 * n + 5 * mm.ncv:]
    nw = int(rest[0])
    walls = rest[1:].reshape(nw, 4)
    s = joemi.Solver(joemi.Mesh.from_faces(mm), joe)
    s.init_uniform()
    s.run(NSTEPS)
    wv = s.wall_values("bump_wall")
    assert wv.shape[0] == nw
    from tests.helpers import compare
    for k, col in ((0, 6), (1, 8), (2, 7), (3, 9)):  # tau_wall, y+, heat flux, mul
        compare(walls[:, k], wv[:, col], 1e-12, f"wall column {col}")
    assert np.all(walls[:, 3] > 0.0) and np.any(walls[:, 0] != 0.0)
