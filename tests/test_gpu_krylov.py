"""The FP32 BiCGStab's residual by its recurrence (GPU_KRYLOV_RESIDUAL =
RECURRENCE, the default for FP32 solves): r = s - omega t, where the
reference recomputes rhs - A phi every iteration (UgpWithCv.cpp:2812-2846),
one product with A fewer per iteration. It changes the linear solves' path
only: a mixed-precision steady solve reaches the same steady state as with
the true residual (GPU_KRYLOV_RESIDUAL = TRUE), and the FP64 solve, which
always takes the reference's form, is not affected by the key."""
import numpy as np
import pytest

from myjoe_amd import cases, joemi

pytestmark = pytest.mark.gpu


def test_recurrence_residual_reaches_the_same_steady_state():
    print("krylov recurrence steady", flush=True)
    mm, joe = cases.bump_case(24, 8, nsteps=1200)
    mesh = joemi.Mesh.from_faces(mm)
    states = []
    for text in (joe + "GPU_KRYLOV_RESIDUAL = TRUE\n", joe):
        s = joemi.Solver(mesh, text, joemi.JM_MIXED)
        s.init_uniform()
        n, hist = s.run(1200)
        assert hist[-1, 4] < 1e-6 * hist[0, 4]
        states.append(s.state())
    for a, b, k in zip(states[0], states[1], ("rho", "rhou", "rhoE")):
        d = float(np.max(np.abs(a - b)) / np.max(np.abs(a)))
        print("recurrence steady", k, f"{d:.1e}", flush=True)
        assert d <= 1e-8, (k, d)


def test_recurrence_key_leaves_the_fp64_solve_alone():
    mm, joe = cases.bump_case(24, 8, nsteps=5)
    mesh = joemi.Mesh.from_faces(mm)
    out = []
    for text in (joe + "GPU_KRYLOV_RESIDUAL = TRUE\n", joe):
        s = joemi.Solver(mesh, text, joemi.JM_FP64)
        s.init_uniform()
        s.run(5)
        out.append(s.state())
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


def test_batched_server_takes_the_recurrence_for_fp32_solves():
    # the batched FP32 BiCGStab takes the recurrence: a mixed request asking
    # for the true residual is refused; FP64 requests take the true residual
    # whatever the key says
    mm, joe = cases.bump_case(24, 8, nsteps=5)
    mesh = joemi.Mesh.from_faces(mm)
    srv = joemi.Server(1 << 30, 2, batched=True)
    with pytest.raises(joemi.JoemiError) as e:
        srv.submit(mesh, joe + "GPU_KRYLOV_RESIDUAL = TRUE\n", joemi.JM_MIXED)
    assert e.value.status == 5 and "RECURRENCE" in str(e.value)
    t = srv.submit(mesh, joe + "GPU_KRYLOV_RESIDUAL = TRUE\n", joemi.JM_FP64)
    srv.wait(t)
    srv.release(t)
    srv.close()
