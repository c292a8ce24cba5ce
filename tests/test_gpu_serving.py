"""The serving loop (jm_server_*): solves admitted by footprint, run
concurrently on their own streams with graph-captured steps, or batched (in
lockstep, each kernel of the step and of the BiCGStab launched once for all
active solves). Each result must be the same bits as the same solve run
alone through the ordinary step."""
import numpy as np
import pytest

from myjoe_amd import cases, joemi

pytestmark = pytest.mark.gpu

REQS = [dict(nx=48, ny=16, thickness=0.10, mach=0.45),
        dict(nx=96, ny=32, thickness=0.07, mach=0.35),
        dict(nx=60, ny=20, thickness=0.12, mach=0.5)]


def _case(r, nsteps=40):
    return cases.bump_case(r["nx"], r["ny"], nsteps=nsteps, thickness=r["thickness"], mach=r["mach"],
                           resid_rel=1e-12)


def _alone(mesh, joe, precision):
    s = joemi.Solver(mesh, joe, precision)
    s.init_uniform()
    n, hist = s.run(100000)
    return s.state(), n, hist


def test_results_same_alone_and_in_the_mix():
    print("serving: mix of three", flush=True)
    built = []
    for k, r in enumerate(REQS):
        mm, joe = _case(r)
        built.append((joemi.Mesh.from_faces(mm), joe, joemi.JM_MIXED if k == 1 else joemi.JM_FP64))
    srv = joemi.Server(mem_budget_bytes=8 << 30, max_active=3)
    tickets = [srv.submit(m, joe, prec) for (m, joe, prec) in built]
    for t in tickets:
        srv.wait(t)
    for t, (m, joe, prec) in zip(tickets, built):
        (rho, rhou, rhoE), rep = srv.result(t)
        assert rep.status == joemi.Server.DONE and rep.steps == 40
        (r2, ru2, rE2), n2, hist2 = _alone(m, joe, prec)
        assert n2 == 40
        assert np.array_equal(rho, r2) and np.array_equal(rhou, ru2) and np.array_equal(rhoE, rE2)
        assert list(rep.resid) == list(hist2[-1])
        srv.release(t)
    srv.close()


def test_admission_respects_the_memory_budget():
    print("serving: budget", flush=True)
    mm, joe = _case(REQS[1], nsteps=10)
    mesh = joemi.Mesh.from_faces(mm)
    # a budget below two footprints: the second solve waits for the first
    fp = mesh.footprint()
    srv = joemi.Server(mem_budget_bytes=int(1.5 * fp), max_active=4)
    t1 = srv.submit(mesh, joe)
    t2 = srv.submit(mesh, joe)
    srv.wait(t2)
    (a, _, _), rep1 = srv.result(t1)
    (b, _, _), rep2 = srv.result(t2)
    assert rep2.queued_s >= rep1.run_s * 0.5
    assert np.array_equal(a, b)
    srv.close()
    small = joemi.Server(mem_budget_bytes=fp // 2, max_active=1)
    with pytest.raises(joemi.JoemiError):
        small.submit(mesh, joe)
    small.close()


def test_bad_request_is_rejected_at_submission():
    mm, joe = _case(REQS[2], nsteps=5)
    mesh = joemi.Mesh.from_faces(mm)
    srv = joemi.Server(mem_budget_bytes=1 << 30, max_active=2)
    with pytest.raises(joemi.JoemiError) as e:
        srv.submit(mesh, joe.replace("bump_wall = WALL", "bump_wall = NOSUCHBC"))
    assert e.value.status == 3
    srv.close()


def _served(built, batched, max_active):
    srv = joemi.Server(mem_budget_bytes=8 << 30, max_active=max_active, batched=batched)
    tickets = [srv.submit(m, joe, prec) for (m, joe, prec) in built]
    for t in tickets:
        srv.wait(t)
    out = []
    for t in tickets:
        state, rep = srv.result(t)
        assert rep.status == joemi.Server.DONE, rep.status
        out.append((state, rep))
        srv.release(t)
    srv.close()
    return out


def test_batched_results_same_alone_and_in_the_mix():
    # batched serving: solves in lockstep, each kernel launched once for all
    # of them; every solve keeps its own blocks and reduction order,
    # so a result is the same bits alone (a batch of one) as in a batch of three
    print("serving: batched mix", flush=True)
    built = []
    for k, r in enumerate(REQS):
        mm, joe = _case(r, nsteps=25)
        built.append((joemi.Mesh.from_faces(mm), joe, joemi.JM_MIXED if k == 1 else joemi.JM_FP64))
    mix = _served(built, True, 3)
    for k, b in enumerate(built):
        (alone,) = _served([b], True, 1)
        (s1, r1), (s2, r2) = mix[k], alone
        assert r1.steps == r2.steps == 25
        for a, c in zip(s1, s2):
            assert np.array_equal(a, c), k
        assert list(r1.resid) == list(r2.resid)


@pytest.mark.parametrize("precision", [joemi.JM_FP64, joemi.JM_MIXED])
def test_batched_step_equals_the_ordinary_step(precision):
    # a solve in a batch is cut into the same blocks and reduced in the same
    # order as its own host-polled solve: the same bits after 25 steps
    print(f"serving: batched vs ordinary precision={precision}", flush=True)
    built = []
    for r in REQS:
        mm, joe = _case(r, nsteps=25)
        built.append((joemi.Mesh.from_faces(mm), joe, precision))
    served = _served(built, True, 3)
    for (state, rep), (mesh, joe, prec) in zip(served, built):
        (r2, ru2, rE2), n2, hist2 = _alone(mesh, joe, prec)
        assert rep.steps == n2 == 25
        for a, b in zip(state, (r2, ru2, rE2)):
            assert np.array_equal(a, b)
        assert list(rep.resid) == list(hist2[-1])


def test_batched_coupled_step_equals_the_ordinary_step():
    # coupled-loop requests (BACKWARD_EULER_COUPLED) batch as well: the same
    # bits as their own host-polled solves, beside segregated requests
    from myjoe_amd import cases

    print("serving: batched coupled", flush=True)
    built = []
    for k, r in enumerate(REQS):
        mm, joe = _case(r, nsteps=25)
        built.append((joemi.Mesh.from_faces(mm), cases.coupled(joe) if k != 1 else joe, joemi.JM_MIXED))
    served = _served(built, True, 3)
    for (state, rep), (mesh, joe, prec) in zip(served, built):
        (r2, ru2, rE2), n2, hist2 = _alone(mesh, joe, prec)
        assert rep.steps == n2 == 25
        for a, b in zip(state, (r2, ru2, rE2)):
            assert np.array_equal(a, b)
        assert list(rep.resid) == list(hist2[-1])


def test_batched_rejects_what_it_does_not_run():
    # the batched step carries no SA scalar and runs the per-entry flux
    # kernel only: such requests fail at submission, not mid-run
    from myjoe_amd import meshgen
    from tests.test_oracle_golden import joe_text, load

    srv = joemi.Server(mem_budget_bytes=1 << 30, max_active=2, batched=True)
    g = load("solve_bump24sa")
    sa_mesh = joemi.Mesh.from_faces(meshgen.bump_channel(24, 8))
    with pytest.raises(joemi.JoemiError) as e:
        srv.submit(sa_mesh, joe_text(g))
    assert e.value.status == 5
    mm, joe = _case(REQS[0], nsteps=5)
    mesh = joemi.Mesh.from_faces(mm)
    with pytest.raises(joemi.JoemiError) as e:
        srv.submit(mesh, joe + "GPU_FLUX_KERNEL = FACE\n")
    assert e.value.status == 5
    t = srv.submit(mesh, joe)  # the same case with the default kernels runs
    srv.wait(t)
    assert srv.result(t)[1].status == joemi.Server.DONE
    srv.close()


def test_batched_request_joining_a_running_group_matches_its_alone_result():
    # one lockstep group: a long solve is stepping before the second request
    # is submitted (its progress is polled until it has taken 10 steps), so
    # the second joins a group already stepping; both results are still the
    # same bits as each solve alone
    import os

    print("serving: batched join", flush=True)
    keep = os.environ.get("JOEMI_SERVER_GROUPS")
    os.environ["JOEMI_SERVER_GROUPS"] = "1"
    try:
        mm_a, joe_a = _case(REQS[1], nsteps=120)
        mm_b, joe_b = _case(REQS[0], nsteps=30)
        ma, mb = joemi.Mesh.from_faces(mm_a), joemi.Mesh.from_faces(mm_b)
        srv = joemi.Server(mem_budget_bytes=8 << 30, max_active=2, batched=True)
        ta = srv.submit(ma, joe_a, joemi.JM_MIXED)
        import time

        t0 = time.time()
        while srv.progress(ta) < 10:
            assert srv.status(ta) <= joemi.Server.RUNNING and time.time() - t0 < 60
            time.sleep(0.001)
        joined_at = srv.progress(ta)
        tb = srv.submit(mb, joe_b, joemi.JM_FP64)
        srv.wait(tb)
        srv.wait(ta)
        (sa, ra), (sb, rb) = srv.result(ta), srv.result(tb)
        srv.release(ta)
        srv.release(tb)
        srv.close()
    finally:
        if keep is None:
            del os.environ["JOEMI_SERVER_GROUPS"]
        else:
            os.environ["JOEMI_SERVER_GROUPS"] = keep
    assert 10 <= joined_at < 120 and ra.steps == 120 and rb.steps == 30
    for (st, rep_), (m, joe, prec) in (((sa, ra), (ma, joe_a, joemi.JM_MIXED)), ((sb, rb), (mb, joe_b, joemi.JM_FP64))):
        ((alone, arep),) = _served([(m, joe, prec)], True, 1)
        for a, c in zip(st, alone):
            assert np.array_equal(a, c)
        assert list(rep_.resid) == list(arep.resid)


def test_served_mixed_solve_reaches_the_oracle_steady_state():
    # a solve served by the batched server in the mixed form (FP32 Jacobian,
    # preconditioner and Krylov vectors; FP64 residual) converges to the FP64
    # oracle's steady state of the 24x8 bump to 1e-8 of each variable's
    # largest magnitude: batching and precision change the path, not the
    # fixed point
    from oracle.oracle import Oracle

    print("serving: mixed steady state", flush=True)
    mm, joe = cases.bump_case(24, 8, nsteps=1200)
    mesh = joemi.Mesh.from_faces(mm)
    o = Oracle(mesh.as_dump(), joe)
    u = cases.U_BUMP
    o.set_state(np.ones(mesh.ncv), np.tile([u, 0.0, 0.0], (mesh.ncv, 1)),
                np.full(mesh.ncv, 1.0 / 0.4 + 0.5 * u * u))
    n, hist = o.run(1200)
    assert hist[-1, 4] < 1e-7 * hist[0, 4]
    ((served, rep),) = _served([(mesh, joe, joemi.JM_MIXED)], True, 1)
    assert rep.steps == 1200
    for a, b, k in zip(served, o.state(), ("rho", "rhou", "rhoE")):
        err = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
        print("served mixed vs oracle", k, f"{err:.1e}", flush=True)
        assert err <= 1e-8, (k, err)


def _uniform_with_a_nan(mesh, joe):
    s = joemi.Solver(mesh, joe, joemi.JM_FP64)
    s.init_uniform()
    rho, rhou, rhoE = (np.array(a) for a in s.state())
    del s
    rho[len(rho) // 2] = np.nan
    return rho, rhou, rhoE


@pytest.mark.parametrize("batched", [False, True])
def test_non_finite_mixed_solve_is_retried_once_in_fp64(batched):
    # a mixed solve whose first residual is NaN (one NaN density in its
    # initial state) is queued again in FP64 from that same state; that run
    # is NaN too and retires, not queued a third time. With fp64_retry=False,
    # or for an FP64 request, there is no retry
    print(f"serving: NaN retry batched={batched}", flush=True)
    mm, joe = _case(REQS[0], nsteps=20)
    mesh = joemi.Mesh.from_faces(mm)
    init = _uniform_with_a_nan(mesh, joe)
    for retry, prec, want in ((True, joemi.JM_MIXED, 1), (False, joemi.JM_MIXED, 0), (True, joemi.JM_FP64, 0)):
        srv = joemi.Server(mem_budget_bytes=4 << 30, max_active=2, batched=batched, fp64_retry=retry)
        t = srv.submit(mesh, joe, prec, state=init)
        srv.wait(t)
        st, rep = srv.result(t)
        assert rep.status == joemi.Server.DONE, (retry, prec)
        assert rep.fp64_retries == want and rep.steps_abandoned == want, (retry, prec, rep.fp64_retries)
        assert rep.steps == 1 and not np.isfinite(rep.resid[4])
        srv.release(t)
        srv.close()


def test_diverging_mixed_solve_finishes_as_its_fp64_solve():
    # request 51 of the serving bench's 64 (seed 7): the 168 x 56 bump at
    # Mach 0.494, bump 9.8 %, CFL ramp to 20. Its mixed solve's FP32 BiCGStab
    # ended step 184 at the iteration limit on a far-off iterate (minimum
    # pressure 0.87 -> 0.05) and the solve went non-finite at step 185, while
    # the FP64 solve stays finite. Served with the retry, the result is the
    # FP64 solve's bits and the abandoned attempt's steps are reported
    print("serving: diverging mixed solve", flush=True)
    r = cases.serving_requests(64, 7, 3000, 12000, rate_per_s=0.0)[51]
    mm, joe = cases.bump_case(r["nx"], r["ny"], nsteps=300, thickness=r["thickness"], mach=r["mach"],
                              resid_rel=1e-12, cfl_max=20.0)
    mesh = joemi.Mesh.from_faces(mm)
    ref, n64, _ = _alone(mesh, joe, joemi.JM_FP64)
    assert n64 == 300 and all(np.all(np.isfinite(a)) for a in ref)
    outcome = {}
    for retry in (False, True):
        srv = joemi.Server(mem_budget_bytes=4 << 30, max_active=2, batched=True, fp64_retry=retry)
        t = srv.submit(mesh, joe, joemi.JM_MIXED)
        srv.wait(t)
        outcome[retry] = srv.result(t)
        srv.release(t)
        srv.close()
    st0, rep0 = outcome[False]
    diverged = not np.isfinite(rep0.resid[4])
    print("mixed alone diverged:", diverged, "at step", rep0.steps, flush=True)
    st1, rep1 = outcome[True]
    assert rep1.fp64_retries == int(diverged)
    if diverged:
        assert rep1.steps_abandoned == rep0.steps and rep1.steps == 300
        for a, b in zip(st1, ref):
            assert np.array_equal(a, b)
    else:
        for a, b in zip(st1, st0):
            assert np.array_equal(a, b)
