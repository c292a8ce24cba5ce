#!/usr/bin/env python3
"""Single-residual throughput on one MI355X.

Workload: [0,1]^3 box of 160^3 hexes, a seeded 50 % split into six Kuhn
tetrahedra each, interior nodes jittered by U(-0.2h, 0.2h); random FP64 state
(rho ~ U(0.5,1.5), p ~ lognormal(0,0.2), u_i ~ N(0,0.3) c). One "step" is one
residual evaluation of the reference's pseudo-time loop -- calcStateVariables,
setBC, calcDt and calcRhs (2nd order, least-squares gradients, smooth
Barth-Jespersen limiter) -- over the whole mesh, with the state resident in
HBM. value = ncv * steps / time in Mcell-it/s for HLLC; Roe, AUSM and the
coupled loop's HLLC residual (calcRhsCoupled) are reported beside it.

    python bench.py [--steps K] [--warmup W] [--n 160] [--gpus 1]
"""
from __future__ import annotations

import argparse
import gc
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "steady solves/min + p50/p99 latency vs plain FP64 GPU loop; residual Mcell-it/s"


def algorithmic_bytes(ncv, nfa, nfa_b, nnz):
    """Compulsory HBM bytes of each residual kernel (every array it needs read
    or written once): prim rows 64 B (rho u v w p x_cv), limited-gradient
    rows 120 B (+8 alpha), face rows 56 B (unit normal, area, x_fa), boundary
    rows 40 B (rho u v w p), face-list entries 8 B, LS coefficients 24 B."""
    nfa_i = nfa - nfa_b
    nfoc = 2 * nfa_i + nfa_b
    return {
        "k_state": ncv * (40 + 40),
        "k_bc": nfa_b * (64 + 56 + 8 + 56),
        "k_grad": ncv * (64 + 4 + 4 + 128) + nnz * 24 + nfoc * 8 + nfa_i * 56 + nfa_b * (40 + 24),
        # fluxes + local time step: + rhou (24 B) and dt (8 B) per cell
        "k_flux": ncv * (64 + 120 + 4 + 8 + 40 + 24 + 8) + nfoc * 8 + nfa * 56 + nfa_b * 40,
        # Jacobian assembly, FP64 blocks: A written once, per cell its prim
        # row, entry range, diagonal index; per entry the face entry and the
        # block index; face rows; boundary rows
        "k_jac": nnz * 25 * 8 + ncv * (64 + 4 + 8) + nfoc * 12 + nfa * 32 + nfa_b * 40,
    }


def measured_traffic(mesh, kernel):
    """HBM bytes per launch of a kernel on this mesh from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (profiles/), when it was taken
    on the same mesh; None otherwise."""
    path = os.path.join(ROOT, "profiles", "residual_160box_hllc_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
    except OSError:
        return None, None
    m = d.get("mesh", {})
    if (m.get("ncv"), m.get("nfa"), m.get("nfa_b"), m.get("nnz")) != (mesh.ncv, mesh.nfa, mesh.nfa_b, mesh.nnz):
        return None, None
    k = d.get("kernels", {}).get(kernel)
    if not k:
        return None, None
    return k["fetch_bytes"] + k["write_bytes"], os.path.relpath(path, ROOT)


def _time_reference(mm, joe, state, reps):
    """Seconds per residual of the reference itself (oracle/_ref/joe_ref_dump,
    built from the reference's sources; its 'time' mode runs state, BC, dt and
    an explicit calcRhs), or None where that binary is absent."""
    import subprocess
    import tempfile

    from myjoe_amd import meshgen
    from oracle import jmdump

    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle", "_ref", "joe_ref_dump")
    if not os.access(exe, os.X_OK):
        return None
    with tempfile.TemporaryDirectory() as wd:
        meshgen.write_msh(mm, os.path.join(wd, "box.msh"))
        with open(os.path.join(wd, "Joe.in"), "w") as f:
            f.write(joe)
        rho, rhou, rhoE = state
        jmdump.write(os.path.join(wd, "state.bin"), {"RHO": rho, "RHOU": rhou.reshape(-1), "RHOE": rhoE})
        r = subprocess.run([exe, "time", "Joe.in", "state.bin", str(reps)], cwd=wd, capture_output=True, text=True,
                           timeout=600)
        for line in r.stdout.splitlines():
            if line.startswith("SECONDS_PER_RESIDUAL"):
                return float(line.split()[1])
    return None


def cpu_baseline(n_box=48, reps=12):
    """CPU residual throughput on a smaller box of the same construction
    (state + BC + dt + explicit calcRhs, HLLC 2nd order LS + limiter): the
    restatement of the reference (bit-identical to it) on every host core
    (one context per thread, up to 16), the headline; beside it one core of
    the restatement and, where oracle/_ref is present, one core of the
    reference itself (it has no threads; the ratio of the two relates the
    restatement's numbers to the reference's)."""
    import threading

    from myjoe_amd import cases, joemi
    from oracle.oracle import Oracle

    mm, joe = cases.box_case(n_box, seed=0)
    mesh = joemi.Mesh.from_faces(mm)
    state = cases.random_state(mesh.ncv, seed=1)
    dump = mesh.as_dump()
    ref_s = _time_reference(mm, joe, state, reps)
    del mm

    def port_run(out, k):
        o = Oracle(dump, joe)
        o.set_state(*state)
        o.residual(False)
        t0 = time.perf_counter()
        for _ in range(reps):
            o.residual(False)
        out[k] = time.perf_counter() - t0

    one = [0.0]
    port_run(one, 0)
    port1 = one[0] / reps
    cores = max(1, min(16, os.cpu_count() or 1))
    dts = [0.0] * cores
    threads = [threading.Thread(target=port_run, args=(dts, k)) for k in range(cores)]
    t0 = time.perf_counter()
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    wall = time.perf_counter() - t0  # context set-up included, so a lower bound of the rate
    rate = mesh.ncv * reps * cores / max(dts) / 1e6
    out = {"value": round(rate, 4), "unit": "Mcell-it/s", "cores": cores, "kind": "port",
           "sample": f"{n_box}^3 split box ({mesh.ncv} cells), {reps} residual evaluations (state, BC, dt, calcRhs; "
                     f"HLLC 2nd order LS+limiter) per core on {cores} host cores, one restatement context per "
                     f"thread ({max(dts):.1f} s, {wall:.1f} s with set-up)",
           "port_1core": round(mesh.ncv / port1 / 1e6, 4)}
    if ref_s is not None:
        out["reference_1core"] = round(mesh.ncv / ref_s / 1e6, 4)
        out["reference_over_port_time_1core"] = round(ref_s / port1, 3)
    return out


def chem_bench(npts=8_000_000, reps=20, cpu_sample=500_000):
    """The FPVA chemistry-table lookup: the reference's LookupCoeff (ten FP32
    coefficients, trilinear with FP64 weights) on a 128 x 32 x 64 table, over
    a field of npts points in cell order (a round jet's mean mixture fraction,
    its variance and a progress variable, smooth along the cells), timed with
    HIP events on the lookup's stream. Compulsory bytes per launch: the points
    (24 B each), the results (80 B each) and the table once (128 x 32 x 64 x
    10 floats = 10.5 MB). The oracle (one core) is timed on cpu_sample points."""
    import numpy as np

    from myjoe_amd import chemtable, joemi

    x1, x2, x3, data, names = chemtable.synthetic_table(128, 32, 64, seed=0)
    m = round(npts ** (1.0 / 3.0))
    n = m * m * m
    g = (np.arange(m) + 0.5) / m
    X, Y, Zc = np.meshgrid(g, g, g, indexing="ij")  # cell centres, x along the jet
    r2 = (Y - 0.5) ** 2 + (Zc - 0.5) ** 2
    width = 0.02 + 0.1 * X
    zm = np.exp(-r2 / (2 * width ** 2)) * (0.05 / (0.05 + X))
    zv = (0.25 * zm * (1 - zm) * np.exp(-4 * X)).reshape(-1)
    pc = (np.clip(X * 1.5, 0, 1) * (1 - np.exp(-r2 / 0.01))).reshape(-1)
    zm = zm.reshape(-1)
    del X, Y, Zc, r2, width
    t = joemi.Chemtable.create(x1, x2, x3, data, names)
    d = [joemi.DeviceArray(v) for v in (zm, zv, pc)]
    out = joemi.DeviceArray(np.zeros((10, n)))
    ms = t.time_lookup(*d, t.COEFF, out, reps=reps)
    sl = np.arange(0, n, max(1, n // 20000))
    k = min(cpu_sample, n)
    sample = {"table": (x1, x2, x3, data), "points": (zm[:k], zv[:k], pc[:k]),
              "check": ((zm[sl], zv[sl], pc[sl]), out.numpy()[:, sl])}
    bytes_ = n * (24 + 80) + 128 * 32 * 64 * 10 * 4
    res = {"workload": f"LookupCoeff (10 FP32 coefficients) on a 128 x 32 x 64 table, {n} points ({m}^3 cells "
                       "of a round jet's Zmean, Zvar, C in cell order)",
           "points": n, "ms": round(ms, 4), "gpoints_per_s": round(n / (ms * 1e-3) / 1e9, 3),
           "bytes_per_launch": int(bytes_), "achieved_GBs": round(bytes_ / (ms * 1e-3) / 1e9, 1),
           "frac_of_hbm_peak": round(bytes_ / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    return res, sample


def chem_cpu_baseline(sample):
    """The oracle's lookup (one core) timed on the sample's points, and the
    device's results at every n/20000-th point checked against it."""
    import numpy as np

    from oracle.oracle import chemtable_lookup

    tab = sample["table"]
    pts = sample["points"]
    t0 = time.perf_counter()
    chemtable_lookup(*tab, *pts, range(10))
    cpu_s = time.perf_counter() - t0
    cpts, dev = sample["check"]
    same = bool(np.array_equal(dev, chemtable_lookup(*tab, *cpts, range(10))))
    return {"cpu_oracle_1core": {"mpoints_per_s": round(pts[0].size / cpu_s / 1e6, 2),
                                 "sample_points": int(pts[0].size)},
            "sampled_points_equal_oracle": same}


def plate_bench(dims=(40, 324, 110, 50), steps=3, reps=10, jac_kernel=None, turb="SA", scal_kernel=None):
    """One implicit step of the 2M-cell flat plate (Mach 0.2, Re 5e6 per unit
    length, Spalart-Allmaras in the coupled loop, 6x6 blocks; with turb =
    "SST" the Menter SST model in the segregated loop, 5x5 blocks and the two
    scalar solves), FP64 and the mixed form: wall-clock ms per step from the
    free stream (host-polled steps, BiCGStab iterations counted), and the
    residual (with SST its kine and omega rows), the Jacobian assembly and
    one BiCGStab iteration timed alone with HIP events, each against its
    compulsory bytes (b = 6 or 5, s = 8, s_J = s_v = 8 or 4):
      Jacobian  nnz b^2 s_J + ncv (b s + 8) + nfa_i (3 s + 24) + nfa_b (3 s + 8)
      iteration nnz (2 b^2 s_J + 8) + ncv (2 b^2 s_J + 24 b s_v + 8)
      residual  780 ncv + 32 nnz + 4 nfoc + 80 nfa_i + 160 nfa_b (the
                inviscid count: a lower bound for this viscous residual)."""
    from myjoe_amd import cases, joemi

    t0 = time.perf_counter()
    mm, joe = cases.plate_case(*dims, turb=turb)
    if jac_kernel:
        joe += f"GPU_JAC_KERNEL = {jac_kernel}\n"
    if scal_kernel and turb == "SST":
        joe += f"GPU_SCAL_KERNEL = {scal_kernel}\n"
    mesh = joemi.Mesh.from_faces(mm)
    del mm
    gc.collect()
    ncv, nfa, nfa_b, nnz = mesh.ncv, mesh.nfa, mesh.nfa_b, mesh.nnz
    nfa_i = nfa - nfa_b
    print(f"# plate: ncv={ncv} nfa={nfa} nfa_b={nfa_b} nnz={nnz} setup {time.perf_counter() - t0:.1f}s",
          file=sys.stderr, flush=True)
    b = 6 if turb == "SA" else 5
    res_bytes = 780 * ncv + 32 * nnz + 4 * (2 * nfa_i + nfa_b) + 80 * nfa_i + 160 * nfa_b
    out = {"workload": f"flat plate {dims[0]}+{dims[1]} x {dims[2]} x {dims[3]} hexes (x ahead of / along the plate, "
                       f"wall-normal at ratio 1.1 with y+ < 1, spanwise), Mach 0.2, Re 5e6 per unit length, "
                       + ("SA in the coupled loop (6x6 blocks)" if turb == "SA" else
                          "Menter SST in the segregated loop (5x5 blocks, kine and omega solved after the flow)")
                       + ", HLLC 2nd order",
           "ncv": ncv, "nfa": nfa, "nfa_b": nfa_b, "nnz": nnz}
    for prec, name in ((joemi.JM_FP64, "fp64"), (joemi.JM_MIXED, "mixed")):
        sj = 8 if prec == joemi.JM_FP64 else 4
        s = joemi.Solver(mesh, joe, prec)
        s.set_state(*cases.plate_uniform(ncv))
        s.step()
        s.sync()
        t1 = time.perf_counter()
        iters = sits = 0
        for _ in range(steps):
            info = s.step()
            iters += info.lin_iters
            sits += info.scal_iters
        s.sync()
        step_ms = (time.perf_counter() - t1) / steps * 1e3
        km = s.time_step_kernels(reps)
        jac_bytes = nnz * b * b * sj + ncv * (b * 8 + 8) + nfa_i * (3 * 8 + 24) + nfa_b * (3 * 8 + 8)
        it_bytes = nnz * (2 * b * b * sj + 8) + ncv * (2 * b * b * sj + 24 * b * sj + 8)
        frac = lambda by, ms: round(by / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        out[name] = {"step_ms": round(step_ms, 3), "mean_bicgstab_iterations": round(iters / steps, 1),
                     **({"mean_scalar_bicgstab_iterations": round(sits / steps, 1)} if turb == "SST" else {}),
                     "kernel_ms": {k: round(v, 4) for k, v in km.items()},
                     "residual_frac_of_hbm_peak": frac(res_bytes, km["residual"]),
                     "jacobian_frac_of_hbm_peak": frac(jac_bytes, km["jacobian"]),
                     "bicgstab_iteration_frac_of_hbm_peak": frac(it_bytes, km["krylov_iteration"]),
                     "bytes": {"residual": res_bytes, "jacobian": jac_bytes, "bicgstab_iteration": it_bytes},
                     "mcell_it_s_residual": round(ncv / (km["residual"] * 1e-3) / 1e6, 1)}
        print(f"# plate {turb} {name}: step {step_ms:.2f} ms ({iters / steps:.1f} BiCGStab its); "
              + " ".join(f"{k}={v:.3f}ms" for k, v in km.items()), file=sys.stderr, flush=True)
        del s
        gc.collect()
    return out


def streams_capacity(built, budget, max_active):
    """Solves/min of the per-stream graph server, all requests present at t = 0."""
    from myjoe_amd import joemi

    srv = joemi.Server(budget, max_active, batched=False)
    t0 = time.perf_counter()
    tickets = [srv.submit(m, joe, joemi.JM_MIXED) for m, joe in built]
    for t in tickets:
        srv.wait(t)
    cap = len(built) / (time.perf_counter() - t0) * 60.0
    for t in tickets:
        srv.release(t)
    srv.close()
    print(f"# serving: per-stream graph server {cap:.1f} solves/min", file=sys.stderr, flush=True)
    return cap


def _strict_json(v):
    """The record with every non-finite float as null (JSON has no NaN)."""
    if isinstance(v, dict):
        return {k: _strict_json(x) for k, x in v.items()}
    if isinstance(v, (list, tuple)):
        return [_strict_json(x) for x in v]
    if isinstance(v, float) and not math.isfinite(v):
        return None
    return v


def serving_bench(n, seed, min_cells, max_cells, max_active, nsteps, resid_rel, budget_gib,
                  legs=("plain", "streams", "batched"), plain_every=4, cfl_max=20.0):
    """Steady solves/min and latency of the served mixed-precision solver
    against a plain FP64 one-at-a-time loop, on steady bump-channel requests
    of random size, bump thickness and Mach (cases.serving_requests), each
    solved until its energy residual has dropped by resid_rel. Capacity: all
    requests present at t = 0. Latency: Poisson arrivals at 0.8 x the served
    capacity, replayed in real time for the server; for the one-at-a-time
    loop the same arrivals are queued in front of its solve times. The plain
    loop solves every plain_every-th request (its rate is one over the mean
    solve time, which a sample of the same random sizes estimates); in its
    queue each measured solve time stands for the plain_every requests from
    it. The served solver is the batched server (lockstep steps, each kernel
    of the step and of the BiCGStab launched once for all active solves); the
    per-stream graph server's capacity is reported beside it. legs names the
    parts to run (plain loop, per-stream graph server, batched server)."""
    import numpy as np

    from myjoe_amd import cases, joemi

    reqs = cases.serving_requests(n, seed, min_cells, max_cells, rate_per_s=0.0)
    built = []
    for r in reqs:
        mm, joe = cases.bump_case(r["nx"], r["ny"], nsteps=nsteps, thickness=r["thickness"], mach=r["mach"],
                                  resid_rel=resid_rel, cfl_max=cfl_max)
        built.append((joemi.Mesh.from_faces(mm), joe))
        del mm
    cells = sum(m.ncv for m, _ in built)
    # plain loop: FP64, one solve at a time, host-polled steps
    service, steps_plain, plain_states = [], [], []
    sampled = list(range(0, n, plain_every))
    for mesh, joe in ([built[k] for k in sampled] if "plain" in legs else []):
        t0 = time.perf_counter()
        s = joemi.Solver(mesh, joe, joemi.JM_FP64)
        s.init_uniform()
        k, _ = s.run(nsteps)
        st = s.state()
        service.append(time.perf_counter() - t0)
        steps_plain.append(k)
        plain_states.append(st)
        del s
    cap_plain = len(service) / sum(service) * 60.0 if service else None
    if service:
        print(f"# serving: plain FP64 loop {cap_plain:.1f} solves/min ({len(service)} of the {n} requests, "
              f"{cells} cells in all)", file=sys.stderr, flush=True)
    budget = int(budget_gib * (1 << 30))
    res = {"workload": f"{n} steady inviscid bump-channel solves, {min_cells}-{max_cells} cells (nx = 3 ny, one "
                       f"cell thick), bump 6-12 %, inlet Mach 0.3-0.5, HLLC 2nd order, CFL ramp to {cfl_max:g}; "
                       f"stop at energy L1 drop {resid_rel:g} or {nsteps} steps", "cells_total": int(cells)}
    if "streams" in legs:
        res["served_per_stream_graphs"] = {"solves_per_min": round(streams_capacity(built, budget, max_active), 2),
                                           "max_active": max_active}
    if "batched" not in legs:
        return res

    # served: batched, mixed precision, all present at t = 0
    srv = joemi.Server(budget, max_active, batched=True)
    t0 = time.perf_counter()
    tickets = [srv.submit(m, joe, joemi.JM_MIXED) for m, joe in built]
    for t in tickets:
        srv.wait(t)
    makespan = time.perf_counter() - t0
    cap_served = n / makespan * 60.0
    steps_served, err, drops = [], {"rho": 0.0, "rhou": 0.0, "rhoE": 0.0}, []
    retries, abandoned = 0, 0
    for k, t in enumerate(tickets):
        st, rep = srv.result(t)
        steps_served.append(rep.steps)
        retries += rep.fp64_retries
        abandoned += rep.steps_abandoned
        drops.append(rep.resid[4] / rep.resid0[4] if rep.resid0[4] > 0 else float("nan"))
        ref = plain_states[sampled.index(k)] if (plain_states and k in sampled) else ()
        for name, a, b in zip(("rho", "rhou", "rhoE"), st, ref):
            d = float(np.max(np.abs(a - b))) / max(float(np.max(np.abs(b))), 1e-300)
            err[name] = d if not (d <= err[name]) else err[name]  # a NaN is kept, not hidden
        srv.release(t)
    print(f"# serving: served mixed {cap_served:.1f} solves/min (makespan {makespan:.2f} s)", file=sys.stderr,
          flush=True)
    # latency under Poisson arrivals at 0.8 x the served capacity
    rate = 0.8 * cap_served / 60.0
    rng = np.random.default_rng(seed + 1)
    arr = np.cumsum(rng.exponential(1.0 / rate, n))
    lat = []
    t0 = time.perf_counter()
    tk = []
    for (m, joe), a in zip(built, arr):
        delay = a - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        tk.append(srv.submit(m, joe, joemi.JM_MIXED))
    for t in tk:
        srv.wait(t)
        _, rep = srv.result(t)
        lat.append(rep.queued_s + rep.run_s)
        srv.release(t)
    srv.close()
    q = lambda v, p: float(np.percentile(np.array(v), p))
    res["served"] = {"solves_per_min": round(cap_served, 2), "precision": "FP64 residual, FP32 Jacobian/Krylov",
                     "mode": "batched: lockstep steps, each step and BiCGStab kernel launched once for all "
                             "active solves (2 groups)",
                     "max_active": max_active, "p50_s": round(q(lat, 50), 3), "p99_s": round(q(lat, 99), 3),
                     "mean_steps": round(float(np.mean(steps_served)), 1)}
    res["arrival_rate_per_min"] = round(rate * 60.0, 2)
    # np.max keeps a NaN (a solve that diverged), reported as null
    dmax = float(np.max(np.array(drops)))
    res["served_energy_resid_drop_max"] = float(f"{dmax:.3g}") if np.isfinite(dmax) else None
    res["served_nonfinite"] = int(np.sum(~np.isfinite(np.array(drops))))
    # mixed solves that went non-finite and were run again in FP64 (the
    # server's retry), and the steps their abandoned attempts took
    res["served_fp64_retries"] = int(retries)
    res["served_steps_abandoned"] = int(abandoned)
    if service:
        # the one-at-a-time loop meets the same arrivals, queued in front of its measured solve times
        done_prev, lat_plain = 0.0, []
        service_all = [service[k // plain_every] for k in range(n)]
        for a, sv in zip(arr, service_all):
            start = max(a, done_prev)
            done_prev = start + sv
            lat_plain.append(done_prev - a)
        res["plain_fp64_loop"] = {"solves_per_min": round(cap_plain, 2), "p50_s": round(q(lat_plain, 50), 3),
                                  "p99_s": round(q(lat_plain, 99), 3),
                                  "mean_steps": round(float(np.mean(steps_plain)), 1),
                                  "solves_timed": len(service)}
        res["speedup_solves_per_min"] = round(cap_served / cap_plain, 3)
        res["served_vs_fp64_max_rel_diff"] = {k: (float(f"{v:.3g}") if np.isfinite(v) else None)
                                              for k, v in err.items()}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=160)
    ap.add_argument("--schemes", default="HLLC,ROE,AUSM,HLLC_COUPLED")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--grad-kernel", default=None, help="GPU_GRAD_KERNEL: ENTRY or CELL")
    ap.add_argument("--flux-kernel", default=None, help="GPU_FLUX_KERNEL: ENTRY or CELL")
    ap.add_argument("--jac-kernel", default=None, help="GPU_JAC_KERNEL: ENTRY or CELL")
    ap.add_argument("--scal-kernel", default=None, help="GPU_SCAL_KERNEL (the SST plate's scalar rows): ENTRY or CELL")
    ap.add_argument("--grad-thread-order", default=None, help="GPU_GRAD_THREAD_ORDER: SORTED or NATIVE")
    ap.add_argument("--cell-order", default=None,
                    help="device cell numbering (GPU_CELL_ORDER): NATIVE, GROUPED, MORTON, MORTON_GROUPED")
    ap.add_argument("--serve", type=int, default=64, help="serving requests (0: skip the serving leg)")
    ap.add_argument("--serve-plain-every", type=int, default=4, help="the plain FP64 loop solves every k-th request")
    ap.add_argument("--serve-cells", default="3000,12000", help="min,max cells of a served request")
    ap.add_argument("--serve-active", type=int, default=16)
    ap.add_argument("--serve-steps", type=int, default=4000)
    ap.add_argument("--serve-cfl", type=float, default=20.0,
                    help="MAX_CFL of the served requests (at 50, 6 of 16 sampled requests diverged)")
    ap.add_argument("--serve-rel", type=float, default=1e-5)
    ap.add_argument("--serve-legs", default="plain,batched",
                    help="serving parts to run: plain (FP64 loop), streams (per-stream graph server), batched")
    ap.add_argument("--plate", type=int, default=1, help="time the 2M-cell flat-plate step (0: skip)")
    ap.add_argument("--chem", type=int, default=8_000_000, help="points of the chemistry-table lookup leg (0: skip)")
    ap.add_argument("--plate-dims", default="40,324,110,50", help="cells ahead of / along the plate, wall-normal, z")
    ap.add_argument("--plate-turb", default="SA,SST", help="turbulence models of the plate leg (SA coupled, SST "
                                                          "segregated)")
    ap.add_argument("--copy-calib-gib", type=float, default=0.0,
                    help="first run a plain device copy of this many GiB (calibrates profiler byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 (multi-GPU sharding is a "
                                   "design in DESIGN.md, not measured)"}))
        sys.exit(2)
    if args.copy_calib_gib > 0:
        import torch

        n = int(args.copy_calib_gib * (1 << 30)) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(f"# copy calibration: {n * 8} bytes read and written per copy, 3 copies", file=sys.stderr, flush=True)
        del src, dst

    from myjoe_amd import cases, joemi

    t_setup = time.perf_counter()
    seed = 0
    while True:
        mm, _ = cases.box_case(args.n, seed=seed)
        mesh = joemi.Mesh.from_faces(mm)
        del mm
        sdn, negtet = mesh.quality()
        if sdn > 0.0:
            break
        print(f"# seed {seed}: s.n min {sdn:.3g} <= 0, reseeding", file=sys.stderr, flush=True)
        seed += 1
    gc.collect()
    state = cases.random_state(mesh.ncv, seed=1)
    print(f"# mesh: ncv={mesh.ncv} nfa={mesh.nfa} nfa_b={mesh.nfa_b} nnz={mesh.nnz} seed={seed} "
          f"min s.n={sdn:.3f} setup {time.perf_counter() - t_setup:.1f}s", file=sys.stderr, flush=True)

    results = {}
    kernels = {}
    for scheme in args.schemes.split(","):
        if scheme == "HLLC_COUPLED":  # calcRhsCoupled's residual (runBackwardEulerCoupled)
            joe = cases.coupled(cases.box_joe_in("HLLC"), keep_symm_bc=True)
        else:
            joe = cases.box_joe_in(scheme)
        if args.cell_order:
            joe += f"GPU_CELL_ORDER = {args.cell_order}\n"
        if args.grad_kernel:
            joe += f"GPU_GRAD_KERNEL = {args.grad_kernel}\n"
        if args.flux_kernel:
            joe += f"GPU_FLUX_KERNEL = {args.flux_kernel}\n"
        if args.jac_kernel:
            joe += f"GPU_JAC_KERNEL = {args.jac_kernel}\n"
        if args.grad_thread_order:
            joe += f"GPU_GRAD_THREAD_ORDER = {args.grad_thread_order}\n"
        s = joemi.Solver(mesh, joe, joemi.JM_FP64)
        s.set_state(*state)
        for _ in range(args.warmup):
            s.calc_rhs(False)
        s.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            s.calc_rhs(False)
        s.sync()
        dt = time.perf_counter() - t0
        results[scheme] = {"ms_per_step": dt / args.steps * 1e3, "mcell_it_s": mesh.ncv * args.steps / dt / 1e6}
        kernels[scheme] = s.time_kernels(reps=10)
        if scheme == args.schemes.split(",")[0]:
            # the Jacobian assembly of an implicit step (FP64 blocks): the
            # residual with A, less the residual alone, wall clock over 5 calls
            s.calc_rhs(True)
            s.sync()
            t0 = time.perf_counter()
            for _ in range(5):
                s.calc_rhs(True)
            s.sync()
            jac_ms = (time.perf_counter() - t0) / 5 * 1e3 - results[scheme]["ms_per_step"]
        print(f"# {scheme}: {results[scheme]['ms_per_step']:.3f} ms/residual, "
              f"{results[scheme]['mcell_it_s']:.1f} Mcell-it/s; kernels "
              + " ".join(f"{k}={v:.3f}ms" for k, v in kernels[scheme].items()), file=sys.stderr, flush=True)
        del s
        gc.collect()

    head = args.schemes.split(",")[0]
    kt = kernels[head]
    ab = algorithmic_bytes(mesh.ncv, mesh.nfa, mesh.nfa_b, mesh.nnz)
    dom = max(("k_grad", "k_flux", "k_state"), key=lambda k: kt.get(k, 0.0))
    achieved = ab[dom] / (kt[dom] * 1e-3) / 1e9
    resid_bytes = ab["k_state"] + ab["k_bc"] + ab["k_grad"] + ab["k_flux"]
    traffic, traffic_src = measured_traffic(mesh, dom) if head == "HLLC" else (None, None)
    # the residual's compulsory bytes as the scope table counts them (state,
    # gradient, limiter, flux passes): 780 ncv + 32 nnz + 4 nfoc + 80 nfa_i + 160 nfa_b
    nfa_i = mesh.nfa - mesh.nfa_b
    survey_bytes = 780 * mesh.ncv + 32 * mesh.nnz + 4 * (2 * nfa_i + mesh.nfa_b) + 80 * nfa_i + 160 * mesh.nfa_b
    out = {
        "metric": METRIC,
        "value": round(results[head]["mcell_it_s"], 2),
        "unit": "Mcell-it/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(results[head]["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"single-residual throughput: {args.n}^3 box, 50% of hexes split into 6 Kuhn tets, "
                        f"interior nodes jittered 0.2h; {head} 2nd order, LS gradients + BJ-mod limiter",
            "ncv": mesh.ncv, "nfa": mesh.nfa, "nfa_b": mesh.nfa_b, "nnz": mesh.nnz,
            "per_scheme_mcell_it_s": {k: round(v["mcell_it_s"], 2) for k, v in results.items()},
            "kernel_ms": {k: round(v, 4) for k, v in kt.items()},
            "jacobian_fp64_ms": round(jac_ms, 3),
            "jacobian_algorithmic_GBs": round(ab["k_jac"] / (jac_ms * 1e-3) / 1e9, 1),
            "residual_algorithmic_GBs": round(resid_bytes / (kt["residual"] * 1e-3) / 1e9, 1),
            "residual_frac_of_hbm_peak_scope_table_bytes": round(
                survey_bytes / (results[head]["ms_per_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        },
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": traffic_src, "algorithmic_bytes_per_launch": int(ab[dom])},
    }
    del mesh
    gc.collect()
    if args.plate:
        dims = tuple(int(x) for x in args.plate_dims.split(","))
        for turb in args.plate_turb.split(","):
            key = "implicit_rans_flat_plate" + ("" if turb == "SA" else "_" + turb.lower())
            out["config"][key] = plate_bench(dims, jac_kernel=args.jac_kernel, turb=turb, scal_kernel=args.scal_kernel)
    chem_sample = None
    if args.chem > 0:
        out["config"]["fpva_chemtable_lookup"], chem_sample = chem_bench(args.chem)
    if args.serve > 0:
        lo, hi = (int(x) for x in args.serve_cells.split(","))
        out["serving"] = serving_bench(args.serve, 7, lo, hi, args.serve_active, args.serve_steps, args.serve_rel, 64.0,
                                       legs=tuple(args.serve_legs.split(",")), plain_every=args.serve_plain_every,
                                       cfl_max=args.serve_cfl)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
        if chem_sample is not None:
            out["config"]["fpva_chemtable_lookup"].update(chem_cpu_baseline(chem_sample))
    print(json.dumps(_strict_json(out), allow_nan=False), flush=True)


if __name__ == "__main__":
    main()
