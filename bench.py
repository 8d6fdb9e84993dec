"""Benchmark: transient PISO on the 2M-cell channel, mixed precision vs FP64.

Workload (one step = one PISO time step of piso_run_mpi at one partition):
  channel 400 x 100 x 50 hex cells (4*pi x 2 x 2*pi, wall-normal tanh
  clustering, adjacent-spacing ratio 1.1), Re_tau = 180 (mu = 1/180,
  body force 1), periodic in x and z, walls in y; initial field = laminar
  parabola of bulk velocity U_b = 15.63 + seeded uniform noise of +-5% U_b
  (seed 180); dt from CFL 0.5; BDF2; 2 PISO correctors; p' by PCG with
  multicolour DILU to a 1e-10 relative residual, momentum by BiCGSTAB with
  multicolour DILU to 1e-10.  --bc twin runs the non-periodic twin instead
  (parabolic inlet / outflow in x, symmetry in z), the case the CPU oracle
  times, since the reference has no periodic boundary.
  * mixed: FP32 matrix / preconditioner / Krylov vectors with FP32
    accumulation inside FP64 reliable updates (p') and FP64 iterative
    refinement (momentum); every solve ends at a 1e-10 FP64 true residual.
  * fp64: the same algorithms entirely in FP64.
value = mixed PISO steps/s; fp64 steps/s, the ratio and p50/p99 step
latency are reported beside it.  Inputs are resident in HBM before timing.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "PISO steps/s @2M cells, mixed vs FP64 at 1e-10 residual; p50/p99 solve latency"
HBM_PEAK = 8.0e12
UB = 15.63


def control_text(nx, lin_max=20000, periodic=True):
    dx = 4 * np.pi / nx
    dt = 0.5 * dx / (1.5 * UB * 1.05)
    bcs = ("bc = 5 wall\nperiodic = 7 6\nperiodic = 9 8\n" if periodic else
           f"bc = 6 velocity-inlet-parabolic {UB!r} 2.0 2 -1.0\nbc = 7 outflow\nbc = 5 wall\nbc = 4 symmetry\n")
    return f"""rho = 1.0
mu = {1.0 / 180.0!r}
transient = true
dt = {dt!r}
n_time_max = 1
n_correct = 2
time_scheme = 2
lin_tol = 1e-10
lin_max = {lin_max}
conv_blend = 0.0
nonorth_corr = 1.0
ppe_precond = jacobi
body_force = 1.0 0.0 0.0
{bcs}"""


def initial_field(mesh, ny, nz, nx):
    from mixsolver_amd.meshgen import tanh_stretch
    rng = np.random.default_rng(180)
    ys = tanh_stretch(ny, 1.1)
    yc = 0.5 * (ys[1:] + ys[:-1])
    Y = np.broadcast_to(yc[None, :, None], (nz, ny, nx)).ravel()
    u0 = np.zeros((mesh.ncells, 3))
    u0[:, 0] = 1.5 * UB * (1.0 - Y ** 2)
    u0 += rng.uniform(-0.05 * UB, 0.05 * UB, size=u0.shape)
    return u0


def run_variant(Case, mesh, ctl, u0, over, steps, warmup):
    cs = Case.from_control(mesh, ctl, **over)
    cs.set_field("u", u0)
    for _ in range(warmup):
        if cs.piso_run(1) != 0:
            raise RuntimeError("warmup step failed")
    lat, hist = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        t1 = time.perf_counter()
        ier = cs.piso_run(1)          # returns after the device is synchronised
        lat.append(time.perf_counter() - t1)
        if ier != 0:
            raise RuntimeError(f"PISO step failed, ier={ier}")
        hist.append(cs.history()[-1])
    total = time.perf_counter() - t0
    return cs, total, np.array(lat), np.array(hist)


def cpu_baseline(mesh, u0, n_needed):
    """Oracle (C port of the reference path, 1 thread) on a bounded sample:
    one PISO step of the 2M-cell non-periodic twin channel with the p'
    PCG-Jacobi solves capped at 40 iterations; the full step is extrapolated
    with the measured per-iteration p' cost to the iteration count the FP64
    Jacobi device solve needed for the same step."""
    from mixsolver_amd.control import parse_control
    from oracle import pyoracle
    ctrl, groups = parse_control(control_text(400, lin_max=40, periodic=False), warn=False)
    ctrl.ppe_precond = 0
    oc = pyoracle.OracleCase(mesh, ctrl, groups)
    oc.set_field("u", u0)
    oc.piso_run(1, mpi=True)
    t_total, t_ppe, iters = oc.timing()
    oc.close()
    per_it = t_ppe / max(iters, 1)
    t_step = (t_total - t_ppe) + per_it * n_needed
    return dict(value=1.0 / t_step, unit="steps/s", cores=1, kind="port",
                sample=(f"oracle PISO step on the 2M-cell non-periodic twin channel (parabolic inlet, outflow, "
                        f"z symmetry) with p' PCG-Jacobi capped at 40 "
                        f"iterations ({t_total:.1f} s, {iters} p' iterations at {per_it * 1e3:.1f} ms); "
                        f"extrapolated to the {n_needed} p' iterations of the FP64 Jacobi device solve"))


def serve_bench(rate, seconds, pool, coarse, seed=4):
    """Serving workload: Poisson arrivals of SPD pressure-correction solves on a
    pool of seeded jittered hex/prism meshes with log-uniform sizes in
    [1e4, 2e6] cells, log-normal face diffusivity (sigma 1), Dirichlet
    boundaries, zero-mean Gaussian RHS, FP32 inner / FP64 reliable updates to
    1e-10.  Each mesh carries two pre-drawn coefficient sets and right-hand
    sides; a request picks a mesh and a set at random.  Latency = completion
    - arrival (submission at the arrival time)."""
    from mixsolver_amd.serve import LaplacianPattern, Server, request_mesh
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(1e4), np.log(2e6), pool))
    kinds = np.where(rng.uniform(size=pool) < 0.5, "hex", "prism")
    pats, data = [], []
    for i in range(pool):
        P = LaplacianPattern(request_mesh(sizes[i], kinds[i], seed=100 + i))
        pats.append(P)
        r2 = np.random.default_rng(1000 + i)
        data.append([(P.values(r2), P.rhs(r2)) for _ in range(2)])
    srv = Server(precond="dilu", coarse=coarse, round_iters=16)
    pids = [srv.add_pattern(P.row_ptr, P.col_idx, P.xc) for P in pats]
    for i in range(pool):            # pattern uploads and first-use set-up, untimed
        srv.result(srv.submit(pids[i], *data[i][0]), timeout_ms=-1)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, int(rate * seconds * 2) + 10))
    arrivals = arrivals[arrivals < seconds]
    picks = rng.integers(0, pool, len(arrivals))
    sets = rng.integers(0, 2, len(arrivals))
    tickets = []
    t0 = time.perf_counter()
    for ta, pi, si in zip(arrivals, picks, sets):
        dt = ta - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        late = (time.perf_counter() - t0) - ta          # submission after the arrival time
        tickets.append((srv.submit(pids[pi], *data[pi][si], tol=1e-10, maxit=20000), pi, late))
    res = [(srv.result(t, timeout_ms=-1), pi, late) for t, pi, late in tickets]
    span = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    # latency = completion - arrival
    lat = np.array([late + r["t_done"] for r, _, late in res])
    wait = np.array([late + r["t_admit"] for r, _, late in res])
    cells_it = float(sum(pats[pi].n * r["it"] for r, pi, _ in res))
    ok = sum(1 for r, _, _ in res if r["ierr"] == 0 and r["res"] <= 1e-10)
    return {
        "workload": f"Poisson {rate:g} req/s for {seconds:g} s; {pool} seeded meshes, log-uniform 1e4..2e6 cells, "
                    "hex/prism, log-normal diffusivity, Dirichlet, DILU"
                    + (f" + {coarse}-aggregate coarse" if coarse else "") + ", FP32 inner / FP64 to 1e-10",
        "requests": len(res), "converged": ok, "offered_rate": rate,
        "throughput_req_per_s": len(res) / span, "makespan_s": span,
        "p50_latency_ms": float(np.percentile(lat, 50) * 1e3), "p99_latency_ms": float(np.percentile(lat, 99) * 1e3),
        "p50_queue_ms": float(np.percentile(wait, 50) * 1e3), "p99_queue_ms": float(np.percentile(wait, 99) * 1e3),
        "cell_iterations_per_s": cells_it / span, "mean_iterations": float(np.mean([r["it"] for r, _, _ in res])),
        "mesh_cells": [int(P.n) for P in pats], "rounds": st["rounds"], "launches": st["launches"],
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", default="400x100x50")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--bc", choices=("periodic", "twin"), default="periodic",
                    help="periodic x/z channel, or its non-periodic twin")
    ap.add_argument("--coarse", type=int, default=512, help="aggregates of the p' coarse space")
    ap.add_argument("--variants", default="mixed,fp64_two_level,fp64",
                    help="comma list of variants to run (profiling runs may select one)")
    ap.add_argument("--serve-seconds", type=float, default=3.0, help="serving run length (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-pool", type=int, default=10)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}): the device library keeps a single "
              "partition, and sharding over several GPUs is not implemented", file=sys.stderr)
        sys.exit(2)
    from mixsolver_amd import meshgen
    from mixsolver_amd.solver import Case
    nx, ny, nz = (int(v) for v in args.size.split("x"))
    periodic = args.bc == "periodic"
    mesh = meshgen.channel(nx, ny, nz, periodic=periodic)
    ctl = control_text(nx, periodic=periodic)
    u0 = initial_field(mesh, ny, nz, nx)
    base = dict(gpu_ppe_precond=2, gpu_mom_precond=2)
    two = dict(base, gpu_coarse=args.coarse)

    variants = set(args.variants.split(","))
    runs = {}
    for name, over in (("mixed", dict(two, gpu_precision=1, gpu_inner_tol=0.1)),
                       ("fp64_two_level", dict(two, gpu_precision=0)),
                       ("fp64", dict(base, gpu_precision=0))):
        if name not in variants:
            continue
        cs, t, lat, h = run_variant(Case, mesh, ctl, u0, over, args.steps, args.warmup)
        if name == "mixed":
            # roofline of the dominant kernel (FP32 SpMV of the mixed p' solve) on the last p' matrix
            runs["kernel"] = cs.time_kernel(fp32=True, reps=100) + cs.time_kernel(fp32=False, reps=50)
            # the same kernels with the last-level cache evicted before each launch
            runs["kernel_cold"] = (cs.time_kernel(fp32=True, reps=20, cold=True)[0],
                                   cs.time_kernel(fp32=False, reps=20, cold=True)[0])
        cs.close()
        runs[name] = (args.steps / t, t, lat, h)
    cells = mesh.ncells
    def pct(name, q):
        return float(np.percentile(runs[name][2], q) * 1e3) if name in runs else None

    def sps(name):
        return runs[name][0] if name in runs else None

    def ratio(a, b):
        return sps(a) / sps(b) if a in runs and b in runs else None

    def iters(name):
        return float(np.mean(runs[name][3][:, 6])) if name in runs else None

    def resmax(name):
        return float(np.max(runs[name][3][:, 7])) if name in runs else None

    lead = "mixed" if "mixed" in runs else sorted(runs.keys() - {"kernel"})[0]
    out = {
        "metric": METRIC,
        "value": sps(lead),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * runs[lead][1] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 inner + f64 refinement" if lead == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": f"channel {nx}x{ny}x{nz} hex ({cells} cells), "
                               + ("periodic x/z" if periodic else "non-periodic twin (inlet/outflow, z symmetry)")
                               + ", transient PISO, 2 correctors, "
                               f"p' PCG + multicolour DILU + {args.coarse}-aggregate coarse correction to "
                               "1e-10, momentum BiCGSTAB + DILU to 1e-10",
                   "cells": cells, "faces": mesh.nfaces, "time_scheme": "BDF2", "parallelism": "1 GPU",
                   "variant": lead},
        # the plain FP64 device path: reference recurrences, one-level DILU
        "fp64_steps_per_s": sps("fp64"),
        "speedup_mixed_vs_fp64": ratio("mixed", "fp64"),
        # the same two-level solver in FP64 (isolates the precision gain)
        "fp64_two_level_steps_per_s": sps("fp64_two_level"),
        "speedup_mixed_vs_fp64_two_level": ratio("mixed", "fp64_two_level"),
        "p50_step_ms": pct(lead, 50),
        "p99_step_ms": pct(lead, 99),
        "fp64_p50_step_ms": pct("fp64", 50),
        "fp64_p99_step_ms": pct("fp64", 99),
        "fp64_two_level_p50_step_ms": pct("fp64_two_level", 50),
        "fp64_two_level_p99_step_ms": pct("fp64_two_level", 99),
        "ppe_iters_per_step_mixed": iters("mixed"),
        "ppe_iters_per_step_fp64": iters("fp64"),
        "ppe_iters_per_step_fp64_two_level": iters("fp64_two_level"),
        "ppe_final_res_mixed_max": resmax("mixed"),
        "ppe_final_res_fp64_max": resmax("fp64"),
        "ppe_final_res_fp64_two_level_max": resmax("fp64_two_level"),
        "cell_iterations_per_s": cells * float(np.sum(runs[lead][3][:, 6])) / runs[lead][1],
    }
    # calibration kernel of the traffic counters (rocprofv3 --pmc runs of this
    # script read its FETCH_SIZE / WRITE_SIZE beside the SpMV's)
    from mixsolver_amd import _lib
    import ctypes
    cp_ms = ctypes.c_double(0.0)
    _lib.check(_lib.lib().ms_time_copy(1 << 30, 20, ctypes.byref(cp_ms)))
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "pcg_spmv_traffic.json")
    if os.path.exists(tpath):      # HBM bytes per launch measured by the counter runs (committed summary)
        with open(tpath) as fh:
            traffic = json.load(fh)["traffic_bytes"].get("k_pcg_spmv<float>")
    if "kernel" in runs:
        t_k, b_k, t_k64, b_k64 = runs["kernel"]
        out["roofline"] = {"bound": "hbm", "kernel": "k_pcg_spmv<float> (SELL-64, difference form, fused p.q)",
                           "achieved": b_k / (t_k * 1e-3) / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                           "frac": b_k / (t_k * 1e-3) / HBM_PEAK, "traffic": traffic,
                           "traffic_source": "profiles/pcg_mixed_channel2M.md (FETCH_SIZE, WRITE_SIZE, copy-calibrated)",
                           "avg_launch_us": t_k * 1e3, "alg_bytes": b_k,
                           "fp64_kernel_GBps": b_k64 / (t_k64 * 1e-3) / 1e9, "fp64_avg_launch_us": t_k64 * 1e3,
                           "copy_1GiB_GBps": 2.0 * (1 << 30) / (cp_ms.value * 1e-3) / 1e9}
        # achievable HBM bandwidth on 4 GiB buffers (16x the last-level cache)
        hbm = {}
        for kind, name, moved in ((0, "copy", 2.0), (1, "copy_nt_stores", 2.0), (2, "read", 1.0)):
            ms = ctypes.c_double(0.0)
            _lib.check(_lib.lib().ms_time_stream(4 << 30, kind, 5, ctypes.byref(ms)))
            hbm[name + "_GBps"] = moved * (4 << 30) / (ms.value * 1e-3) / 1e9
        out["roofline"]["hbm_achievable_4GiB"] = hbm
        if "kernel_cold" in runs:
            c32, c64 = runs["kernel_cold"]
            out["roofline"].update({"cold_avg_launch_us": c32 * 1e3, "cold_achieved": b_k / (c32 * 1e-3) / 1e9,
                                    "cold_frac": b_k / (c32 * 1e-3) / HBM_PEAK,
                                    "fp64_cold_avg_launch_us": c64 * 1e3,
                                    "fp64_cold_GBps": b_k64 / (c64 * 1e-3) / 1e9})
    if args.serve_seconds > 0:
        # the request server's coarse inverse runs in one workgroup (k <= 256)
        out["serving"] = serve_bench(args.serve_rate, args.serve_seconds, args.serve_pool, min(args.coarse, 256))
    if not args.no_cpu:
        # the oracle runs the non-periodic twin (the reference has no periodic
        # boundary); p' iterations its Jacobi PCG needs for the first step come
        # from the device FP64 Jacobi run of the same twin
        twin = mesh if not periodic else meshgen.channel(nx, ny, nz)
        jac, _, _, h_jac = run_variant(Case, twin, control_text(nx, periodic=False), u0,
                                       dict(gpu_ppe_precond=0, gpu_mom_precond=2, gpu_precision=0), 1, 0)
        jac.close()
        out["cpu_baseline"] = cpu_baseline(twin, u0, int(h_jac[0, 6]))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
