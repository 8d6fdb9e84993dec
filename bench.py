"""Benchmark: transient PISO on the 2M-cell channel, mixed precision vs FP64.

Workload (one step = one PISO time step of piso_run_mpi at one partition):
  channel 400 x 100 x 50 hex cells (4*pi x 2 x 2*pi, wall-normal tanh
  clustering, adjacent-spacing ratio 1.1), Re_tau = 180 (mu = 1/180,
  body force 1), periodic in x and z, walls in y; initial field = laminar
  parabola of bulk velocity U_b = 15.63 + seeded uniform noise of +-5% U_b
  (seed 180); dt from CFL 0.5; BDF2; 2 PISO correctors; p' to a 1e-10
  relative FP64 residual, momentum by BiCGSTAB with multicolour DILU to
  1e-10.  --bc twin runs the non-periodic twin instead (parabolic inlet /
  outflow in x, symmetry in z), the case the CPU oracle times, since the
  reference has no periodic boundary.  Variants (all reach 1e-10):
  * mixed_schur (value): the mesh's cell graph has two colours; colour 0 is
    eliminated and PCG runs on the Schur complement of colour 1 with the
    Jacobi preconditioner (the iterates of multicolour DILU-PCG) and the
    aggregate coarse correction, FP32 inner iterations inside FP64 reliable
    updates; the FP32 iterations apply S matrix-free (two launches over the
    full matrix's rows; MIXSOLVER_SCHUR_MF=0 uses the formed S);
  * fp64_schur: the same solver in FP64 (the precision ratio's reference);
  * mixed / fp64_two_level: PCG with multicolour DILU and the coarse
    correction on the full system, mixed and FP64;
  * fp64: PCG with multicolour DILU, FP64, no coarse correction.
p50/p99 step latency and the iteration counts are reported per variant.
Inputs are resident in HBM before timing.  After the variants: the serving
workload (Poisson arrivals of p' solves, at 200 and 100 req/s) and the
parameter-sweep workload (32 porous LTNE cases of 16M cells, forked from one
resident system, per-step p50/p99 latency; --no-sweep skips it,
--sweep-active 32,4,1 compares in-flight limits).

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


def run_variant(Case, mesh, ctl, u0, over, steps, warmup, probe=False):
    """One PISO run of warmup + steps steps (the first step starts the run,
    every later one continues it), the last `steps` timed.  probe=True times the
    last FP32 SpMV launch of every 16-iteration chunk of the p' solves
    inside the timed region (start / stop events on the library stream; on
    from the warm-up on, so the chunk graphs are captured before the timed
    region) and returns (launches, ms, rows) as a fifth element."""
    from mixsolver_amd.solver import spmv_probe
    cs = Case.from_control(mesh, ctl, **over)
    cs.set_field("u", u0)
    if probe:
        spmv_probe(True)
    # one run of warmup + steps PISO steps: the first step starts the run
    # (first-order start, apc = 1), every later one continues it (BDF2)
    started = False

    def step():
        nonlocal started
        ier = cs.piso_continue(1) if started else cs.piso_run(1)
        started = True
        return ier

    for _ in range(warmup):
        if step() != 0:
            raise RuntimeError("warmup step failed")
    lat, hist = [], []
    cs.solver_stats(reset=True)       # counters of the timed steps only
    if probe:
        spmv_probe(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        t1 = time.perf_counter()
        ier = step()                  # returns after the device is synchronised
        lat.append(time.perf_counter() - t1)
        if ier != 0:
            raise RuntimeError(f"PISO step failed, ier={ier}")
        hist.append(cs.history()[-1])
    total = time.perf_counter() - t0
    if probe:
        return cs, total, np.array(lat), np.array(hist), spmv_probe(False)
    return cs, total, np.array(lat), np.array(hist)


def step_bytes(mesh, periodic, coarse, n1, nnz_s, ppe_its, mom_its, correctors=2, spmv_red=None, matfree=False):
    """Algorithmic HBM bytes of one PISO step on the red-black reduced p'
    path (mixed precision), summed over its launches with the step's logged
    iteration counts.  Each array is read once or written once per launch;
    int32 indices; FP64 assembly, FP32 Krylov work.  n1 / nnz_s: rows and
    nonzeros of the reduced matrix S; spmv_red: algorithmic bytes of the
    FP32 product with S in use (the matrix-free pair when matfree)."""
    n = mesh.ncells
    nper = int(np.count_nonzero(np.isin(mesh.zone, (6, 7, 8, 9)))) if periodic else 0
    nfi = mesh.n_interior + nper // 2                 # couplings after joining the periodic pairs
    nfb = mesh.nfaces - mesh.n_interior - nper
    nf = nfi + nfb
    nnz = n + 2 * nfi
    spmv32 = nnz * 8 + 4 * (n + 1) + 2 * n * 4
    spmv64 = nnz * 12 + 4 * (n + 1) + 3 * n * 8       # r = b - A x
    dilu32 = (nnz - n) * 8 + 4 * n * 4
    bicg_it = 2 * spmv32 + 2 * dilu32 + 20 * n * 4
    mom_asm = nfi * 56 + nfb * 72 + n * 128 + nnz * 8 + 2 * n * 8
    ppe_asm = nfi * 40 + nfb * 48 + 24 * n + nnz * 8 + 8 * n
    grad = nf * 40 + 200 * n
    rhie = nfi * 48 + 96 * n + nf * 8
    corr = nf * 40 + 64 * n
    spmv_s = spmv_red if spmv_red else nnz_s * 8 + 4 * (n1 + 1) + 2 * n1 * 4
    red_it = spmv_s + 44 * n1                         # product with S, update (8 arrays), p update (3)
    schur_form = nnz * 12 + nnz_s * 8 + nnz_s * 12    # S from A, then its FP32 copy
    if matfree:
        schur_form += nnz * 12 + nnz * 4              # the matrix-free weights from A
    coarse_setup = 16 * coarse ** 3 + 8 * coarse ** 2  # Gauss-Jordan: k launches over the k x k matrix
    per_corr = ppe_asm + rhie + corr + grad + 3 * spmv64   # residual, reduced rhs, back substitution
    parts = {
        "momentum_assembly": 3 * mom_asm, "momentum_solves": mom_its * bicg_it,
        "p_iterations": ppe_its * red_it, "correctors": correctors * per_corr,
        "schur_and_coarse_setup": schur_form + coarse_setup, "gradients": grad,
    }
    return float(sum(parts.values())), {k: float(v) for k, v in parts.items()}


def cpu_leg(cap):
    """One PISO step of the oracle (the C restatement of the reference path)
    on the 2M-cell non-periodic twin channel, p' PCG-Jacobi capped at `cap`
    iterations per solve (0 = the control's 20000, a full step).  Prints
    JSON: total and p' seconds, p' iterations.  Run in a child process so the
    serial or the OpenMP oracle library (ORACLE_THREADS) can be chosen."""
    from mixsolver_amd import meshgen
    from mixsolver_amd.control import parse_control
    from oracle import pyoracle
    nx, ny, nz = 400, 100, 50
    mesh = meshgen.channel(nx, ny, nz)
    u0 = initial_field(mesh, ny, nz, nx)
    ctrl, groups = parse_control(control_text(nx, lin_max=cap if cap > 0 else 20000, periodic=False), warn=False)
    ctrl.ppe_precond = 0
    oc = pyoracle.OracleCase(mesh, ctrl, groups)
    oc.set_field("u", u0)
    oc.piso_run(1, mpi=True)
    t_total, t_ppe, iters = oc.timing()
    print(json.dumps({"t_total": t_total, "t_ppe": t_ppe, "iters": iters}))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n_needed, full=False):
    """The oracle on this host's cores, at 1 thread and on all the
    cores this process may use, on one PISO step of the 2M-cell
    non-periodic twin channel.  Bounded sample (default): the p' PCG-Jacobi
    solves capped at 200 iterations each; the step is scaled with the
    measured per-iteration p' cost to the iteration count the FP64 Jacobi
    device solve needed for the same step.  full=True: the whole step, no
    cap, no scaling."""
    import subprocess
    ncores = len(os.sched_getaffinity(0))
    cap = 0 if full else 200
    legs = {}
    for threads in (1, ncores):
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), ORACLE_THREADS="1" if threads > 1 else "0",
                   ORACLE_PROGRESS="1" if full else "0")
        # the leg's stderr (progress of a full step) goes straight through
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-leg", str(cap)], env=env,
                           stdout=subprocess.PIPE, text=True, check=True)
        leg = json.loads(r.stdout.strip().splitlines()[-1])
        per_it = leg["t_ppe"] / max(leg["iters"], 1)
        t_step = leg["t_total"] if full else (leg["t_total"] - leg["t_ppe"]) + per_it * n_needed
        legs[threads] = (t_step, leg, per_it)
        print(f"bench: cpu leg on {threads} thread(s): {t_step:.2f} s/step", file=sys.stderr, flush=True)
    t1, leg1, pit1 = legs[1]
    tn, legn, pitn = legs[ncores]
    how = ("one full step, measured" if full else
           f"p' PCG-Jacobi capped at 200 iterations per solve ({leg1['iters']} p' iterations at "
           f"{pit1 * 1e3:.1f} ms on 1 thread, {pitn * 1e3:.1f} ms on {ncores}), the step scaled to the "
           f"{n_needed} p' iterations of the FP64 Jacobi device solve")
    return dict(value=1.0 / tn, unit="steps/s", cores=ncores, kind="port",
                one_thread_steps_per_s=1.0 / t1, cpu_model=cpu_model(),
                sample=(f"oracle (C restatement of the reference path) PISO step on the 2M-cell non-periodic twin "
                        f"channel (parabolic inlet, outflow, z symmetry): {how}; {ncores} cores "
                        f"(OpenMP in the Krylov vector work), 1-thread figure beside it"))


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


SWEEP_CTL = """rho = 1.0
mu = 0.05
transient = true
dt = 0.002
n_time_max = 1
n_correct = 2
time_scheme = 2
lin_tol = 1e-10
lin_max = 20000
conv_blend = 0.5
nonorth_corr = 1.0
ppe_precond = jacobi
k_cond = 0.1
cp = 1.0
thermal_model = ltne
bc = 6 velocity-inlet-parabolic 0.05 2.0 2 -1.0 1.0
bc = 7 outflow
bc = 5 wall
bc = 4 symmetry
tbc = 5 1 0.5
cell_zone = 8 porous perm=0.01 perm_yy=0.02 inertial=0.5 porosity=0.4 disp_l=0.01 disp_t=0.002 bj_alpha=0.3 k_s=2.0 cp_s=0.8 rho_s=3.0 h_sf=5.0 a_sf=10.0
"""


def sweep_bench(size, ncases, nsteps, max_active, coarse, seed=7):
    """Parameter-sweep workload: a porous channel (8 x 2 x 1, nx x ny x nz
    hex cells, one porous zone: Darcy-Forchheimer source, anisotropic
    permeability, dispersion, LTNE fluid and solid energy equations coupled
    by h_sf a_sf, wall at fixed temperature) with a seeded porosity field in
    [0.4, 0.9] (Gaussian noise smoothed over 10 cells); ncases cases with
    seeded inlet mean speeds uniform in [0.01, 0.1] m/s (parabolic inlet),
    starting from rest, nsteps PISO steps each (dt 0.002, 2 correctors,
    BDF2), p' through the red-black reduced system in mixed precision to
    1e-10; all cases submitted at once and admitted by device-memory budget
    (90% of the free HBM after the base system)."""
    from scipy.ndimage import gaussian_filter
    from mixsolver_amd import meshgen, zones
    from mixsolver_amd.control import parse_control
    from mixsolver_amd.serve import Sweep
    from mixsolver_amd.solver import Case
    nx, ny, nz = (int(v) for v in size.split("x"))
    t0 = time.perf_counter()
    m = meshgen.hex_box(np.linspace(0.0, 8.0, nx + 1), np.linspace(-1.0, 1.0, ny + 1), np.linspace(0.0, 1.0, nz + 1),
                        zone_ids=dict(xmin=6, xmax=7, ymin=5, ymax=5, zmin=4, zmax=4))
    m = meshgen.with_cell_zones(m, np.full(m.ncells, 8), {8: ("porous", "bed")})
    ctrl, _ = parse_control(SWEEP_CTL, warn=False)
    cz = zones.resolve_cell_zones(m, ctrl.cell_zones)
    f = zones.setup_porous_fields(m, ctrl.cell_zones, cz)
    g = gaussian_filter(np.random.default_rng(seed).standard_normal((nz, ny, nx)), 10.0, mode="wrap").ravel()
    f["porosity"] = 0.4 + 0.5 * (g - g.min()) / (g.max() - g.min())
    c2, g2 = parse_control(SWEEP_CTL, warn=False)
    c2.cell_zones = []
    for k, v in dict(gpu_ppe_precond=0, gpu_ppe_solver=2, gpu_mom_precond=2, gpu_coarse=coarse, gpu_precision=1,
                     gpu_inner_tol=0.1).items():
        setattr(c2, k, v)
    base = Case(m, c2, g2)
    base.set_cell_types(cz)
    for k, v in f.items():
        base.set_field(k, v)
    t_setup = time.perf_counter() - t0
    print(f"bench: sweep base {m.ncells} cells set up in {t_setup:.1f} s", file=sys.stderr, flush=True)
    speeds = np.random.default_rng(seed + 1).uniform(0.01, 0.1, ncases)
    runs = []
    for act in max_active:
        sw = Sweep(base, budget_bytes=0, max_active=act)
        t1 = time.perf_counter()
        ids = [sw.submit(nsteps, zone=6, speed=float(v)) for v in speeds]
        res = []
        for i in ids:
            res.append(sw.result(i, wait_s=3600.0))
            print(f"bench: sweep (at most {act} in flight) case {len(res)}/{ncases} done at "
                  f"{time.perf_counter() - t1:.1f} s", file=sys.stderr, flush=True)
        span = time.perf_counter() - t1
        st = sw.stats()
        sw.close()
        steps = np.concatenate([r["step_ms"] for r in res])
        later = np.concatenate([r["step_ms"][1:] for r in res]) if nsteps > 1 else steps
        first = min(res, key=lambda r: r["t_admit"])
        bad = [dict(speed=float(v), ier=r["ier"], steps=r["steps"], ppe_res=r["history"][:, 7].tolist(),
                    ppe_its=r["history"][:, 6].tolist())
               for v, r in zip(speeds, res)
               if not (r["ier"] == 0 and r["steps"] == nsteps and (r["history"][:, 7] <= 1e-10).all())]
        nsteps_all = int(sum(r["steps"] for r in res))
        lat = [r["t_done"] - r["t_submit"] for r in res]
        runs.append({
            "max_in_flight": act, "peak_cases_in_flight": st["peak_active"],
            "converged_cases": ncases - len(bad), "unconverged": bad,
            "ppe_fp64_solves": int(sum(r["stats"]["ppe_fp64_solves"] for r in res)),
            "fp32_stalled_cases": int(sum(r["stats"]["fp32_stalled"] for r in res)),
            "ppe_iters_per_step": float(np.mean(np.concatenate([r["history"][:, 6] for r in res]))),
            "p50_step_ms": float(np.percentile(steps, 50)), "p99_step_ms": float(np.percentile(steps, 99)),
            "p50_later_step_ms": float(np.percentile(later, 50)),
            "p99_later_step_ms": float(np.percentile(later, 99)),
            "first_step_alone_ms": float(first["step_ms"][0]),
            "makespan_s": span, "case_steps_per_s": nsteps_all / span,
            "cell_steps_per_s": nsteps_all * m.ncells / span,
            "p50_case_latency_s": float(np.percentile(lat, 50)), "p99_case_latency_s": float(np.percentile(lat, 99)),
            "footprint_GB_per_case": st["footprint_bytes"] / 1e9, "budget_GB": st["budget_bytes"] / 1e9,
        })
    nbytes = base.info["device_bytes"]
    base.close()
    return {
        "workload": f"porous LTNE channel {size} hex ({m.ncells} cells), seeded porosity [0.4, 0.9] (smoothed, "
                    f"correlation 10 cells), Darcy-Forchheimer, fluid + solid energy (h_sf a_sf); {ncases} cases, "
                    f"seeded inlet mean speeds U(0.01, 0.1) m/s, {nsteps} PISO steps each from rest; p' reduced "
                    f"system + {coarse}-aggregate coarse, FP32 inner / FP64 to 1e-10; all submitted at t = 0, "
                    "admitted by HBM budget (90% of the free HBM)",
        "cells": int(m.ncells), "cases": ncases, "steps_per_case": nsteps, "base_setup_s": t_setup,
        "base_GB": nbytes / 1e9, "runs": runs,
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
    ap.add_argument("--coarse", type=int, default=256, help="aggregates of the p' coarse space")
    ap.add_argument("--variants", default="mixed_schur,fp64_schur,mixed,fp64_two_level,fp64",
                    help="comma list of variants to run (profiling runs may select one)")
    ap.add_argument("--serve-seconds", type=float, default=3.0, help="serving run length (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-pool", type=int, default=10)
    ap.add_argument("--probe", choices=("timed", "extra", "off"), default="timed",
                    help="in-loop SpMV timing of the lead variant: inside its timed region, over as many extra "
                         "steps after it, or none")
    ap.add_argument("--cpu-full", action="store_true", help="CPU baseline on one full step (minutes), no scaling")
    ap.add_argument("--no-sweep", action="store_true", help="skip the parameter-sweep workload")
    ap.add_argument("--sweep-size", default="800x200x100")
    ap.add_argument("--sweep-cases", type=int, default=32)
    ap.add_argument("--sweep-steps", type=int, default=2)
    ap.add_argument("--sweep-active", default="4",
                    help="most cases in flight, one sweep per value (the HBM budget also bounds them)")
    ap.add_argument("--cpu-leg", type=int, default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_leg is not None:
        cpu_leg(args.cpu_leg)
        return
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
    lead_name = next((k for k in ("mixed_schur", "mixed", "fp64_schur", "fp64_two_level", "fp64") if k in variants), None)
    runs, fields = {}, {}
    schur = dict(two, gpu_ppe_precond=0, gpu_ppe_solver=2)
    for name, over in (("mixed_schur", dict(schur, gpu_precision=1, gpu_inner_tol=0.1)),
                       ("fp64_schur", dict(schur, gpu_precision=0)),
                       ("mixed", dict(two, gpu_precision=1, gpu_inner_tol=0.1)),
                       ("fp64_two_level", dict(two, gpu_precision=0)),
                       ("fp64", dict(base, gpu_precision=0))):
        if name not in variants:
            continue
        print(f"bench: {name}: set-up and {args.warmup} + {args.steps} steps", file=sys.stderr, flush=True)
        probe_here = name == lead_name and name.startswith("mixed") and args.probe == "timed"
        res = run_variant(Case, mesh, ctl, u0, over, args.steps, args.warmup, probe=probe_here)
        cs, t, lat, h = res[:4]
        if probe_here:
            runs["inloop"] = res[4] + ("timed region",)
        runs[name + "_stats"] = cs.solver_stats()
        if name in ("mixed", "fp64_two_level", "mixed_schur", "fp64_schur"):
            # fields after the same steps: the mixed result's error against FP64
            fields[name] = (cs.get("u"), cs.get("p"))
        if name == "mixed_schur":
            runs["kernel_schur"] = (cs.time_kernel(fp32=True, reps=100, reduced=True)
                                    + cs.time_kernel(fp32=False, reps=50, reduced=True))
            runs["kernel_schur_cold"] = (cs.time_kernel(fp32=True, reps=20, cold=True, reduced=True)[0],
                                         cs.time_kernel(fp32=False, reps=20, cold=True, reduced=True)[0])
        if name == "mixed":
            # roofline of the dominant kernel (FP32 SpMV of the mixed p' solve) on the last p' matrix
            runs["kernel"] = cs.time_kernel(fp32=True, reps=100) + cs.time_kernel(fp32=False, reps=50)
            # the same kernels with the last-level cache evicted before each launch
            runs["kernel_cold"] = (cs.time_kernel(fp32=True, reps=20, cold=True)[0],
                                   cs.time_kernel(fp32=False, reps=20, cold=True)[0])
        if name == lead_name and name.startswith("mixed") and args.probe == "extra":
            from mixsolver_amd.solver import spmv_probe
            spmv_probe(True)
            tx = time.perf_counter()
            for _ in range(args.steps):
                if cs.piso_continue(1) != 0:
                    raise RuntimeError("PISO step failed")
            tx = time.perf_counter() - tx
            runs["inloop"] = spmv_probe(False) + (f"{args.steps} steps after the timed region "
                                                  f"({args.steps / tx:.3f} steps/s with the probe)",)
        cs.close()
        runs[name] = (args.steps / t, t, lat, h)
        print(f"bench: {name} {args.steps / t:.3f} steps/s", file=sys.stderr, flush=True)
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

    lead = next(k for k in ("mixed_schur", "mixed", "fp64_schur", "fp64_two_level", "fp64") if k in runs)
    ref64 = "fp64_schur" if lead == "mixed_schur" else "fp64_two_level"

    def field_err():
        if lead not in fields or ref64 not in fields:
            return None
        (um, pm), (uf, pf) = fields[lead], fields[ref64]
        pm, pf = pm - pm.mean(), pf - pf.mean()     # p is fixed up to a constant (periodic channel)
        return {"u": float(np.linalg.norm(um - uf) / np.linalg.norm(uf)),
                "p": float(np.linalg.norm(pm - pf) / np.linalg.norm(pf))}
    solver_desc = {
        "mixed_schur": "p' red-black reduced PCG (colour 0 eliminated, Jacobi on the Schur complement = "
                       "multicolour DILU iterates) + {k}-aggregate coarse correction, FP32 inner (S applied "
                       "matrix-free) / FP64 reliable updates to 1e-10",
        "fp64_schur": "p' red-black reduced PCG + {k}-aggregate coarse correction, FP64, to 1e-10",
        "mixed": "p' PCG + multicolour DILU + {k}-aggregate coarse correction, FP32 inner / FP64 reliable updates "
                 "to 1e-10",
        "fp64_two_level": "p' PCG + multicolour DILU + {k}-aggregate coarse correction, FP64, to 1e-10",
        "fp64": "p' PCG + multicolour DILU, FP64, to 1e-10",
    }
    lead_stats = runs.get(lead + "_stats")
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
        "dtype": "f32 inner + f64 refinement" if lead.startswith("mixed") else "f64",
        "data": "synthetic",
        "config": {"workload": f"channel {nx}x{ny}x{nz} hex ({cells} cells), "
                               + ("periodic x/z" if periodic else "non-periodic twin (inlet/outflow, z symmetry)")
                               + ", transient PISO, 2 correctors, " + solver_desc[lead].format(k=args.coarse)
                               + ", momentum BiCGSTAB + DILU to 1e-10",
                   "cells": cells, "faces": mesh.nfaces, "time_scheme": "BDF2", "parallelism": "1 GPU",
                   "variant": lead},
        # precision gain at equal algorithm: the lead (mixed) against the same solver in FP64
        "fp64_same_solver": ref64,
        "speedup_precision": ratio(lead, ref64),
        # rel-L2 of the lead's fields against the same solver in FP64 after the same steps
        "mixed_vs_fp64_rel_l2": field_err(),
        # p' solves of the lead run that ran or finished in FP64 (stall guard)
        "mixed_fp64_fallback_solves": lead_stats["ppe_fp64_solves"] if lead_stats else None,
        "mixed_fp32_stalled": lead_stats["fp32_stalled"] if lead_stats else None,
        "variants": {k: {"steps_per_s": sps(k), "p50_step_ms": pct(k, 50), "p99_step_ms": pct(k, 99),
                         "ppe_iters_per_step": iters(k), "ppe_final_res_max": resmax(k), "solver": solver_desc[k]
                         .format(k=args.coarse) if k != "fp64" else solver_desc[k]}
                     for k in ("mixed_schur", "fp64_schur", "mixed", "fp64_two_level", "fp64") if k in runs},
        # the plain FP64 device path (one-level DILU) and the DILU two-level variants
        "fp64_steps_per_s": sps("fp64"),
        "speedup_vs_fp64_one_level": ratio(lead, "fp64"),
        "p50_step_ms": pct(lead, 50),
        "p99_step_ms": pct(lead, 99),
        "ppe_iters_per_step": iters(lead),
        "ppe_final_res_max": resmax(lead),
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
    matfree = os.environ.get("MIXSOLVER_SCHUR_MF", "1") != "0"
    kname = ("k_mf_stage1 + k_mf_stage2: the reduced matrix S applied matrix-free" if matfree else
             "k_pcg_spmv<float> on the reduced matrix S") if "kernel_schur" in runs else "k_pcg_spmv<float>"
    if os.path.exists(tpath):      # HBM bytes per launch measured by the counter runs (committed summary)
        with open(tpath) as fh:
            traffic = json.load(fh)["traffic_bytes"].get(kname)
    kkey = "kernel_schur" if "kernel_schur" in runs else ("kernel" if "kernel" in runs else None)
    if kkey:
        t_k, b_k, t_k64, b_k64 = runs[kkey]
        # the kernel's average launch inside the PISO steps (probe), else the
        # back-to-back launches on the last p' matrix
        t_avg, src = t_k, "100 back-to-back launches on the last p' matrix, HIP events"
        if "inloop" in runs and runs["inloop"][0] > 0:
            nl, ms_l, rows_l, where = runs["inloop"]
            t_avg = ms_l / nl
            src = (f"in-loop: {nl} launches of the p' solves ({rows_l} rows; the last of each 16-iteration "
                   f"chunk) over {where}, start/stop events of each launch (hipExtLaunchKernelGGL) on the "
                   "library stream")
        kdesc = (" (two launches over the full matrix's SELL-64 rows, difference form, fused p.q; one product)"
                 if kkey == "kernel_schur" and matfree else " (SELL-64, difference form, fused p.q)")
        out["roofline"] = {"bound": "hbm", "kernel": kname + kdesc,
                           "achieved": b_k / (t_avg * 1e-3) / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                           "frac": b_k / (t_avg * 1e-3) / HBM_PEAK, "traffic": traffic,
                           "traffic_source": "profiles/ (FETCH_SIZE, WRITE_SIZE, copy-calibrated)",
                           "avg_launch_us": t_avg * 1e3, "avg_launch_source": src,
                           "back_to_back_avg_launch_us": t_k * 1e3, "alg_bytes": b_k,
                           "fp64_kernel_GBps": b_k64 / (t_k64 * 1e-3) / 1e9, "fp64_avg_launch_us": t_k64 * 1e3,
                           "copy_1GiB_GBps": 2.0 * (1 << 30) / (cp_ms.value * 1e-3) / 1e9}
        # achievable HBM bandwidth on 4 GiB buffers (16x the last-level cache)
        hbm = {}
        for kind, name, moved in ((0, "copy", 2.0), (1, "copy_nt_stores", 2.0), (2, "read", 1.0)):
            ms = ctypes.c_double(0.0)
            _lib.check(_lib.lib().ms_time_stream(4 << 30, kind, 5, ctypes.byref(ms)))
            hbm[name + "_GBps"] = moved * (4 << 30) / (ms.value * 1e-3) / 1e9
        out["roofline"]["hbm_achievable_4GiB"] = hbm
        if kkey + "_cold" in runs:
            c32, c64 = runs[kkey + "_cold"]
            out["roofline"].update({"cold_avg_launch_us": c32 * 1e3, "cold_achieved": b_k / (c32 * 1e-3) / 1e9,
                                    "cold_frac": b_k / (c32 * 1e-3) / HBM_PEAK,
                                    "fp64_cold_avg_launch_us": c64 * 1e3,
                                    "fp64_cold_GBps": b_k64 / (c64 * 1e-3) / 1e9})
        if kkey == "kernel_schur" and "kernel" in runs:
            t_a, b_a = runs["kernel"][0], runs["kernel"][1]
            out["roofline"]["full_matrix_spmv"] = {"avg_launch_us": t_a * 1e3, "alg_bytes": b_a,
                                                   "frac": b_a / (t_a * 1e-3) / HBM_PEAK}
    if lead == "mixed_schur" and "kernel_schur" in runs and lead_stats:
        # whole-step roofline: algorithmic bytes of every launch of a step
        # (logged iteration counts) over the measured step time
        n1 = runs["inloop"][2] if "inloop" in runs and runs["inloop"][2] > 0 else cells // 2
        nnz_s = (runs["kernel_schur"][3] - 4.0 * (n1 + 1) - 16.0 * n1) / 12.0
        b_step, parts = step_bytes(mesh, periodic, args.coarse, n1, nnz_s, iters(lead),
                                   lead_stats["mom_iters"] / args.steps, spmv_red=runs["kernel_schur"][1],
                                   matfree=matfree)
        t_step = runs[lead][1] / args.steps
        out["roofline_step"] = {"bound": "hbm", "B_step": b_step, "t_step_ms": 1e3 * t_step,
                                "achieved": b_step / t_step / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                                "frac": b_step / t_step / HBM_PEAK, "parts": parts,
                                "ppe_iters_per_step": iters(lead),
                                "mom_iters_per_step": lead_stats["mom_iters"] / args.steps}
    if args.serve_seconds > 0:
        # the request server's coarse inverse runs in one workgroup (k <= 256)
        out["serving"] = serve_bench(args.serve_rate, args.serve_seconds, args.serve_pool, min(args.coarse, 256))
        # the same stream at half the offered rate, below saturation: the
        # steady-state latency (the offered-rate run above saturates one GPU,
        # so its latency grows with the run length; its throughput is the
        # saturation throughput)
        out["serving_below_saturation"] = serve_bench(args.serve_rate / 2, args.serve_seconds, args.serve_pool,
                                                      min(args.coarse, 256))
        print("bench: serving done", file=sys.stderr, flush=True)
    if not args.no_sweep and args.sweep_cases > 0:
        out["sweep"] = sweep_bench(args.sweep_size, args.sweep_cases, args.sweep_steps,
                                   [int(v) for v in args.sweep_active.split(",")], args.coarse)
        print("bench: sweep done", file=sys.stderr, flush=True)
    if not args.no_cpu:
        # the oracle runs the non-periodic twin (the reference has no periodic
        # boundary); p' iterations its Jacobi PCG needs for the first step come
        # from the device FP64 Jacobi run of the same twin
        twin = mesh if not periodic else meshgen.channel(nx, ny, nz)
        jac, _, _, h_jac = run_variant(Case, twin, control_text(nx, periodic=False), u0,
                                       dict(gpu_ppe_precond=0, gpu_mom_precond=2, gpu_precision=0), 1, 0)
        jac.close()
        out["cpu_baseline"] = cpu_baseline(int(h_jac[0, 6]), full=args.cpu_full)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
