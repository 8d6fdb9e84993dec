"""Write golden vectors from the reference's own FP64 Fortran path.

Runs oracle/_ref/ref_harness (built from /root/reference by
oracle/build_ref.sh) on small generated meshes and stores inputs and outputs
as compressed .npz fixtures next to this script.  Only needs to run where the
reference sources exist; the fixtures are committed data.

  python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from mixsolver_amd import cas, meshgen  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")

CAVITY_CTL = """rho = 1.0
mu = 0.001
alpha_u = 0.7
alpha_p = 0.3
outer_max = 1000
outer_tol = 1e-8
lin_tol = 1e-8
lin_max = 1000
conv_blend = {blend}
nonorth_corr = {nonorth}
ppe_precond = {precond}
bc = 5 wall
bc = 4 symmetry
lid = 5 2 1.0 1.0 0.0 0.0
"""

CHANNEL_CTL = """rho = 1.0
mu = {mu}
transient = true
dt = {dt}
n_time_max = {nsteps}
n_correct = 2
n_out_every = 1000000
time_scheme = 2
lin_tol = 1e-10
lin_max = 2000
conv_blend = {blend}
nonorth_corr = 1.0
ppe_precond = {precond}
body_force = 1.0 0.0 0.0
bc = 6 velocity-inlet-parabolic 1.0 2.0 2 -1.0 1.0
bc = 7 outflow
bc = 5 wall
bc = 4 symmetry
tbc = 5 1 0.5
k_cond = 0.01
cp = 1.0
"""

PIMPLE_CTL = """rho = 1.0
mu = 0.02
transient = true
pimple = true
n_outer_iter = 3
alpha_u = 0.7
dt = 0.08
n_time_max = 3
n_correct = 1
n_out_every = 1000000
time_scheme = 2
lin_tol = 1e-10
lin_max = 2000
conv_blend = 0.5
nonorth_corr = 1.0
ppe_precond = ic0
body_force = 1.0 0.0 0.0
bc = 6 velocity-inlet-parabolic 1.0 2.0 2 -1.0 1.0
bc = 7 outflow
bc = 5 wall
bc = 4 symmetry
tbc = 5 1 0.5
k_cond = 0.01
cp = 1.0
"""

PRISM_CTL = """rho = 1.0
mu = 0.02
transient = true
dt = 0.05
n_time_max = 2
n_correct = 2
n_out_every = 1000000
time_scheme = 2
lin_tol = 1e-10
lin_max = 2000
conv_blend = 0.5
nonorth_corr = 1.0
ppe_precond = ic0
bc = 6 pressure-far-field 0.0 1.0 0.0 0.0
bc = 5 wall
bc = 4 symmetry
tbc = 5 1 1.0
k_cond = 0.05
cp = 1.0
"""

OUTLET_CTL = """rho = 1.2
mu = 0.01
alpha_u = 0.6
alpha_p = 0.4
outer_max = 50
outer_tol = 1e-9
lin_tol = 1e-9
lin_max = 500
conv_blend = 1.0
nonorth_corr = 1.0
ppe_precond = jacobi
bc = 6 velocity-inlet 1.0 0.0 0.0 1.0
bc = 7 pressure-outlet 0.0
bc = 5 wall
bc = 4 slip-wall
boussinesq = true
beta = 0.01
gravity = 0.0 -9.81 0.0
tref = 0.5
k_cond = 0.1
cp = 1.0
tbc = 5 2 3.0
"""


def mesh_arrays(m):
    out = dict(x=m.x, fptr=m.fptr, fnodes=m.fnodes, c0=m.c0, c1=m.c1, zone=m.zone,
               ncells=np.int64(m.ncells), ctype=np.int64(m.ctype))
    if m.czone is not None:
        ids = sorted(m.czone_names)
        out.update(czone=m.czone, czone_ids=np.array(ids, np.int64),
                   czone_cond=np.array([m.czone_names[i][0] for i in ids]),
                   czone_name=np.array([m.czone_names[i][1] for i in ids]))
    return out


def run(args, cwd):
    r = subprocess.run([HARNESS] + args, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(r.stdout[-2000:] + r.stderr[-2000:])
    return r.stdout


class Reader:
    def __init__(self, path):
        self.b = open(path, "rb").read()
        self.o = 0

    def take(self, dt, n):
        a = np.frombuffer(self.b, dtype=dt, count=n, offset=self.o).copy()
        self.o += a.nbytes
        return a


def channel_box(nx, ny, nz):
    return meshgen.channel(nx, ny, nz, lx=4.0, lz=1.0, ratio=1.1)


def seeded_fields(m, seed, umag=1.0):
    rng = np.random.default_rng(seed)
    nc = m.ncells
    u = rng.uniform(-umag, umag, size=(nc, 3))
    p = rng.normal(size=nc)
    T = rng.uniform(0.0, 1.0, size=nc)
    return u, p, T


def write_fields(path, *arrs):
    with open(path, "wb") as f:
        for a in arrs:
            np.ascontiguousarray(a).tofile(f)


def driver_log(text):
    """The diagnostic lines the reference's drivers print (SIMPLE '  it=',
    PISO / PIMPLE '  step=' and the '    Tmin=' line after each step)."""
    keep = ("  it=", "  step=", "    Tmin=")
    return "\n".join(ln for ln in text.splitlines() if ln.startswith(keep))


def case_run(name, m, zone_kind, ctl, mode, nsteps, fields=None, want_ts=False):
    with tempfile.TemporaryDirectory() as td:
        cas.write_cas(os.path.join(td, "m.cas"), m, zone_kind)
        open(os.path.join(td, "c.control"), "w").write(ctl)
        inf = ""
        if fields is not None:
            inf = os.path.join(td, "in.bin")
            write_fields(inf, *fields)
        log = run(["m.cas", "c.control", mode, "out.bin", inf, str(nsteps)], td)
        r = Reader(os.path.join(td, "out.bin"))
        nc, nf = m.ncells, m.nfaces
        ier = r.take(np.int32, 1)[0]
        out = dict(ier=np.int64(ier), u=r.take(np.float64, 3 * nc).reshape(nc, 3), p=r.take(np.float64, nc),
                   flux=r.take(np.float64, nf), T=r.take(np.float64, nc),
                   gp=r.take(np.float64, 3 * nc).reshape(nc, 3), gu=r.take(np.float64, 9 * nc).reshape(nc, 9),
                   apc=r.take(np.float64, nc), gt=r.take(np.float64, 3 * nc).reshape(nc, 3))
        if want_ts:
            out["T_s"] = r.take(np.float64, nc)
        # the reference's printed per-iteration / per-step diagnostics
        out["log"] = np.array(driver_log(log))
        return out, log


def restart_fixture():
    """Serial piso_run from a restored BDF2 state (ts_order = 2 with u_old /
    u_old_old and T_old / T_old_old given): mod_uns_simple.f90:408-414 keep
    the first step second order."""
    mc = channel_box(12, 8, 4)
    zkc = {4: "symmetry", 5: "wall", 6: "velocity-inlet", 7: "outflow"}
    ctl = CHANNEL_CTL.format(mu=0.05, dt=0.02, nsteps=2, blend=0.5, precond="ic0")
    u0, p0, T0 = seeded_fields(mc, 181, umag=0.3)
    u0[:, 0] += 1.0
    rng = np.random.default_rng(182)
    u1 = u0 + rng.uniform(-0.02, 0.02, u0.shape)
    u2 = u1 + rng.uniform(-0.02, 0.02, u0.shape)
    T1 = T0 + rng.uniform(-0.05, 0.05, T0.shape)
    T2 = T1 + rng.uniform(-0.05, 0.05, T0.shape)
    out, log = case_run("chanrs", mc, zkc, ctl, "pisor", 2,
                        (u0, p0, T0, u1, u2, T1, T2, np.array([2], np.int32)), want_ts=True)
    save("channel_piso_restart", control=ctl, nsteps=np.int64(2), **mesh_arrays(mc),
         in_u=u0, in_p=p0, in_T=T0, in_u_old=u1, in_u_old_old=u2, in_T_old=T1, in_T_old_old=T2,
         in_ts_order=np.int64(2), **{"out_" + k: v for k, v in out.items()})


def geom_run(m, zone_kind, ctl):
    with tempfile.TemporaryDirectory() as td:
        cas.write_cas(os.path.join(td, "m.cas"), m, zone_kind)
        open(os.path.join(td, "c.control"), "w").write(ctl)
        run(["m.cas", "c.control", "geom", "out.bin"], td)
        r = Reader(os.path.join(td, "out.bin"))
        nc, nf, ncf = r.take(np.int32, 3)
        g = dict(xf=r.take(np.float64, 3 * nf).reshape(nf, 3), sf=r.take(np.float64, 3 * nf).reshape(nf, 3),
                 area=r.take(np.float64, nf), vol=r.take(np.float64, nc), xc=r.take(np.float64, 3 * nc).reshape(nc, 3),
                 lf=r.take(np.float64, nf))
        g["cf_ptr"] = r.take(np.int32, nc + 1)
        g["cf"] = r.take(np.int32, ncf)
        g["c2c_ptr"] = r.take(np.int32, nc + 1)
        g["c2c"] = r.take(np.int32, int(g["c2c_ptr"][-1]))
        nnz = r.take(np.int32, 1)[0]
        g["row_ptr"] = r.take(np.int32, nc + 1)
        g["col_idx"] = r.take(np.int32, nnz)
        g["fgrp"] = r.take(np.int32, nf)
        return g


def funcs_run(m, zone_kind, ctl, seed, ts_order):
    rng = np.random.default_rng(seed)
    nc, nf = m.ncells, m.nfaces
    u, p, T = seeded_fields(m, seed)
    flux = rng.normal(size=nf) * 0.1
    u_old = u + 0.05 * rng.normal(size=u.shape)
    u_oo = u + 0.05 * rng.normal(size=u.shape)
    T_old = T + 0.05 * rng.normal(size=nc)
    T_oo = T + 0.05 * rng.normal(size=nc)
    apc = rng.uniform(1.0, 3.0, size=nc)
    pp = rng.normal(size=nc) * 0.01
    inputs = dict(u=u, p=p, T=T, flux=flux, u_old=u_old, u_old_old=u_oo, T_old=T_old, T_old_old=T_oo,
                  apc=apc, pp=pp, ts_order=np.int64(ts_order))
    with tempfile.TemporaryDirectory() as td:
        cas.write_cas(os.path.join(td, "m.cas"), m, zone_kind)
        open(os.path.join(td, "c.control"), "w").write(ctl)
        inf = os.path.join(td, "in.bin")
        write_fields(inf, u, p, T, flux, u_old, u_oo, T_old, T_oo, apc, pp, np.array([ts_order], np.int32))
        run(["m.cas", "c.control", "funcs", "out.bin", inf], td)
        r = Reader(os.path.join(td, "out.bin"))
        g = geom_run(m, zone_kind, ctl)
        nnz = len(g["col_idx"])
        out = dict(gp=r.take(np.float64, 3 * nc).reshape(nc, 3), gu=r.take(np.float64, 9 * nc).reshape(nc, 9),
                   gt=r.take(np.float64, 3 * nc).reshape(nc, 3))
        for comp in (1, 2, 3):
            out[f"mom{comp}_val"] = r.take(np.float64, nnz)
            out[f"mom{comp}_rhs"] = r.take(np.float64, nc)
            out[f"mom{comp}_ap"] = r.take(np.float64, nc)
        out["tem_val"] = r.take(np.float64, nnz)
        out["tem_rhs"] = r.take(np.float64, nc)
        out["tem_ap"] = r.take(np.float64, nc)
        out["flux_rc"] = r.take(np.float64, nf)
        out["flux_out"] = r.take(np.float64, nf)
        out["cor_u"] = r.take(np.float64, 3 * nc).reshape(nc, 3)
        out["cor_p"] = r.take(np.float64, nc)
        out["cor_flux"] = r.take(np.float64, nf)
        out["cor_dpp"] = r.take(np.float64, 3 * nc).reshape(nc, 3)
    return inputs, out


def funcs2_run(m, zone_kind, ctl, seed, ts_order):
    """Per-function vectors with porous / solid zones (harness mode funcs2)."""
    rng = np.random.default_rng(seed)
    nc, nf = m.ncells, m.nfaces
    u, p, T = seeded_fields(m, seed)
    flux = rng.normal(size=nf) * 0.1
    u_old = u + 0.05 * rng.normal(size=u.shape)
    u_oo = u + 0.05 * rng.normal(size=u.shape)
    T_old = T + 0.05 * rng.normal(size=nc)
    T_oo = T + 0.05 * rng.normal(size=nc)
    apc = rng.uniform(1.0, 3.0, size=nc)
    pp = rng.normal(size=nc) * 0.01
    Ts = T + 0.2 * rng.normal(size=nc)
    Ts_old = Ts + 0.05 * rng.normal(size=nc)
    Ts_oo = Ts + 0.05 * rng.normal(size=nc)
    inputs = dict(u=u, p=p, T=T, flux=flux, u_old=u_old, u_old_old=u_oo, T_old=T_old, T_old_old=T_oo,
                  apc=apc, pp=pp, ts_order=np.int64(ts_order), T_s=Ts, T_s_old=Ts_old, T_s_old_old=Ts_oo)
    with tempfile.TemporaryDirectory() as td:
        cas.write_cas(os.path.join(td, "m.cas"), m, zone_kind)
        open(os.path.join(td, "c.control"), "w").write(ctl)
        inf = os.path.join(td, "in.bin")
        write_fields(inf, u, p, T, flux, u_old, u_oo, T_old, T_oo, apc, pp, np.array([ts_order], np.int32),
                     Ts, Ts_old, Ts_oo)
        run(["m.cas", "c.control", "funcs2", "out.bin", inf], td)
        r = Reader(os.path.join(td, "out.bin"))
        g = geom_run(m, zone_kind, ctl)
        nnz = len(g["col_idx"])
        out = dict(cztype=r.take(np.int32, nc), perm_dir=r.take(np.float64, 3 * nc).reshape(nc, 3))
        for k in ("inertial", "porosity", "disp_l", "disp_t", "bj_alpha", "k_s", "cp_s", "rho_s", "h_sf", "a_sf",
                  "q_src"):
            out[k] = r.take(np.float64, nc)
        out.update(gp=r.take(np.float64, 3 * nc).reshape(nc, 3), gu=r.take(np.float64, 9 * nc).reshape(nc, 9),
                   gt=r.take(np.float64, 3 * nc).reshape(nc, 3), gts=r.take(np.float64, 3 * nc).reshape(nc, 3))
        for comp in (1, 2, 3):
            out[f"mom{comp}_val"] = r.take(np.float64, nnz)
            out[f"mom{comp}_rhs"] = r.take(np.float64, nc)
            out[f"mom{comp}_ap"] = r.take(np.float64, nc)
        for t in ("tem", "sol"):
            out[f"{t}_val"] = r.take(np.float64, nnz)
            out[f"{t}_rhs"] = r.take(np.float64, nc)
            out[f"{t}_ap"] = r.take(np.float64, nc)
        out["flux_rc"] = r.take(np.float64, nf)
        out["flux_out"] = r.take(np.float64, nf)
        out["cor_u"] = r.take(np.float64, 3 * nc).reshape(nc, 3)
        out["cor_p"] = r.take(np.float64, nc)
        out["cor_flux"] = r.take(np.float64, nf)
        out["cor_dpp"] = r.take(np.float64, 3 * nc).reshape(nc, 3)
        out["cor_u_old"] = r.take(np.float64, 3 * nc).reshape(nc, 3)
    return inputs, out


POROUS_CTL = """rho = 1.0
mu = 0.05
alpha_u = 0.7
alpha_p = 0.3
transient = true
dt = 0.02
n_time_max = 2
n_correct = 2
n_out_every = 1000000
time_scheme = 2
outer_max = 3
lin_tol = 1e-10
lin_max = 2000
conv_blend = 0.5
nonorth_corr = 1.0
ppe_precond = jacobi
k_cond = 0.1
cp = 1.0
{model}
bc = 6 velocity-inlet-parabolic 1.0 2.0 2 -1.0 1.0
bc = 7 outflow
bc = 5 wall
bc = 4 symmetry
tbc = 5 1 0.5
tbc_plane = 5 2 -1.0 2 3.0
{zones}
"""

PLUG_LTNE = ("thermal_model = ltne",
             "cell_zone = 8 porous perm=0.01 perm_yy=0.02 inertial=0.5 porosity=0.4 disp_l=0.01 disp_t=0.002 "
             "bj_alpha=0.3 k_s=2.0 cp_s=0.8 rho_s=3.0 h_sf=5.0 a_sf=10.0")
PLUG_UNIFIED = ("porous_model = unified",
                "cell_zone = 8 porous perm=0.01 inertial=0.5 porosity=0.4 disp_l=0.01 disp_t=0.002 k_s=2.0 "
                "cp_s=0.8 rho_s=3.0")
# porous zone typed by a VC tag in its name, coefficients matched by name;
# a solid block with a heat source (conjugate heat transfer)
SOLID_CHT = ("", "cell_zone = plug perm=0.02 inertial=0.2 porosity=0.5 k_s=1.5 cp_s=0.9 rho_s=2.5\n"
                  "cell_zone = 9 solid k_s=5.0 cp_s=0.5 rho_s=2.0 q_src=4.0")


def zoned_channel(solid=False):
    """10x6x3 jittered channel with a porous plug (cells 3 <= i < 6) and,
    with solid=True, a solid block (i >= 7, j < 2) -- zones contiguous."""
    m = meshgen.hex_box(np.linspace(0, 2.0, 11), meshgen.tanh_stretch(6, 1.1), np.linspace(0, 0.6, 4),
                        zone_ids=dict(xmin=6, xmax=7, ymin=5, ymax=5, zmin=4, zmax=4), jitter=0.15, seed=21)
    idx = np.arange(m.ncells)
    I, J = idx % 10, (idx // 10) % 6
    cz = np.full(m.ncells, 2, np.int32)
    cz[(I >= 3) & (I < 6)] = 8
    names = {2: ("fluid", "free"), 8: ("fluid", "plug")}
    if solid:
        cz[(I >= 7) & (J < 2)] = 9
        names = {2: ("fluid", "free"), 8: ("vc:porous", "plug"), 9: ("solid", "block")}
    return meshgen.with_cell_zones(m, cz, names)


def porous_fixtures():
    zk = {4: "symmetry", 5: "wall", 6: "velocity-inlet", 7: "outflow"}
    for tag, (model, zones), solid in (("plug_ltne", PLUG_LTNE, False), ("plug_unified", PLUG_UNIFIED, False),
                                       ("solid_cht", SOLID_CHT, True)):
        m = zoned_channel(solid)
        ctl = POROUS_CTL.format(model=model, zones=zones)
        fx = {}
        for tso in (0, 2):
            inputs, out = funcs2_run(m, zk, ctl, seed=31 + tso, ts_order=tso)
            fx.update({f"ts{tso}_in_" + k: v for k, v in inputs.items()})
            fx.update({f"ts{tso}_out_" + k: v for k, v in out.items()})
        u0, p0, T0 = seeded_fields(m, 41, umag=0.3)
        u0[:, 0] += 1.0
        simple, _ = case_run(tag, m, zk, ctl, "simple", 3, (u0, p0, T0), want_ts=True)
        piso, _ = case_run(tag, m, zk, ctl, "piso", 2, (u0, p0, T0), want_ts=True)
        save(f"porous_{tag}", control=ctl, **mesh_arrays(m), in_u=u0, in_p=p0, in_T=T0,
             **{"simple_" + k: v for k, v in simple.items()}, **{"piso_" + k: v for k, v in piso.items()}, **fx)


def solve_run(row_ptr, col_idx, val, b, x0, tol, kmax, maxit):
    n, nnz = len(b), len(val)
    with tempfile.TemporaryDirectory() as td:
        inf = os.path.join(td, "in.bin")
        write_fields(inf, np.array([n, nnz], np.int32), row_ptr.astype(np.int32), col_idx.astype(np.int32),
                     val, b, x0, np.array([tol]), np.array([kmax], np.int32))
        run(["-", "-", "solve", "out.bin", inf, str(maxit)], td)
        r = Reader(os.path.join(td, "out.bin"))
        out = {}
        out["ic0_ierr"] = np.int64(r.take(np.int32, 1)[0])
        out["ic0_L"] = r.take(np.float64, nnz)
        out["ilu0_ierr"] = np.int64(r.take(np.int32, 1)[0])
        out["ilu0_LU"] = r.take(np.float64, nnz)
        for s in ("cg_jacobi", "cg_ic0", "bicgstab_ilu0"):
            its, errs, res = [], [], []
            for _ in range(kmax):
                its.append(r.take(np.int32, 1)[0])
                errs.append(r.take(np.int32, 1)[0])
                res.append(r.take(np.float64, 1)[0])
            out[f"{s}_hist_it"] = np.array(its)
            out[f"{s}_hist_ierr"] = np.array(errs)
            out[f"{s}_hist_res"] = np.array(res)
            out[f"{s}_it"] = np.int64(r.take(np.int32, 1)[0])
            out[f"{s}_ierr"] = np.int64(r.take(np.int32, 1)[0])
            out[f"{s}_res"] = np.float64(r.take(np.float64, 1)[0])
            out[f"{s}_x"] = r.take(np.float64, n)
    return out


def save(name, **kw):
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **kw)
    print("wrote", path, os.path.getsize(path), "bytes")


REF_CAVITY = "/root/reference/cluster_cases/cavity_Re1000"


def cavity_known_answer():
    """The reference's shipped Re=1000 lid-driven cavity (80x80x1 hex mesh,
    unMesh.cas + unMesh.control of cluster_cases/cavity_Re1000): its README
    (lines 179-180) records max|u| = 9.11512E-01 and the Ghia centreline
    relative L2 error 2.89706E-02 for this control (outer_max = 1000, no
    early exit).  The harness re-runs the reference's simple_run for 1000
    iterations and the fixture keeps both the published and the full-precision
    numbers."""
    from mixsolver_amd.post import ghia_compare
    from mixsolver_amd.control import parse_control
    from oracle import pyoracle
    m = cas.read_cas(os.path.join(REF_CAVITY, "unMesh.cas"))
    ctl = "\n".join(l for l in open(os.path.join(REF_CAVITY, "unMesh.control")).read().splitlines()
                    if not l.strip().startswith("out_format")) + "\nn_out_every = 1000000\n"
    with tempfile.TemporaryDirectory() as td:
        cas.write_cas(os.path.join(td, "m.cas"), m, {4: "symmetry", 5: "wall"})
        open(os.path.join(td, "c.control"), "w").write(ctl)
        run(["m.cas", "c.control", "simple", "out.bin", "", "1000"], td)
        r = Reader(os.path.join(td, "out.bin"))
        ier = r.take(np.int32, 1)[0]
        u = r.take(np.float64, 3 * m.ncells).reshape(m.ncells, 3)
    ctrl, groups = parse_control(ctl, warn=False)
    xc = pyoracle.OracleCase(m, ctrl, groups).geometry()["xc"].reshape(-1, 3)
    l2, _ = ghia_compare(xc, u, 1000)
    return m, ctl, int(ier), float(np.linalg.norm(u, axis=1).max()), float(l2)


def cavity128_fixture():
    """Config 1: the Re = 1000 lid-driven cavity on 128 x 128 x 1 uniform hex
    cells (16,384), plain FP64 steady SIMPLE to outer_tol = 1e-8 (PCG-Jacobi
    p', BiCGSTAB-ILU(0) momentum, under-relaxation 0.7 / 0.3), run by the
    reference's simple_run until it converges.  Keeps the iteration count,
    the full final fields and the printed diagnostics."""
    m = meshgen.cavity(128)
    zk = {4: "symmetry", 5: "wall"}
    ctl = CAVITY_CTL.format(blend=0.0, nonorth=1.0, precond="jacobi").replace("outer_max = 1000",
                                                                             "outer_max = 20000")
    out, log = case_run("cav128", m, zk, ctl, "simple", 20000)
    its = [int(ln.split()[0]) for ln in (l.split("CONVERGED after ")[1] for l in log.splitlines()
                                           if "CONVERGED after" in l)]
    save("cavity128_simple_c1", control=ctl, **mesh_arrays(m), iters=np.int64(its[0] if its else -1),
         **{"out_" + k: v for k, v in out.items() if k in ("ier", "u", "p", "flux", "log")})


def channel2m_twin_fixture():
    """Config 2's non-periodic twin at full size: the 400 x 100 x 50 channel
    (2,000,000 hex cells, tanh wall clustering 1.1) with a parabolic inlet,
    outflow and z symmetry, seeded with the benchmark's initial field
    (laminar parabola + 5% uniform noise, seed 180), one PISO step of the
    reference's piso_run (2 correctors, BDF1 first step, p' PCG-Jacobi and
    momentum BiCGSTAB-ILU(0) to 1e-10).  The fields are too large to keep,
    so the fixture holds a signature: the printed step line, the norms of
    u, p and flux, and u, p at 1,000 seeded probe cells."""
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
    import bench
    nx, ny, nz = 400, 100, 50
    m = meshgen.channel(nx, ny, nz)
    u0 = bench.initial_field(m, ny, nz, nx)
    ctl = bench.control_text(nx, periodic=False) + "n_out_every = 1000000\n"
    zk = {4: "symmetry", 5: "wall", 6: "velocity-inlet", 7: "outflow"}
    out, log = case_run("chan2m", m, zk, ctl, "piso", 1, (u0, np.zeros(m.ncells), np.zeros(m.ncells)))
    t_ref = [float(ln.split()[1]) for ln in log.splitlines() if ln.startswith("HARNESS_DRIVER_SECONDS")]
    probes = np.sort(np.random.default_rng(2000).choice(m.ncells, 1000, replace=False))
    save("channel2m_twin_signature", control=ctl, nx=np.int64(nx), ny=np.int64(ny), nz=np.int64(nz),
         out_ier=out["ier"], out_log=out["log"], probes=probes, probe_u=out["u"][probes], probe_p=out["p"][probes],
         norm_u=np.float64(np.linalg.norm(out["u"])), norm_p=np.float64(np.linalg.norm(out["p"])),
         norm_flux=np.float64(np.linalg.norm(out["flux"])), sum_p=np.float64(out["p"].sum()),
         ref_driver_seconds=np.float64(t_ref[0] if t_ref else -1.0))


REF_BJ = "/root/reference/cluster_cases/bj_porous"
BJ_README = dict(  # cluster_cases/bj_porous/README.md:51-54 (table), :65 (iterations), :69 (it=1 line)
    readme_alpha=np.array([0.0, 1.0, 2.0]),
    readme_uF1=np.array([8.5512e-03, 1.1768e-02, 1.0038e-02]),
    readme_uF1_ref=np.array([8.4977e-03, 1.2117e-02, 1.0447e-02]),
    readme_uD=np.array([2.9946e-03, 2.9180e-03, 2.9592e-03]),
    readme_uD_ref=np.float64(2.7086e-03),
    readme_uB_class=np.array([3.3857e-02, 7.9000e-03, 5.5402e-03]),
    readme_rms_pct=np.array([0.22, 0.94, 1.20]),
    readme_iters=np.array([754, 882, 809]),
    readme_it1_a1=np.array([3.110e-18, 2.120e-04, 1.36e-09]))  # mass-imbal, du_max, lin-res


def bj_known_answer():
    """The reference's shipped Beavers-Joseph fluid/porous channel (100x40x1
    hex mesh, upper half fluid, lower half a porous bed tagged 'VC:porous' in
    the mesh; controls variants/bj_a{0,1,2}.control: continuity and stress
    jump alpha = 1, 2).  Its README publishes, per alpha, the converged
    iteration count, the first-fluid-cell and bed velocities uF1 / uD and
    the two-layer closed form; the harness re-runs the reference's
    simple_run to convergence (outer_max 2000) and the fixture keeps its
    velocity field, iteration count and it=1 log line beside the published
    numbers."""
    import re
    m = cas.read_cas(os.path.join(REF_BJ, "unMesh.cas"))
    zk = {5: "symmetry", 6: "wall", 7: "wall", 8: "pressure-outlet", 9: "pressure-outlet"}
    out = {}
    for a in range(3):
        ctl = open(os.path.join(REF_BJ, "variants", f"bj_a{a}.control")).read()
        with tempfile.TemporaryDirectory() as td:
            cas.write_cas(os.path.join(td, "m.cas"), m, zk)
            open(os.path.join(td, "c.control"), "w").write(ctl)
            log = run(["m.cas", "c.control", "simple", "out.bin", "", "2000"], td)
            r = Reader(os.path.join(td, "out.bin"))
            ier = r.take(np.int32, 1)[0]
            u = r.take(np.float64, 3 * m.ncells).reshape(m.ncells, 3)
        it = int(re.search(r"CONVERGED after\s+(\d+)", log).group(1))
        l1 = re.search(r"it=\s+1\s+mass-imbal=\s*(\S+)\s+du_max=\s*(\S+).*lin-it=\s*(\d+)\s+lin-res=\s*(\S+)", log)
        out.update({f"control_a{a}": ctl, f"ref_ier_a{a}": np.int64(ier), f"ref_u_a{a}": u,
                    f"ref_iters_a{a}": np.int64(it),
                    f"ref_it1_a{a}": np.array([float(l1.group(1)), float(l1.group(2)), float(l1.group(4))])})
    return m, out


REF_HC = "/root/reference/cases/heat_conduction"
HC_CASES = dict(c1=("c1_1d.cas", "c1_1d_fixedT"), c2=("c1_1d.cas", "c2_1d_adiabatic"),
                c3=("c3_annulus.cas", "c3_annulus_fixedT"), c4=("c3_annulus.cas", "c4_annulus_adiabatic"),
                c5=("c5_1d_src.cas", "c5_1d_source"))
# cases/heat_conduction/README.md:80-84: max |T_s - T_exact|, RMS, tolerance
HC_README = dict(c1=(0.0, 0.0, 1e-6), c2=(0.0, 0.0, 1e-6), c3=(3.912e-01, 2.869e-01, 1.0),
                 c4=(0.0, 0.0, 1e-6), c5=(5.208e-02, 3.294e-02, 0.1))


def heat_conduction_known_answer():
    """The reference's pure-conduction family (cases/heat_conduction: all
    cells solid, k_s = 15, fixed-temperature / adiabatic walls, a volumetric
    source zone in c5): its shipped meshes and controls re-run by the
    reference's simple_run (one outer iteration reaches the tolerance), with
    the solid temperature kept beside the README's published maximum and RMS
    errors against the closed-form profiles."""
    import re
    out = {}
    for c, (mf, cf) in HC_CASES.items():
        m = cas.read_cas(os.path.join(REF_HC, mf))
        ctl = open(os.path.join(REF_HC, cf + ".control")).read()
        with tempfile.TemporaryDirectory() as td:
            shutil.copy(os.path.join(REF_HC, mf), os.path.join(td, "m.cas"))
            open(os.path.join(td, "c.control"), "w").write(ctl)
            log = run(["m.cas", "c.control", "simple", "out.bin", "", "3000"], td)
            r = Reader(os.path.join(td, "out.bin"))
            nc, nf = m.ncells, m.nfaces
            ier = r.take(np.int32, 1)[0]
            r.take(np.float64, 3 * nc + nc + nf + nc + 3 * nc + 9 * nc + nc + 3 * nc)
            Ts = r.take(np.float64, nc)
        it = int(re.search(r"CONVERGED after\s+(\d+)", log).group(1))
        out.update({f"{c}_{k}": v for k, v in mesh_arrays(m).items()})
        out.update({f"{c}_control": ctl, f"{c}_ref_ier": np.int64(ier), f"{c}_ref_iters": np.int64(it),
                    f"{c}_ref_T_s": Ts, f"{c}_readme": np.array(HC_README[c])})
    return out


REF_CYL = "/root/reference/cluster_cases/cylinder_Re40"
# cluster_cases/cylinder_Re40/README.md:48-56 (this bundle's column)
CYL_README = dict(readme_cp_front=1.5018, readme_cp_base=-0.4925, readme_cp_min=-0.9863, readme_phi_cp_min=-82.9,
                  readme_sep=np.array([51.7, 51.6]), readme_wake_D=2.3420, readme_umax=1.216703,
                  readme_max_reverse=-0.10211, readme_cd_pressure=1.0766, readme_dcp_mean=0.0082,
                  readme_dcp_max=0.1146)


def cylinder_known_answer():
    """The reference's shipped Re=40 cylinder (3928-cell O-grid, far field
    at r = 10, limited second-order upwind, IC0 PCG, 8000 SIMPLE iterations:
    the case stops at the iteration cap by design).  The harness re-runs
    simple_run, then compute_forces / write_cp_distribution /
    compute_separation on the final fields; the fixture keeps the final u
    and p, the force outputs (full precision) and the pressure / viscous
    parts (printed), the Cp table and separation lines as written, the
    archived wall-Cp baseline shipped with the case and the README's
    regression indicators."""
    import re
    m = cas.read_cas(os.path.join(REF_CYL, "unMesh.cas"))
    ctl = open(os.path.join(REF_CYL, "unMesh.control")).read()
    nc, nf = m.ncells, m.nfaces
    with tempfile.TemporaryDirectory() as td:
        shutil.copy(os.path.join(REF_CYL, "unMesh.cas"), os.path.join(td, "m.cas"))
        open(os.path.join(td, "c.control"), "w").write(ctl)
        run(["m.cas", "c.control", "simple", "out.bin", "", "8000"], td)
        r = Reader(os.path.join(td, "out.bin"))
        ier = r.take(np.int32, 1)[0]
        u = r.take(np.float64, 3 * nc).reshape(nc, 3)
        p = r.take(np.float64, nc)
        write_fields(os.path.join(td, "in.bin"), u, p, np.zeros(nc))
        log = run(["m.cas", "c.control", "forces", "fo.bin", "in.bin", "1"], td)
        forces = np.fromfile(os.path.join(td, "fo.bin"), np.float64)
        cp = np.loadtxt(os.path.join(td, "cp.dat"), comments="#")

    def num(label):
        return float(re.search(re.escape(label) + r"\s*:\s*(\S+)", log).group(1))

    printed = np.array([num("Pressure  Fx (drag)"), num("Viscous   Fx (drag)"), num("CD_p (pressure)"),
                        num("CD_v (viscous)")])
    sep = np.array([num("Top separation theta"), num("Bot separation theta"),
                    num("Wake length (x_wake - xcyl)/D")])
    base = np.loadtxt(os.path.join(REF_CYL, "reference_cylinder_cp.dat"), comments="#")
    return dict(**mesh_arrays(m), control=ctl, nsteps=np.int64(8000), ref_ier=np.int64(ier), ref_u=u, ref_p=p,
                ref_forces=forces, ref_forces_printed=printed, ref_cp_table=cp, ref_separation_printed=sep,
                archived_cp=base[:, :2], **CYL_README)


REF_CHAN = "/root/reference/cluster_cases/channel_lowspeed"
# cluster_cases/channel_lowspeed/README.md:46-49 (it=1 line, iterations), :56-66 (table)
CHAN_README = dict(readme_u_bulk_50=1.567710e-04, readme_u_bulk_out=1.565399e-04, readme_umax_ratio_out=1.5006,
                   readme_umag_max=2.354425e-04, readme_umag_max_at=np.array([298.5, 5.116, 5.0]),
                   readme_profile_err_max=0.484e-2, readme_profile_err_l1=0.248e-2, readme_dpdx_fit=-3.472454e-10,
                   readme_entrance_length=55.5, readme_mass_drift=1.015e-2, readme_p_min=-1.317361e-07,
                   readme_p_rise=1.1633e-07, readme_iterations=np.int64(773),
                   readme_it1=np.array([5.113e-13, 2.840e-04]), readme_it1_lin=np.int64(278))


def channel_known_answer():
    """The reference's shipped low-speed planar channel (300 x 100 x 1 hexes,
    wall-clustered in y, mass-flow inlet 1.846e-4, outflow, Re_H = 100, IC0
    PCG, SIMPLE to an absolute du of 1e-10).  The mesh is the tensor grid
    x = 0..300 (unit cells), z = {0, 10} and the 101 y levels kept here
    (the shipped nodes sit on them within 1e-14).  The fixture keeps the
    control, the reference's convergence record, its per-layer bulk velocity
    and mean pressure, the outlet-layer profile, the regression indicators of
    the case's check on the final fields and the README's table."""
    import re
    sys.path.insert(0, os.path.dirname(HERE))
    import casecheck
    from mixsolver_amd.control import parse_control
    from oracle import pyoracle
    m = cas.read_cas(os.path.join(REF_CHAN, "unMesh.cas"))
    ctl = open(os.path.join(REF_CHAN, "unMesh.control")).read()
    y = np.sort(m.x[:, 1])
    ys = np.array([g.mean() for g in np.split(y, np.nonzero(np.diff(y) > 1e-9)[0] + 1)])
    ys[0], ys[-1] = 0.0, 10.0
    nc = m.ncells
    with tempfile.TemporaryDirectory() as td:
        shutil.copy(os.path.join(REF_CHAN, "unMesh.cas"), os.path.join(td, "m.cas"))
        open(os.path.join(td, "c.control"), "w").write(ctl)
        log = run(["m.cas", "c.control", "simple", "out.bin", "", "8000"], td)
        r = Reader(os.path.join(td, "out.bin"))
        ier = r.take(np.int32, 1)[0]
        u = r.take(np.float64, 3 * nc).reshape(nc, 3)
        p = r.take(np.float64, nc)
    ctrl, groups = parse_control(ctl, warn=False)
    xc = pyoracle.OracleCase(m, ctrl, groups).geometry()["xc"].reshape(-1, 3)
    chk = casecheck.channel_check(xc, u, p)
    it = [int(k) for k in re.findall(r"CONVERGED after\s+(\d+)", log)]
    it1 = re.search(r"it=\s*1\s+mass-imbal=\s*(\S+)\s+du_max=\s*(\S+).*?lin-it=\s*(\d+)", log)
    lv = np.unique(np.round(xc[:, 0], 9))
    mo = np.abs(xc[:, 0] - lv[-1]) < 1e-6
    o = np.argsort(xc[mo, 1])
    return dict(ys=ys, control=ctl, ref_ier=np.int64(ier), ref_iterations=np.int64(it[0] if it else -1),
                ref_it1=np.array([float(it1.group(1)), float(it1.group(2))]) if it1 else np.zeros(2),
                ref_it1_lin=np.int64(it1.group(3)) if it1 else np.int64(-1),
                ref_out_y=xc[mo, 1][o], ref_out_u=u[mo, 0][o],
                **{"ref_" + k: np.asarray(v) for k, v in chk.items()}, **CHAN_README)


def vtu_fixtures():
    """The reference's vtk_write files (inline ASCII and appended binary) for
    three small cases -- a cavity (quad faces, fluid only), a jittered prism
    O-grid (triangle and quad faces) and the conduction bar with a source zone
    (solid cells: temperature_solid and heat_source_solid written) -- with
    seeded fields that include a negative zero and extreme exponents.  The
    fixture keeps the inputs and the SHA-256 digest and size of each file."""
    import hashlib
    out = {}
    mp = meshgen.cylinder_prisms(ntheta=12, nr=3, r_out=4.0, layers=1, lz=0.2, jitter=0.15, seed=7)
    hc = os.path.join(REF_HC, "c5_1d_src.cas")
    cases = [("cav", meshgen.cavity(8), {4: "symmetry", 5: "wall"},
              CAVITY_CTL.format(blend=0.0, nonorth=1.0, precond="jacobi"), None),
             ("prism", mp, {4: "symmetry", 5: "wall", 6: "pressure-far-field"}, PRISM_CTL, None)]
    if os.path.exists(hc):
        cases.append(("solid", cas.read_cas(hc), None, open(os.path.join(REF_HC, "c5_1d_source.control")).read(), hc))
    for tag, m, zk, ctl, src in cases:
        u, p, T = seeded_fields(m, 11)
        p[0], p[1], p[2] = -0.0, 1.0e-300, -2.5e+200
        u[0, 1] = 0.0
        with tempfile.TemporaryDirectory() as td:
            if src:
                shutil.copy(src, os.path.join(td, "m.cas"))
            else:
                cas.write_cas(os.path.join(td, "m.cas"), m, zk)
            open(os.path.join(td, "c.control"), "w").write(ctl)
            write_fields(os.path.join(td, "in.bin"), u, p, T)
            run(["m.cas", "c.control", "vtu", "o.bin", "in.bin", "1"], td)
            for kind in ("a", "b"):
                data = open(os.path.join(td, f"{kind}.vtu"), "rb").read()
                out[f"{tag}_{kind}_sha256"] = hashlib.sha256(data).hexdigest()
                out[f"{tag}_{kind}_size"] = np.int64(len(data))
        out.update({f"{tag}_{k}": v for k, v in mesh_arrays(m).items()})
        out.update({f"{tag}_control": ctl, f"{tag}_u": u, f"{tag}_p": p, f"{tag}_T": T})
    return out


FMT_PROGRAM = """program fmt_probe
   implicit none
   integer, parameter :: dp = kind(1.0d0)
   integer :: n, i, u
   real(dp), allocatable :: v(:)
   open(newunit=u, file='v.bin', access='stream', form='unformatted', status='old')
   read(u) n
   allocate(v(n))
   read(u) v
   close(u)
   open(newunit=u, file='es.txt', status='replace')
   do i = 1, n
      write(u,'(es24.16)') v(i)
   end do
   close(u)
   open(newunit=u, file='es3.txt', status='replace')
   do i = 1, n
      write(u,'(1x,es24.16e3)') v(i)
   end do
   close(u)
end program fmt_probe
"""


def fortran_formats():
    """flang's own ES24.16 and ES24.16E3 output (the edit descriptors of the
    field dump and of the VTU files) for a set of values that exercise the
    rounding and exponent rules: signed zeros, denormals, the extremes,
    values that round up to the next power of ten at the 17th digit, two-
    and three-digit exponents on both sides.  Built from the program above
    with the image's amdflang (no reference code involved)."""
    fc = os.environ.get("FC", "/opt/rocm/llvm/bin/amdflang")
    rng = np.random.default_rng(99)
    v = np.concatenate([
        [0.0, -0.0, 1.0, -1.0, 0.1, 1.0 / 3.0, 2.0 / 3.0, 5e-324, 2.2250738585072014e-308, 1.7976931348623157e308,
         1e100, 1e-100, 1e99, 1e-99, 9.99999999999999999e99, 9.9999999999999999e-100, 9.999999999999999e99,
         0.99999999999999999, 123456789.123456789, -9.5, 0.5, 1e16, 1e17, 9.87654321e-5],
        rng.normal(size=40) * 10.0 ** rng.integers(-320, 300, size=40),
        rng.uniform(-1, 1, size=20)])
    with tempfile.TemporaryDirectory() as td:
        open(os.path.join(td, "p.f90"), "w").write(FMT_PROGRAM)
        subprocess.run([fc, "-O1", "-o", os.path.join(td, "p"), os.path.join(td, "p.f90")], check=True,
                       capture_output=True)
        with open(os.path.join(td, "v.bin"), "wb") as f:
            np.array([len(v)], np.int32).tofile(f)
            v.tofile(f)
        subprocess.run([os.path.join(td, "p")], cwd=td, check=True, capture_output=True)
        es = open(os.path.join(td, "es.txt")).read().split("\n")[: len(v)]
        es3 = open(os.path.join(td, "es3.txt")).read().split("\n")[: len(v)]
    return dict(values=v, es24_16=np.array(es), es24_16e3=np.array(es3))


IFACE_CASES = {   # name: (p_dir, slip, zhang, T_ltne)
    "iface_legacy": (0, 0, 0, 0),
    "iface_zhang": (0, 0, 1, 0),
    "iface_pdir_slip_ltne": (1, 1, 0, 1),
}


def iface_data(m, seed):
    """Seeded coupling-interface data over all faces."""
    rng = np.random.default_rng(seed)
    nf = m.nfaces
    vel = np.zeros((nf, 3))
    vel[:, 0] = 1.0 + 0.2 * rng.uniform(-1, 1, nf)
    vel[:, 1:] = 0.1 * rng.normal(size=(nf, 2))
    tmode = np.zeros(nf, np.int32)
    tmode[::3] = 1
    return dict(vel=vel, p=0.1 * rng.normal(size=nf), T=rng.uniform(0.5, 1.5, nf), Ts=rng.uniform(0.5, 1.5, nf),
                q=rng.normal(size=nf), slip_C=rng.uniform(0.01, 0.2, nf), peer_vel=rng.normal(size=(nf, 3)),
                T_mode=tmode)


def write_iface(f, sw, d):
    np.array(sw, np.int32).tofile(f)
    for k in ("vel", "p", "T", "Ts", "q", "slip_C", "peer_vel"):
        np.ascontiguousarray(d[k], np.float64).tofile(f)
    np.ascontiguousarray(d["T_mode"], np.int32).tofile(f)


def iface_fixtures():
    """Coupling-interface faces (mod_uns_bc.f90:282-323, :460-474, :527-545,
    :671-701, :879-893; mod_uns_simple.f90:1144-1206): the inlet zone of the
    zoned LTNE channel becomes a named interface zone without a bc line; the
    harness applies seeded per-face values through set_interface_* and the
    four switches, then writes the per-function outputs (ts_order 0 and 2)
    and a 3-iteration serial simple_run."""
    m = zoned_channel()
    zk = {4: "symmetry", 5: "wall", 6: "interface", 7: "outflow"}
    ctl = "\n".join(l for l in POROUS_CTL.format(model=PLUG_LTNE[0], zones=PLUG_LTNE[1]).splitlines()
                    if not l.startswith("bc = 6 "))
    for name, sw in IFACE_CASES.items():
        d = iface_data(m, 400 + len(name))
        fx = {}
        for tso in (0, 2):
            rng = np.random.default_rng(500 + tso)
            nc, nf = m.ncells, m.nfaces
            u, p, T = seeded_fields(m, 500 + tso)
            u[:, 0] += 1.0
            flux = rng.normal(size=nf) * 0.1
            u_old = u + 0.05 * rng.normal(size=u.shape)
            u_oo = u + 0.05 * rng.normal(size=u.shape)
            T_old = T + 0.05 * rng.normal(size=nc)
            T_oo = T + 0.05 * rng.normal(size=nc)
            apc = rng.uniform(1.0, 3.0, size=nc)
            pp = rng.normal(size=nc) * 0.01
            Ts = T + 0.2 * rng.normal(size=nc)
            Ts_old = Ts + 0.05 * rng.normal(size=nc)
            Ts_oo = Ts + 0.05 * rng.normal(size=nc)
            ins = dict(u=u, p=p, T=T, flux=flux, u_old=u_old, u_old_old=u_oo, T_old=T_old, T_old_old=T_oo,
                       apc=apc, pp=pp, T_s=Ts, T_s_old=Ts_old, T_s_old_old=Ts_oo)
            with tempfile.TemporaryDirectory() as td:
                cas.write_cas(os.path.join(td, "m.cas"), m, zk)
                open(os.path.join(td, "c.control"), "w").write(ctl)
                inf = os.path.join(td, "in.bin")
                with open(inf, "wb") as f:
                    for a in (u, p, T, flux, u_old, u_oo, T_old, T_oo, apc, pp, np.array([tso], np.int32), Ts,
                              Ts_old, Ts_oo):
                        np.ascontiguousarray(a).tofile(f)
                    write_iface(f, sw, d)
                run(["m.cas", "c.control", "ifuncs", "out.bin", inf], td)
                r = Reader(os.path.join(td, "out.bin"))
                g = geom_run(m, zk, ctl)
                nnz = len(g["col_idx"])
                r.take(np.int32, nc)
                r.take(np.float64, 3 * nc + 11 * nc)
                out = dict(gp=r.take(np.float64, 3 * nc).reshape(nc, 3), gu=r.take(np.float64, 9 * nc).reshape(nc, 9),
                           gt=r.take(np.float64, 3 * nc).reshape(nc, 3), gts=r.take(np.float64, 3 * nc).reshape(nc, 3))
                for comp in (1, 2, 3):
                    out[f"mom{comp}_val"] = r.take(np.float64, nnz)
                    out[f"mom{comp}_rhs"] = r.take(np.float64, nc)
                    out[f"mom{comp}_ap"] = r.take(np.float64, nc)
                for t in ("tem", "sol"):
                    out[f"{t}_val"] = r.take(np.float64, nnz)
                    out[f"{t}_rhs"] = r.take(np.float64, nc)
                    out[f"{t}_ap"] = r.take(np.float64, nc)
                out["flux_rc"] = r.take(np.float64, nf)
                out["flux_out"] = r.take(np.float64, nf)
                out["cor_u"] = r.take(np.float64, 3 * nc).reshape(nc, 3)
                out["cor_p"] = r.take(np.float64, nc)
                out["cor_flux"] = r.take(np.float64, nf)
                out["cor_dpp"] = r.take(np.float64, 3 * nc).reshape(nc, 3)
                out["cor_u_old"] = r.take(np.float64, 3 * nc).reshape(nc, 3)
            fx.update({f"ts{tso}_in_{k}": v for k, v in ins.items()})
            fx.update({f"ts{tso}_out_{k}": v for k, v in out.items()})
        # serial simple_run, 3 iterations, from the ts_order-0 fields
        with tempfile.TemporaryDirectory() as td:
            cas.write_cas(os.path.join(td, "m.cas"), m, zk)
            open(os.path.join(td, "c.control"), "w").write(ctl)
            inf = os.path.join(td, "in.bin")
            with open(inf, "wb") as f:
                for a in (fx["ts0_in_u"], fx["ts0_in_p"], fx["ts0_in_T"]):
                    np.ascontiguousarray(a).tofile(f)
                write_iface(f, sw, d)
            run(["m.cas", "c.control", "isimple", "out.bin", inf, "3"], td)
            r = Reader(os.path.join(td, "out.bin"))
            nc, nf = m.ncells, m.nfaces
            sim = dict(ier=np.int64(r.take(np.int32, 1)[0]), u=r.take(np.float64, 3 * nc).reshape(nc, 3),
                       p=r.take(np.float64, nc), flux=r.take(np.float64, nf), T=r.take(np.float64, nc))
            r.take(np.float64, 3 * nc + 9 * nc + nc + 3 * nc)
            sim["T_s"] = r.take(np.float64, nc)
        save(name, control=ctl, switches=np.array(sw, np.int64), **mesh_arrays(m),
             **{"iface_" + k: v for k, v in d.items()}, **fx,
             **{"simple_" + k: v for k, v in sim.items()})


def main():
    if not os.path.exists(HARNESS):
        sys.exit("build the reference harness first: oracle/build_ref.sh")

    # 1. cavity, Jacobi PCG, orthogonal-corrected diffusion, first-order upwind
    m = meshgen.cavity(16)
    zk = {4: "symmetry", 5: "wall"}
    ctl = CAVITY_CTL.format(blend=0.0, nonorth=1.0, precond="jacobi")
    g = geom_run(m, zk, ctl)
    out, log = case_run("cav16", m, zk, ctl, "simple", 5)
    save("cavity16_simple_jacobi", control=ctl, nsteps=np.int64(5), **mesh_arrays(m),
         **{"geom_" + k: v for k, v in g.items()}, **{"out_" + k: v for k, v in out.items()})

    # 2. cavity, IC0 PCG + limited second-order upwind (the shipped cavity controls)
    ctl = CAVITY_CTL.format(blend=1.0, nonorth=0.0, precond="ic0")
    out, log = case_run("cav16b", m, zk, ctl, "simple", 4)
    save("cavity16_simple_ic0_blend", control=ctl, nsteps=np.int64(4), **mesh_arrays(m),
         **{"out_" + k: v for k, v in out.items()})

    # 3. channel, transient PISO (BDF2), parabolic inlet + outflow, seeded field
    mc = channel_box(12, 8, 4)
    zkc = {4: "symmetry", 5: "wall", 6: "velocity-inlet", 7: "outflow"}
    ctl = CHANNEL_CTL.format(mu=0.05, dt=0.02, nsteps=3, blend=0.0, precond="jacobi")
    u0, p0, T0 = seeded_fields(mc, 180, umag=0.3)
    u0[:, 0] += 1.0
    out, log = case_run("chan", mc, zkc, ctl, "piso", 3, (u0, p0, T0))
    gch = geom_run(mc, zkc, ctl)
    save("channel_piso_jacobi", control=ctl, nsteps=np.int64(3), **mesh_arrays(mc),
         in_u=u0, in_p=p0, in_T=T0, **{"geom_" + k: v for k, v in gch.items()},
         **{"out_" + k: v for k, v in out.items()})

    # 3b. the same channel under PIMPLE: 3 outer iterations, relaxed momentum, IC0, blending
    out, log = case_run("chanpim", mc, zkc, PIMPLE_CTL, "pimple", 3, (u0, p0, T0))
    save("channel_pimple_ic0", control=PIMPLE_CTL, nsteps=np.int64(3), **mesh_arrays(mc),
         in_u=u0, in_p=p0, in_T=T0, **{"out_" + k: v for k, v in out.items()})

    # 3c. the reference's shipped Re=1000 cavity and its published answer
    if os.path.isdir(REF_CAVITY):
        mk, ctlk, ierk, umax, l2 = cavity_known_answer()
        save("cavity_re1000_known_answer", control=ctlk, nsteps=np.int64(1000), **mesh_arrays(mk),
             out_ier=np.int64(ierk), ref_umax=umax, ref_ghia_l2=l2,
             readme_umax=0.911512, readme_ghia_l2=0.0289706)

    # 3i. coupling-interface faces
    iface_fixtures()

    # 3h. flang's ES24.16 / ES24.16E3 output (field dump and VTU numbers)
    save("fortran_formats", **fortran_formats())

    # 3g. the reference's VTU files (ASCII and binary) for three small cases
    save("vtu_files", **vtu_fixtures())

    # 3f. the reference's shipped Re=40 cylinder: forces, Cp, separation, published indicators
    if os.path.isdir(REF_CYL):
        save("cylinder_re40_known_answer", **cylinder_known_answer())

    # 3e. the reference's pure-conduction family and its published errors
    if os.path.isdir(REF_HC):
        save("heat_conduction_known_answer", **heat_conduction_known_answer())

    # 3d. the reference's shipped Beavers-Joseph porous channel and its published answers
    if os.path.isdir(REF_BJ):
        mb, ob = bj_known_answer()
        save("bj_porous_known_answer", **mesh_arrays(mb), **ob, **BJ_README)

    # 4. jittered prism O-grid, PISO with IC0, blended convection, far field
    mp = meshgen.cylinder_prisms(ntheta=24, nr=6, r_out=4.0, layers=2, lz=0.2, jitter=0.15, seed=200)
    zkp = {4: "symmetry", 5: "wall", 6: "pressure-far-field"}
    u0, p0, T0 = seeded_fields(mp, 200, umag=0.2)
    u0[:, 0] += 1.0
    out, log = case_run("prism", mp, zkp, PRISM_CTL, "piso", 2, (u0, p0, T0))
    gp = geom_run(mp, zkp, PRISM_CTL)
    save("prism_piso_ic0", control=PRISM_CTL, nsteps=np.int64(2), **mesh_arrays(mp),
         in_u=u0, in_p=p0, in_T=T0, **{"geom_" + k: v for k, v in gp.items()},
         **{"out_" + k: v for k, v in out.items()})

    # 5. per-function vectors on the jittered prism mesh and on the outlet box
    for tag, mm, zz, cc, tso in (("prism", mp, zkp, PRISM_CTL, 2),
                                 ("outlet", None, None, OUTLET_CTL, 0),
                                 ("channel", mc, zkc, CHANNEL_CTL.format(mu=0.05, dt=0.02, nsteps=3,
                                                                        blend=1.0, precond="jacobi"), 1)):
        if mm is None:
            mm = meshgen.hex_box(np.linspace(0, 2, 9), np.linspace(0, 1, 6), np.linspace(0, 0.5, 4),
                                 zone_ids=dict(xmin=6, xmax=7, ymin=5, ymax=5, zmin=4, zmax=4),
                                 jitter=0.2, seed=7)
            zz = {4: "wall", 5: "wall", 6: "velocity-inlet", 7: "pressure-outlet"}
            mo = mm
        inputs, out = funcs_run(mm, zz, cc, seed=11, ts_order=tso)
        save(f"funcs_{tag}", control=cc, **mesh_arrays(mm), **{"in_" + k: v for k, v in inputs.items()},
             **{"out_" + k: v for k, v in out.items()})

    # 6. pressure-outlet box, steady SIMPLE (Dirichlet-p rows, no pin, Boussinesq, heat flux)
    u0, p0, T0 = seeded_fields(mo, 5, umag=0.1)
    out, log = case_run("outlet", mo, {4: "wall", 5: "wall", 6: "velocity-inlet", 7: "pressure-outlet"},
                        OUTLET_CTL, "simple", 3, (u0, p0, T0))
    save("outlet_simple_jacobi", control=OUTLET_CTL, nsteps=np.int64(3), **mesh_arrays(mo),
         in_u=u0, in_p=p0, in_T=T0, **{"out_" + k: v for k, v in out.items()})

    # 7. linear solver histories on a pressure-correction-like SPD matrix
    rp, ci = g["row_ptr"], g["col_idx"]
    n = len(rp) - 1
    rng = np.random.default_rng(3)
    val = np.zeros(len(ci))
    for i in range(n):
        for k in range(rp[i] - 1, rp[i + 1] - 1):
            j = ci[k] - 1
            val[k] = -rng.uniform(0.5, 1.5) if j != i else 0.0
    # symmetrise and make diagonally dominant
    A = np.zeros((n, n))
    for i in range(n):
        for k in range(rp[i] - 1, rp[i + 1] - 1):
            A[i, ci[k] - 1] = val[k]
    A = 0.5 * (A + A.T)
    np.fill_diagonal(A, -A.sum(axis=1) + 0.05)
    for i in range(n):
        for k in range(rp[i] - 1, rp[i + 1] - 1):
            val[k] = A[i, ci[k] - 1]
    # non-symmetric twin for BiCGSTAB (convective skew)
    b = rng.normal(size=n)
    x0 = np.zeros(n)
    out = solve_run(rp, ci, val, b, x0, 1e-10, 25, 1000)
    save("solvers_spd", row_ptr=rp, col_idx=ci, val=val, b=b, x0=x0, tol=np.float64(1e-10),
         kmax=np.int64(25), maxit=np.int64(1000), **out)
    # nonsymmetric: the reference's own x-momentum matrix of the channel
    # per-function case (upwind convection + diffusion + BDF time term)
    fc = np.load(os.path.join(HERE, "funcs_channel.npz"))
    gch = geom_run(mc, zkc, CHANNEL_CTL.format(mu=0.05, dt=0.02, nsteps=3, blend=1.0, precond="jacobi"))
    rpn, cin = gch["row_ptr"], gch["col_idx"]
    valn = fc["out_mom1_val"]
    bn = fc["out_mom1_rhs"]
    x0 = rng.normal(size=len(bn)) * 0.1
    out = solve_run(rpn, cin, valn, bn, x0, 1e-12, 10, 1000)
    save("solvers_nonsym", row_ptr=rpn, col_idx=cin, val=valn, b=bn, x0=x0, tol=np.float64(1e-12),
         kmax=np.int64(10), maxit=np.int64(1000), **out)


if __name__ == "__main__":
    if sys.argv[1:] == ["--porous"]:
        porous_fixtures()
    elif sys.argv[1:] == ["--iface"]:
        iface_fixtures()
    elif sys.argv[1:] == ["--fmt"]:
        save("fortran_formats", **fortran_formats())
    elif sys.argv[1:] == ["--vtu"]:
        save("vtu_files", **vtu_fixtures())
    elif sys.argv[1:] == ["--chan"]:
        save("channel_lowspeed_known_answer", **channel_known_answer())
    elif sys.argv[1:] == ["--cyl"]:
        save("cylinder_re40_known_answer", **cylinder_known_answer())
    elif sys.argv[1:] == ["--hc"]:
        save("heat_conduction_known_answer", **heat_conduction_known_answer())
    elif sys.argv[1:] == ["--c2"]:
        channel2m_twin_fixture()
    elif sys.argv[1:] == ["--c1"]:
        cavity128_fixture()
    elif sys.argv[1:] == ["--restart"]:
        restart_fixture()
    elif sys.argv[1:] == ["--bj"]:
        mb, ob = bj_known_answer()
        save("bj_porous_known_answer", **mesh_arrays(mb), **ob, **BJ_README)
    else:
        main()
        porous_fixtures()
        restart_fixture()
