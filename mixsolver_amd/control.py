"""``unMesh.control`` parser and the C-ABI control / boundary-group structs.

Keys, defaults and error codes follow ``read_control`` / ``parse_bc`` /
``parse_lid`` / ``parse_tbc`` / ``parse_tbc_plane``
(mod_uns_control.f90:148-280 defaults, :325-553 keys, :1010-1152 bc grammar,
:1158-1197 lid filter, :1198-1300 thermal filters).  Unknown keywords only
warn, as in the reference (:540-541), which is how the device-path keys below
stay compatible with the reference reader:

  gpu_ppe_precond  jacobi | ic0 | dilu | ilu0 | mc_ilu0 | mc_ic0
                                                  (default: follow ppe_precond)
  gpu_mom_precond  ilu0 | jacobi | dilu | mc_ilu0 (default: ilu0, the reference)
  gpu_ppe_solver   pcg | bicgstab | schur         (default pcg, the reference's
                   cg_jacobi / cg_ic0; bicgstab solves p' with BiCGSTAB and an
                   ILU(0)-type preconditioner; schur eliminates one colour of a
                   two-colour cell graph and runs PCG with the Jacobi
                   preconditioner and the coarse space on the remaining Schur
                   complement: the iterates of multicolour DILU-PCG at the cost
                   of one matrix product per iteration)
  gpu_precision    fp64 | mixed                   (default fp64)
  gpu_inner_tol    mixed PCG: reliable-update trigger, the FP32 recursive residual
                   falls below this fraction of its value at the last FP64 update
                   (default 0.1)
  gpu_reorder      none | rcm                     (default none)
  gpu_coarse       p' PCG two-level preconditioner: number of cell aggregates
                   (recursive coordinate bisection), 0 = off, at most 1024
                   (default 0)
  periodic         zoneA zoneB [tx ty tz]: join the faces of zoneA to the faces
                   of zoneB translated by (tx, ty, tz) (default: the difference
                   of the two zones' mean face centroids); up to 4 lines; the
                   two zones take no bc line.  The reference has no periodic
                   boundary, so a control file using it runs on the device path
                   only.

``ilu0``/``ic0`` select natural-order factorizations (parity with the
reference); ``dilu`` selects the multicolour DILU preconditioner;
``mc_ilu0``/``mc_ic0`` select ILU(0) / IC(0) factorizations in the
colour-blocked cell numbering (multicolour ILU(0): the rows of one colour are
factored and swept together).  For p' the factorization is IC(0) under PCG
and ILU(0) under BiCGSTAB.
"""
from __future__ import annotations

import ctypes as C
import warnings

import numpy as np

BC_NONE, BC_WALL, BC_SYMMETRY, BC_VINLET, BC_POUTLET, BC_FARFIELD = 0, 1, 2, 3, 4, 5
BC_SLIPWALL, BC_MASSINLET, BC_OUTFLOW, BC_INTERFACE, BC_VINLET_PARAB = 6, 7, 8, 9, 10

PRECOND = {"jacobi": 0, "ic0": 1, "ilu0": 1, "dilu": 2, "mc_ilu0": 3, "mc_ic0": 3}
PPE_SOLVER = {"pcg": 0, "cg": 0, "bicgstab": 1, "schur": 2}
PRECISION = {"fp64": 0, "mixed": 1}

_BC_NAMES = {
    "wall": BC_WALL, "symmetry": BC_SYMMETRY,
    "velocity-inlet": BC_VINLET, "velocity_inlet": BC_VINLET, "inlet": BC_VINLET,
    "velocity-inlet-parabolic": BC_VINLET_PARAB, "velocity_inlet_parabolic": BC_VINLET_PARAB,
    "parabolic-inlet": BC_VINLET_PARAB, "parabolic_inlet": BC_VINLET_PARAB,
    "inlet-parabolic": BC_VINLET_PARAB,
    "pressure-outlet": BC_POUTLET, "pressure_outlet": BC_POUTLET, "outlet": BC_POUTLET,
    "pressure-far-field": BC_FARFIELD, "pressure_far_field": BC_FARFIELD,
    "farfield": BC_FARFIELD, "far-field": BC_FARFIELD,
    "slip-wall": BC_SLIPWALL, "slipwall": BC_SLIPWALL, "slip": BC_SLIPWALL,
    "mass-flow-inlet": BC_MASSINLET, "mass_flow_inlet": BC_MASSINLET,
    "mass-inlet": BC_MASSINLET, "mass_inlet": BC_MASSINLET,
    "outflow": BC_OUTFLOW, "fully-developed-outlet": BC_OUTFLOW,
    "fully_developed_outlet": BC_OUTFLOW,
}


class MsCtrl(C.Structure):
    """ms_ctrl_t (include/mixsolver.h)."""
    _fields_ = [
        ("rho", C.c_double), ("mu", C.c_double), ("alpha_u", C.c_double), ("alpha_p", C.c_double),
        ("outer_tol", C.c_double), ("lin_tol", C.c_double), ("conv_blend", C.c_double),
        ("nonorth_corr", C.c_double),
        ("dt", C.c_double), ("cp", C.c_double), ("k_cond", C.c_double), ("tref", C.c_double),
        ("beta", C.c_double),
        ("gravity", C.c_double * 3), ("body_force", C.c_double * 3),
        ("init_T", C.c_double), ("t_pert", C.c_double),
        ("outer_max", C.c_int), ("lin_max", C.c_int), ("ppe_precond", C.c_int), ("transient", C.c_int),
        ("time_scheme", C.c_int), ("n_time_max", C.c_int), ("n_correct", C.c_int), ("n_out_every", C.c_int),
        ("pimple", C.c_int), ("n_outer_iter", C.c_int), ("boussinesq", C.c_int), ("thermal_model", C.c_int),
        ("inlet_ramp", C.c_int), ("nbc", C.c_int),
        ("gpu_ppe_precond", C.c_int), ("gpu_mom_precond", C.c_int), ("gpu_precision", C.c_int),
        ("gpu_ppe_solver", C.c_int),
        ("gpu_inner_tol", C.c_double),
        ("gpu_reorder", C.c_int), ("gpu_coarse", C.c_int),
        ("nperiodic", C.c_int), ("periodic_zone", (C.c_int * 2) * 4), ("periodic_auto", C.c_int * 4),
        ("periodic_shift", (C.c_double * 3) * 4),
        ("porous_unified", C.c_int),
    ]


class MsBcGroup(C.Structure):
    """ms_bcgroup_t (include/mixsolver.h)."""
    _fields_ = [
        ("zone", C.c_int), ("btype", C.c_int),
        ("uvel", C.c_double * 3), ("pval", C.c_double), ("mdot", C.c_double), ("uspeed", C.c_double),
        ("prfl_mean", C.c_double), ("prfl_span", C.c_double), ("prfl_origin", C.c_double),
        ("prfl_axis", C.c_int), ("has_lid", C.c_int),
        ("lid_dir", C.c_int), ("ttype", C.c_int),
        ("lid_coord", C.c_double), ("lid_tol", C.c_double), ("lid_vel", C.c_double * 3),
        ("tval", C.c_double), ("qval", C.c_double),
        ("ntbc_plane", C.c_int), ("tbc_pdir", C.c_int * 4), ("tbc_pttype", C.c_int * 4),
        ("tbc_pcoord", C.c_double * 4), ("tbc_tol", C.c_double), ("tbc_ptval", C.c_double * 4),
        ("tbc_pqval", C.c_double * 4),
    ]


class ControlError(ValueError):
    def __init__(self, code, msg):
        super().__init__(f"control error {code}: {msg}")
        self.code = code


def default_ctrl() -> MsCtrl:
    c = MsCtrl()
    c.rho, c.mu, c.alpha_u, c.alpha_p = 1.0, 0.01, 0.7, 0.3
    c.outer_max, c.outer_tol, c.lin_tol, c.lin_max = 2000, 1.0e-6, 1.0e-8, 200
    c.conv_blend, c.nonorth_corr = 0.0, 1.0
    c.ppe_precond = PRECOND["ic0"]
    c.transient, c.time_scheme, c.dt, c.n_time_max = 0, 1, 0.0, 0
    c.n_correct, c.n_out_every, c.pimple, c.n_outer_iter = 2, 0, 0, 1
    c.cp, c.k_cond, c.tref, c.beta = 1005.0, 0.026, 0.0, 0.0
    c.boussinesq, c.thermal_model, c.inlet_ramp = 0, 0, 1
    c.gpu_ppe_precond, c.gpu_mom_precond, c.gpu_precision = -1, 1, 0
    c.gpu_ppe_solver, c.gpu_inner_tol, c.gpu_reorder, c.gpu_coarse = 0, 0.1, 0, 0
    c.cell_zones = []          # parsed cell_zone lines (host side, see zones.py)
    return c


CZ_AUTO, CZ_FLUID, CZ_POROUS, CZ_SOLID = 0, 1, 2, 3
_CZ_KEYS = ("perm", "perm_xx", "perm_yy", "perm_zz", "disp_l", "disp_t", "bj_alpha", "inertial",
            "porosity", "k_s", "cp_s", "rho_s", "h_sf", "a_sf", "q_src")


def _parse_cell_zone(val):
    """parse_cell_zone (mod_uns_control.f90:811-1005):
    ``<id|name> [fluid|porous|solid] [key=value ...]``.  Without a type token
    the line supplies coefficients only (CZ_AUTO)."""
    toks = val.split()
    if not toks:
        raise ControlError(9, f"bad cell_zone line: {val}")
    cz = dict(id=0, name="", ztype=CZ_AUTO, **{k: 0.0 for k in _CZ_KEYS})
    cz["porosity"] = 1.0
    try:
        cz["id"] = int(toks[0])
    except ValueError:
        cz["name"] = toks[0][:32]
    rest = toks[1:]
    if rest and "=" not in rest[0]:
        t = rest[0].lower()
        if t not in ("fluid", "porous", "solid"):
            raise ControlError(9, f"cell_zone type must be fluid, porous or solid: {rest[0]}")
        cz["ztype"] = {"fluid": CZ_FLUID, "porous": CZ_POROUS, "solid": CZ_SOLID}[t]
        rest = rest[1:]
    for tok in rest:
        if "=" not in tok:
            raise ControlError(9, f"bad cell_zone option (need key=value): {tok}")
        k, v = tok.split("=", 1)
        k = k.strip().lower()
        if k not in _CZ_KEYS:
            raise ControlError(9, f"unknown cell_zone option: {k}")
        try:
            cz[k] = float(v.replace("d", "e").replace("D", "e"))
        except ValueError:
            raise ControlError(9, f"bad {k} value: {v}") from None
    return cz


def _floats(toks, n, code, val):
    try:
        return [float(t.replace("d", "e").replace("D", "e")) for t in toks[:n]] if len(toks) >= n else None
    except ValueError:
        raise ControlError(code, val)


def _bool(val, key):
    v = val.strip().lower()
    if v in ("true", "1", "yes", "on"):
        return 1
    if v in ("false", "0", "no", "off"):
        return 0
    raise ControlError(2, f"bad {key} value: {val}")


def _parse_bc(val: str) -> MsBcGroup:
    g = MsBcGroup()
    g.prfl_axis, g.prfl_span = 2, 1.0
    toks = val.replace(",", " ").split()
    if len(toks) < 2:
        raise ControlError(6, val)
    g.zone = int(toks[0])
    name = toks[1]
    if name not in _BC_NAMES:
        raise ControlError(7, val)
    bt = _BC_NAMES[name]
    g.btype = bt
    rest = toks[2:]
    if bt == BC_VINLET:
        f = _floats(rest, 3, 6, val)
        if f is None:
            raise ControlError(6, val)
        g.uvel[:] = f
        if len(rest) >= 4:
            g.tval = float(rest[3])
    elif bt == BC_VINLET_PARAB:
        if len(rest) < 4:
            raise ControlError(6, val)
        g.prfl_mean, g.prfl_span = float(rest[0]), float(rest[1])
        g.prfl_axis, g.prfl_origin = int(rest[2]), float(rest[3])
        if len(rest) >= 5:
            g.tval = float(rest[4])
        if g.prfl_span <= 0.0 or not 1 <= g.prfl_axis <= 3:
            raise ControlError(6, val)
    elif bt == BC_POUTLET:
        f = _floats(rest, 1, 6, val)
        if f is None:
            raise ControlError(6, val)
        g.pval = f[0]
    elif bt == BC_FARFIELD:
        f = _floats(rest, 4, 6, val)
        if f is None:
            raise ControlError(6, val)
        g.pval = f[0]
        g.uvel[:] = f[1:4]
    elif bt == BC_MASSINLET:
        f = _floats(rest, 1, 6, val)
        if f is None or f[0] <= 0.0:
            raise ControlError(6, val)
        g.mdot = f[0]
        if len(rest) >= 2:
            g.tval = float(rest[1])
    return g


def parse_control(text: str, warn=True):
    """Parse control text -> (MsCtrl, [MsBcGroup]).  Raises ControlError with
    the reference's ier code on malformed input."""
    ctrl = default_ctrl()
    groups: list[MsBcGroup] = []
    ppe_gpu = None
    for raw in text.splitlines():
        line = raw.split("#", 1)[0].strip()
        if not line:
            continue
        if "=" not in line:
            raise ControlError(2, f'control line without "=": {line}')
        key, val = line.split("=", 1)
        key, val = key.strip().lower(), val.strip()
        toks = val.split()
        try:
            if key in ("rho", "mu", "alpha_u", "alpha_p", "outer_tol", "lin_tol", "conv_blend",
                       "nonorth_corr", "dt", "cp", "k_cond", "tref", "beta", "gpu_inner_tol"):
                setattr(ctrl, key, float(toks[0].replace("d", "e").replace("D", "e")))
            elif key == "init_t":
                ctrl.init_T = float(toks[0])
            elif key == "t_pert":
                ctrl.t_pert = float(toks[0])
            elif key in ("outer_max", "lin_max", "n_time_max", "n_correct", "n_out_every",
                         "n_outer_iter", "inlet_ramp"):
                setattr(ctrl, key, int(toks[0]))
            elif key == "ppe_precond":
                ctrl.ppe_precond = 1 if val.strip().lower() == "ic0" else 0
            elif key == "transient":
                ctrl.transient = _bool(val, key)
            elif key == "pimple":
                ctrl.pimple = _bool(val, key)
            elif key == "boussinesq":
                ctrl.boussinesq = _bool(val, key)
            elif key == "time_scheme":
                ts = int(toks[0])
                if ts not in (1, 2):
                    raise ControlError(2, f"time_scheme must be 1 or 2, got {ts}")
                ctrl.time_scheme = ts
            elif key == "gravity":
                ctrl.gravity[:] = [float(t) for t in toks[:3]]
            elif key == "body_force":
                if len(toks) < 3:
                    raise ControlError(9, val)
                ctrl.body_force[:] = [float(t) for t in toks[:3]]
            elif key == "thermal_model":
                v = val.lower()
                if v not in ("lte", "ltne"):
                    raise ControlError(2, f"thermal_model must be lte or ltne: {val}")
                ctrl.thermal_model = 0 if v == "lte" else 1
            elif key == "bc":
                if len(groups) >= 32:
                    raise ControlError(3, "too many bc lines (max 32)")
                groups.append(_parse_bc(val))
            elif key == "lid":
                z, d = int(toks[0]), int(toks[1])
                coord = float(toks[2])
                vel = [float(t) for t in toks[3:6]]
                if not 1 <= d <= 3 or len(vel) < 3:
                    raise ControlError(8, val)
                ib = [i for i, g in enumerate(groups) if g.zone == z and g.btype == BC_WALL]
                if not ib:
                    raise ControlError(8, f"lid zone {z} is not a wall bc")
                g = groups[ib[-1]]
                g.has_lid, g.lid_dir, g.lid_coord = 1, d, coord
                g.lid_vel[:] = vel
            elif key == "tbc":
                z, tt = int(toks[0]), int(toks[1])
                if not 0 <= tt <= 2:
                    raise ControlError(9, val)
                ib = [i for i, g in enumerate(groups) if g.zone == z]
                if not ib:
                    raise ControlError(9, val)
                g = groups[ib[-1]]
                g.ttype = tt
                g.tval = float(toks[2]) if tt == 1 else 0.0
                g.qval = float(toks[2]) if tt == 2 else 0.0
            elif key == "tbc_plane":
                z, d, coord, tt = int(toks[0]), int(toks[1]), float(toks[2]), int(toks[3])
                ib = [i for i, g in enumerate(groups) if g.zone == z and g.btype == BC_WALL]
                if not ib or not 1 <= d <= 3 or not 0 <= tt <= 2:
                    raise ControlError(10, val)
                g = groups[ib[-1]]
                n = g.ntbc_plane
                if n >= 4:
                    raise ControlError(10, val)
                g.tbc_pdir[n], g.tbc_pcoord[n], g.tbc_pttype[n] = d, coord, tt
                g.tbc_ptval[n] = float(toks[4]) if tt == 1 else 0.0
                g.tbc_pqval[n] = float(toks[4]) if tt == 2 else 0.0
                g.ntbc_plane = n + 1
            elif key == "gpu_ppe_precond":
                ppe_gpu = PRECOND[val.lower()]
            elif key == "gpu_mom_precond":
                ctrl.gpu_mom_precond = PRECOND[val.lower()]
            elif key == "gpu_ppe_solver":
                ctrl.gpu_ppe_solver = PPE_SOLVER[val.lower()]
            elif key == "gpu_precision":
                ctrl.gpu_precision = PRECISION[val.lower()]
            elif key == "gpu_reorder":
                ctrl.gpu_reorder = {"none": 0, "rcm": 1}[val.lower()]
            elif key == "gpu_coarse":
                ctrl.gpu_coarse = int(val)
                if not 0 <= ctrl.gpu_coarse <= 1024:
                    raise ValueError(val)
            elif key == "porous_model":
                v = val.lower()
                if v in ("partitioned", "partition", "block"):
                    ctrl.porous_unified = 0
                elif v in ("unified", "one-domain", "onedomain", "single-domain"):
                    ctrl.porous_unified = 1
                else:
                    raise ControlError(2, f"porous_model must be partitioned or unified: {val}")
            elif key == "cell_zone":
                if len(ctrl.cell_zones) >= 32:
                    raise ControlError(9, "too many cell_zone lines (max 32)")
                ctrl.cell_zones.append(_parse_cell_zone(val))
            elif key == "periodic":
                k = ctrl.nperiodic
                if k >= 4 or len(toks) not in (2, 5):
                    raise ControlError(2, f"bad periodic value (zoneA zoneB [tx ty tz], max 4 lines): {val}")
                za, zb = int(toks[0]), int(toks[1])
                if za == zb:
                    raise ControlError(2, f"periodic zones must differ: {val}")
                ctrl.periodic_zone[k][0], ctrl.periodic_zone[k][1] = za, zb
                ctrl.periodic_auto[k] = 1 if len(toks) == 2 else 0
                if len(toks) == 5:
                    ctrl.periodic_shift[k][:] = [float(t) for t in toks[2:5]]
                ctrl.nperiodic = k + 1
            elif key in ("out_format", "out_fmt", "out_cells", "out_cell",
                         "restart", "restart_file", "partition_file", "dump_file",
                         "mesh_scale"):
                pass          # output / restart / zone keys: host-side, no device effect
            else:
                if warn:
                    warnings.warn(f"unknown control keyword: {key}")
        except (IndexError, ValueError, KeyError) as e:
            if isinstance(e, ControlError):
                raise
            raise ControlError(2, f"bad value for {key}: {val}") from None
    if not groups:
        raise ControlError(5, "control file defines no boundary conditions")
    pz = {ctrl.periodic_zone[k][j] for k in range(ctrl.nperiodic) for j in range(2)}
    if any(g.zone in pz for g in groups):
        raise ControlError(2, "a periodic zone also has a bc line")
    ctrl.nbc = len(groups)
    ctrl.gpu_ppe_precond = ctrl.ppe_precond if ppe_gpu is None else ppe_gpu
    return ctrl, groups


def interface_zones(mesh):
    """Face zones whose condition or user name contains 'interface' (any
    case): the coupling surfaces of zone_is_interface (mod_uns_bc.f90:
    902-918)."""
    out = []
    for z, (cond, name) in sorted(getattr(mesh, "zones", {}).items()):
        if "interface" in (str(cond).strip().lower() + " " + str(name).strip().lower()):
            out.append(int(z))
    return out


def with_interface_groups(mesh, ctrl, groups):
    """groups plus one BC_INTERFACE group per interface zone of the mesh that
    has no bc line (build_bc, mod_uns_bc.f90:282-323, gives the boundary
    faces of such zones a coupling-interface group after the control's
    groups); sets ctrl.nbc."""
    have = {int(g.zone) for g in groups}
    bnd = set(int(z) for z in np.unique(mesh.zone[mesh.c1 == 0]))
    out = list(groups)
    for z in interface_zones(mesh):
        if z in bnd and z not in have:
            g = MsBcGroup()
            g.zone, g.btype = z, BC_INTERFACE
            out.append(g)
    ctrl.nbc = len(out)
    return out


def groups_array(groups):
    arr = (MsBcGroup * max(1, len(groups)))()
    for i, g in enumerate(groups):
        arr[i] = g
    return arr
