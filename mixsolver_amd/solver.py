"""Host-side mirror of the reference's unstructured SIMPLE/PISO interface.

``Case`` owns one device system (mesh, fields, matrices resident on the
MI355X) and exposes the reference's routines under their names:

  simple_run(nsteps)        mod_uns_simple.f90:88        (serial steady driver)
  piso_run(nsteps)          mod_uns_simple_mpi.f90:298   (transient driver, np=1)
  pimple_run(nsteps)        mod_uns_simple_mpi.f90:580   (PIMPLE driver, np=1)
  compute_gradients()       mod_uns_fields.f90:164
  momentum_assembly(comp)   mod_uns_simple.f90:928
  temperature_assembly()    mod_uns_simple.f90:1748
  flux_rhiechow()           mod_uns_simple.f90:1413
  outflow_mass_rescale()    mod_uns_simple.f90:1536
  ppe_assembly()            mod_uns_simple.f90:1552 / mod_uns_simple_mpi.f90:882
  correct_fields(pp)        mod_uns_simple.f90:1649

Fields are exchanged in the reference's natural cell order and Fortran
layouts (u(3,ncells) interleaved).  Driver status codes (ier) are the
reference's: 20-23 SIMPLE breakdowns, 30-36 PISO, 60 non-finite.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, dp, ip
from .control import groups_array, parse_control

_WIDTH = dict(u=3, gp=3, gu=9, gt=3, gts=3, u_old=3, u_old_old=3, perm_dir=3, xf=3, sf=3, xc=3)


class Case:
    def __init__(self, mesh, ctrl, groups, device=0):
        L = _lib.lib()
        check(L.ms_init(device))
        self.mesh = mesh
        self.nc, self.nf = mesh.ncells, mesh.nfaces
        self.ctrl = ctrl
        from .control import with_interface_groups
        self._garr = groups_array(with_interface_groups(mesh, ctrl, groups))
        h = C.c_void_p()
        x = np.ascontiguousarray(mesh.x, np.float64)
        fptr = np.ascontiguousarray(mesh.fptr, np.int64)
        check(L.ms_system_create(C.byref(h), mesh.nnodes, dp(x), self.nf,
                                 fptr.ctypes.data_as(C.POINTER(C.c_longlong)),
                                 ip(np.ascontiguousarray(mesh.fnodes, np.int32)),
                                 ip(np.ascontiguousarray(mesh.c0, np.int32)),
                                 ip(np.ascontiguousarray(mesh.c1, np.int32)),
                                 ip(np.ascontiguousarray(mesh.zone, np.int32)),
                                 self.nc, C.byref(ctrl), self._garr))
        self.h = h
        info = np.zeros(8, np.int64)
        check(L.ms_system_info(h, info.ctypes.data_as(C.POINTER(C.c_longlong))))
        self.nnz = int(info[2])
        self.info = dict(ncells=int(info[0]), nfaces=int(info[1]), nnz=int(info[2]), ncolors=int(info[3]),
                         nlevels=int(info[4]), device_bytes=int(info[5]), nstore=int(info[6]),
                         renumbered=bool(info[7]))
        from .zones import apply_zones
        apply_zones(self, mesh, ctrl)

    def set_cell_types(self, cztype):
        """Block type of every cell (1 fluid, 2 porous, 3 solid), natural order."""
        t = None if cztype is None else np.ascontiguousarray(cztype, np.int32)
        if t is not None and t.size != self.nc:
            raise ValueError(f"cztype: expected {self.nc} values, got {t.size}")
        check(_lib.lib().ms_set_cell_types(self.h, None if t is None else ip(t)))
        self.cztype = None if t is None else t.copy()

    def set_interface(self, what, values):
        """Coupling-interface data over all faces (natural face order):
        vel (3,nf), p, T, Ts, q, slip_C, peer_vel (3,nf), T_mode (0/1); see
        ms_set_interface."""
        v = np.ascontiguousarray(values, np.float64).ravel()
        check(_lib.lib().ms_set_interface(self.h, what.encode(), dp(v), v.size))

    def set_interface_modes(self, p_dir=False, slip=False, zhang=False, T_ltne=False):
        """The coupling driver's interface switches (set_iface_*)."""
        check(_lib.lib().ms_set_interface_modes(self.h, int(p_dir), int(slip), int(zhang), int(T_ltne)))

    def nan_scan(self):
        """(ifin, field name, cell 1-based, value, 'NaN' | 'Inf') of
        nan_scan over the current fields."""
        f, c, k, i = C.c_int(0), C.c_int(0), C.c_int(0), C.c_int(0)
        v = C.c_double(0.0)
        check(_lib.lib().ms_nan_scan(self.h, C.byref(f), C.byref(c), C.byref(v), C.byref(k), C.byref(i)))
        names = ("u_x", "u_y", "u_z", "p", "T", "T_s")
        return (i.value, names[f.value] if i.value else "(none)", c.value, v.value,
                {1: "NaN", 2: "Inf"}.get(k.value, "(none)"))

    def get_ts_order(self):
        t = C.c_int(0)
        check(_lib.lib().ms_get_ts_order(self.h, C.byref(t)))
        return t.value

    def write_dump(self, path, source_mesh="", generated=None):
        """The system's u, p, T and time history as the reference's field
        dump v1 (mixsolver_amd.restart)."""
        from .restart import FIELDS, write_dump
        write_dump(path, self.nc, {k: self.get(k) for k in FIELDS}, self.get_ts_order(), source_mesh, generated)

    def read_dump(self, path):
        """Restore u, p, T, the time history and ts_order from a field dump;
        returns the reference's ier (0 ok) and the dump's source_mesh."""
        from .restart import FIELDS, read_dump
        ier, f, ts, src = read_dump(path, self.nc)
        if ier == 0:
            for k in FIELDS:
                self.set_field(k, f[k])
            self.set_ts_order(ts)
        return ier, src

    def write_vtu(self, path, binary=False):
        """The current fields as the reference's vtk_write file (see
        mixsolver_amd.vtu): ASCII Float64 or appended binary Float32."""
        from .vtu import optional_arrays, write_vtu
        with_ts, with_q = optional_arrays(getattr(self, "cztype", None), self.get("h_sf"), self.get("a_sf"))
        write_vtu(path, self.mesh, self.get("xc"), self.get("p"), self.get("u"), self.get("T"),
                  T_s=self.get("T_s") if with_ts else None, q_src=self.get("q_src") if with_q else None,
                  binary=binary)

    @classmethod
    def from_control(cls, mesh, control_text, device=0, **overrides):
        ctrl, groups = parse_control(control_text, warn=False)
        for k, v in overrides.items():
            setattr(ctrl, k, v)
        return cls(mesh, ctrl, groups, device)

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().ms_system_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # fields -----------------------------------------------------------------
    def _size(self, name):
        n = self.nf if name in ("flux", "xf", "sf", "area", "lf") else self.nc
        return n * _WIDTH.get(name, 1)

    def set_field(self, name, values):
        a = np.ascontiguousarray(np.asarray(values, np.float64).ravel())
        if a.size != self._size(name):
            raise ValueError(f"{name}: expected {self._size(name)} values, got {a.size}")
        check(_lib.lib().ms_set_field(self.h, name.encode(), dp(a), a.size))

    def get(self, name):
        n = self._size(name)
        a = np.empty(n, np.float64)
        check(_lib.lib().ms_get_field(self.h, name.encode(), dp(a), n))
        w = _WIDTH.get(name, 1)
        return a.reshape(-1, w) if w > 1 else a

    def set_ts_order(self, t):
        check(_lib.lib().ms_set_ts_order(self.h, int(t)))

    def topology(self):
        rp = np.zeros(self.nc + 1, np.int32)
        ci = np.zeros(self.nnz, np.int32)
        fg = np.zeros(self.nf, np.int32)
        check(_lib.lib().ms_get_topology(self.h, ip(rp), ip(ci), ip(fg)))
        return dict(row_ptr=rp, col_idx=ci, fgrp=fg)

    # drivers ------------------------------------------------------------------
    def simple_run(self, nsteps=0):
        itd, ier = C.c_int(0), C.c_int(0)
        check(_lib.lib().ms_simple_run(self.h, int(nsteps), C.byref(itd), C.byref(ier)))
        return ier.value, itd.value

    def piso_run(self, nsteps=0):
        ier = C.c_int(0)
        check(_lib.lib().ms_piso_run(self.h, int(nsteps), C.byref(ier)))
        return ier.value

    def pimple_run(self, nsteps=0):
        ier = C.c_int(0)
        check(_lib.lib().ms_pimple_run(self.h, int(nsteps), C.byref(ier)))
        return ier.value

    def piso_continue(self, nsteps):
        """The next nsteps steps of the last piso_run, without restarting it."""
        ier = C.c_int(0)
        check(_lib.lib().ms_piso_continue(self.h, int(nsteps), C.byref(ier)))
        return ier.value

    def pimple_continue(self, nsteps):
        """The next nsteps steps of the last pimple_run, without restarting it."""
        ier = C.c_int(0)
        check(_lib.lib().ms_pimple_continue(self.h, int(nsteps), C.byref(ier)))
        return ier.value

    def history(self):
        n = C.c_int(0)
        check(_lib.lib().ms_get_history(self.h, None, 0, C.byref(n)))
        out = np.zeros(max(1, n.value) * 8)
        check(_lib.lib().ms_get_history(self.h, dp(out), n.value, C.byref(n)))
        return out[: n.value * 8].reshape(-1, 8)

    STAT_KEYS = ("ppe_solves", "ppe_iters_fp32", "ppe_iters_fp64", "ppe_fp64_solves", "fp32_stalled",
                 "stall_window", "mom_solves", "mom_iters")

    def solver_stats(self, reset=False):
        """Solve counters since creation or the last reset: p' solves, p'
        iterations in FP32 and in FP64, p' solves that ran or finished in FP64
        on a mixed-precision system, whether the FP32 p' iteration stalled
        (sticky), the stall window in force, momentum component solves and
        their iterations."""
        out = np.zeros(8, np.int64)
        check(_lib.lib().ms_solver_stats(self.h, out.ctypes.data_as(C.POINTER(C.c_longlong)), int(reset)))
        return dict(zip(self.STAT_KEYS, (int(v) for v in out)))

    FORCE_KEYS = ("Fx", "Fy", "Fz", "CD", "CL", "u_inf", "A_ref", "p_inf", "Fx_p", "Fy_p", "Fz_p", "Fx_v", "Fy_v",
                  "Fz_v")

    def compute_forces(self):
        """Pressure and viscous force on the wall zones with the drag / lift
        coefficients of compute_forces (mod_uns_forces.f90:43-185), reduced
        on the device: dict over FORCE_KEYS."""
        out = np.zeros(14)
        check(_lib.lib().ms_compute_forces(self.h, dp(out)))
        return dict(zip(self.FORCE_KEYS, out.tolist()))

    def cp_distribution(self):
        """Rows of write_cp_distribution (mod_uns_forces.f90:195-254), one per
        wall face: (n, 6) theta deg, Cp, Cp_exact, x, y, z."""
        n = C.c_int(0)
        check(_lib.lib().ms_cp_distribution(self.h, None, 0, C.byref(n)))
        rows = np.zeros(6 * max(1, n.value))
        check(_lib.lib().ms_cp_distribution(self.h, dp(rows), n.value, C.byref(n)))
        return rows[: 6 * n.value].reshape(-1, 6)

    def separation(self):
        """compute_separation (mod_uns_forces.f90:275-407): separation angles
        (rad, from +x) of the upper and lower wall, whether found, and the
        wake length."""
        out = np.zeros(5)
        check(_lib.lib().ms_separation(self.h, dp(out)))
        return dict(theta_top=out[0], found_top=bool(out[1]), theta_bot=out[2], found_bot=bool(out[3]),
                    wake_length=out[4])

    def time_kernel(self, fp32=True, reps=50, cold=False, reduced=False):
        """(average ms, algorithmic bytes) of the PCG SpMV kernel on the
        current matrix (HIP events on the library stream); cold = each
        launch timed alone after the last-level cache was evicted; reduced =
        the red-black reduced p' matrix (gpu_ppe_solver = schur)."""
        t, b = C.c_double(0.0), C.c_double(0.0)
        which = (0 if fp32 else 1) + (2 if cold else 0) + (4 if reduced else 0)
        check(_lib.lib().ms_time_kernel(self.h, which, int(reps), C.byref(t), C.byref(b)))
        return t.value, b.value

    def coarse_info(self):
        """(k, nchunks, aggregate of every cell) of the p' coarse space."""
        k, nch = C.c_int(0), C.c_int(0)
        agg = np.zeros(self.nc, np.int32)
        check(_lib.lib().ms_coarse_info(self.h, C.byref(k), C.byref(nch), ip(agg)))
        return k.value, nch.value, agg

    # per-call routines ---------------------------------------------------------
    def compute_gradients(self):
        check(_lib.lib().ms_compute_gradients(self.h))

    def momentum_assembly(self, comp):
        val, rhs, ap = np.zeros(self.nnz), np.zeros(self.nc), np.zeros(self.nc)
        check(_lib.lib().ms_momentum_assembly(self.h, int(comp), dp(val), dp(rhs), dp(ap)))
        return val, rhs, ap

    def solid_temperature_assembly(self):
        val, rhs, ap = np.zeros(self.nnz), np.zeros(self.nc), np.zeros(self.nc)
        check(_lib.lib().ms_solid_temperature_assembly(self.h, dp(val), dp(rhs), dp(ap)))
        return val, rhs, ap

    def temperature_assembly(self):
        val, rhs, ap = np.zeros(self.nnz), np.zeros(self.nc), np.zeros(self.nc)
        check(_lib.lib().ms_temperature_assembly(self.h, dp(val), dp(rhs), dp(ap)))
        return val, rhs, ap

    def ppe_assembly(self):
        val, rhs = np.zeros(self.nnz), np.zeros(self.nc)
        check(_lib.lib().ms_ppe_assembly(self.h, dp(val), dp(rhs)))
        return val, rhs

    def flux_rhiechow(self):
        check(_lib.lib().ms_flux_rhiechow(self.h))

    def outflow_mass_rescale(self):
        check(_lib.lib().ms_outflow_mass_rescale(self.h))

    def correct_fields(self, pp):
        pp = np.ascontiguousarray(pp, np.float64)
        dpp = np.zeros(3 * self.nc)
        check(_lib.lib().ms_correct_fields(self.h, dp(pp), dp(dpp)))
        return dpp.reshape(-1, 3)


def set_graphs(on: bool) -> None:
    """Replay the mixed-precision PCG's 16-iteration chunks from a captured
    HIP graph (default) or launch their kernels one by one; results are
    identical either way."""
    check(_lib.lib().ms_set_graphs(1 if on else 0))


def set_fused_momentum(on: bool) -> None:
    """Assemble the three momentum components' interior faces in one pass
    where that is exact (default); off: one pass per component.  Results are
    identical either way."""
    check(_lib.lib().ms_set_fused_momentum(int(bool(on))))


def set_schur_matfree(on: bool) -> None:
    """FP32 iterations of the reduced p' solve apply S matrix-free (default)
    or the formed S."""
    check(_lib.lib().ms_set_schur_matfree(int(bool(on))))


def set_fused_p_update(on: bool) -> None:
    """With the matrix-free S, apply the reduced PCG's p update inside the
    next product or by a launch of its own (default); results identical."""
    check(_lib.lib().ms_set_fused_p_update(int(bool(on))))


def set_stall_window(iters: int) -> None:
    """Iterations without a halving of the FP32 residual after which the
    mixed-precision PCG hands over to FP64 (0 = the default 1024)."""
    check(_lib.lib().ms_set_stall_window(int(iters)))


def spmv_probe(on=None):
    """In-loop timing of the mixed PCG's FP32 SpMV launches: returns
    (launches, total_ms, rows) counted so far, then turns the probe on
    (True, counters zeroed) or off (False), or leaves it (None)."""
    n, ms, rows = C.c_longlong(0), C.c_double(0.0), C.c_longlong(0)
    check(_lib.lib().ms_spmv_probe(-1 if on is None else int(bool(on)), C.byref(n), C.byref(ms), C.byref(rows)))
    return n.value, ms.value, rows.value


def cell_ordering(row_ptr, col_idx, xc=None, rcm=False, colored=False, coarse=0):
    """The device numbering of a cell graph (1-based CSR pattern): returns
    (n2o, color, ncolors) with n2o[new] = old (0-based), as ms_system_create
    builds it for gpu_reorder / DILU colouring / gpu_coarse."""
    rp = np.ascontiguousarray(row_ptr, np.int32)
    ci = np.ascontiguousarray(col_idx, np.int32)
    n = len(rp) - 1
    n2o, color, nc = np.zeros(n, np.int32), np.zeros(n, np.int32), C.c_int(0)
    x = None if xc is None else dp(np.ascontiguousarray(xc, np.float64))
    check(_lib.lib().ms_cell_ordering(n, ip(rp), ip(ci), x, int(rcm), int(colored), int(coarse), ip(n2o), ip(color),
                                      C.byref(nc)))
    return n2o, color, nc.value


def coarse_aggregates(xc, k):
    """Recursive coordinate bisection of cell centres xc (n,3) into k
    aggregates of equal size (+-1): the aggregation of the p' coarse space
    (host only, no device needed)."""
    xc = np.ascontiguousarray(xc, np.float64).reshape(-1, 3)
    agg = np.zeros(len(xc), np.int32)
    check(_lib.lib().ms_coarse_aggregates(len(xc), dp(xc), int(k), ip(agg)))
    return agg


def periodic_match(mesh, control_text):
    """Face pairing of the control text's ``periodic`` lines on ``mesh``:
    partner[f] = the face joined with boundary face f (0-based, both
    directions), -1 elsewhere.  Host only (no device needed); raises on the
    errors the device system create would raise."""
    ctrl, _ = parse_control(control_text, warn=False)
    x = np.ascontiguousarray(mesh.x, np.float64)
    fptr = np.ascontiguousarray(mesh.fptr, np.int64)
    partner = np.zeros(mesh.nfaces, np.int32)
    check(_lib.lib().ms_periodic_match(mesh.nnodes, dp(x), mesh.nfaces,
                                       fptr.ctypes.data_as(C.POINTER(C.c_longlong)),
                                       ip(np.ascontiguousarray(mesh.fnodes, np.int32)),
                                       ip(np.ascontiguousarray(mesh.c0, np.int32)),
                                       ip(np.ascontiguousarray(mesh.c1, np.int32)),
                                       ip(np.ascontiguousarray(mesh.zone, np.int32)),
                                       mesh.ncells, C.byref(ctrl), ip(partner)))
    return partner
