"""Continuous-batching solve server (host side) and the synthetic request
workload of the serving benchmark.

``Server`` wraps the library's request server (``ms_server_*``): patterns
(one per mesh) are registered once and stay resident on the device; each
request brings matrix values in the pattern's CSR order and a right-hand side,
and is solved to an FP64 true relative residual ``tol`` by mixed-precision PCG
(FP32 inner iterations, FP64 reliable updates) with Jacobi or multicolour DILU
and an optional aggregate coarse correction.  Every admitted request is packed
into every launch; a request's result does not depend on the other requests in
flight.

``Sweep`` wraps the parameter-sweep server (``ms_sweep_*``): whole
transient PISO cases of one mesh that differ in an inlet speed, forked from
one resident ``Case`` (they share its mesh, geometry, face constants, zone
properties and matrix patterns on the device and own their fields, matrices
and solver state), admitted by device-memory budget and run concurrently,
one host thread and HIP stream per case in flight.  A case's results do not
depend on the mix.

``LaplacianPattern`` builds the pressure-correction-like SPD systems of the
workload on a jittered hex or prism mesh: face coefficients
``a_f = k_f |S_f| / |d_f|`` with log-normal face diffusivity ``k_f``,
Dirichlet boundary faces (``p = 0``, the face coefficient goes to the
diagonal), and a zero-mean Gaussian right-hand side.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib, meshgen
from ._lib import check, dp, ip

PRECOND = {"jacobi": 0, "dilu": 2}


SWEEP_STATS = ("submitted", "done", "failed", "queued", "active", "peak_active", "footprint_bytes",
               "budget_bytes", "base_device_bytes", "steps_run")
CASE_STATE = ("queued", "running", "done", "failed")


class Sweep:
    """Parameter sweep over whole transient cases forked from ``base`` (a
    ``Case`` with a transient control; its fields are the cases' initial
    fields).  ``budget_bytes`` 0: 90% of the free device memory."""

    def __init__(self, base, budget_bytes=0, max_active=4):
        self.base = base
        self.h = C.c_void_p()
        check(_lib.lib().ms_sweep_create(C.byref(self.h), base.h, int(budget_bytes), int(max_active)))
        self._cases = {}

    def submit(self, nsteps, group=0, uvel=None, speed=0.0, keep_fields=False, zone=None):
        """Queue a case of ``nsteps`` PISO steps; ``group`` (1-based index of
        the base's boundary groups, or ``zone``: the group of that boundary
        zone id) is the inlet whose speed it sets: ``uvel`` for a velocity
        inlet or far field, ``speed`` for a parabolic (mean speed) or mass
        inlet.  Returns its id."""
        if zone is not None:
            ids = [int(g.zone) for g in self.base._garr]
            if int(zone) not in ids:
                raise ValueError(f"no boundary group for zone {zone}")
            group = ids.index(int(zone)) + 1
        uv = None
        if uvel is not None:
            uv = np.ascontiguousarray(uvel, np.float64).reshape(3)
        i = C.c_longlong(0)
        check(_lib.lib().ms_sweep_submit(self.h, int(group), dp(uv) if uv is not None else None, float(speed),
                                         int(nsteps), int(bool(keep_fields)), C.byref(i)))
        self._cases[i.value] = (int(nsteps), bool(keep_fields), uv)
        return i.value

    def result(self, case_id, wait_s=0.0):
        """dict(state, steps, ier, step_ms, history, t_submit, t_admit, t_done,
        stats, u, p, T, T_s) once the case is done; dict(state) before.  A
        failed case raises MixSolverError with its message."""
        nsteps, keep, _ = self._cases[case_id]
        nc = self.base.nc
        st, steps, ier = C.c_int(0), C.c_int(0), C.c_int(0)
        step_ms, hist, times = np.zeros(nsteps), np.zeros(8 * nsteps), np.zeros(3)
        stats = np.zeros(8, np.int64)
        f = dict(u=np.zeros(3 * nc), p=np.zeros(nc), T=np.zeros(nc), T_s=np.zeros(nc)) if keep else {}
        check(_lib.lib().ms_sweep_result(
            self.h, C.c_longlong(case_id), float(wait_s), C.byref(st), C.byref(steps), C.byref(ier), dp(step_ms),
            dp(hist), dp(times), stats.ctypes.data_as(C.POINTER(C.c_longlong)),
            *(dp(f[k]) if keep else None for k in ("u", "p", "T", "T_s"))))
        out = dict(state=CASE_STATE[st.value])
        if st.value != 2:
            return out
        n = steps.value
        out.update(steps=n, ier=ier.value, step_ms=step_ms[:n], history=hist[: 8 * n].reshape(-1, 8),
                   t_submit=times[0], t_admit=times[1], t_done=times[2],
                   stats=dict(zip(("ppe_solves", "ppe_iters_fp32", "ppe_iters_fp64", "ppe_fp64_solves",
                                   "fp32_stalled", "stall_window", "mom_solves", "mom_iters"),
                                  (int(v) for v in stats))))
        if keep:
            f["u"] = f["u"].reshape(-1, 3)
            out.update(f)
        return out

    def stats(self):
        out = np.zeros(10, np.int64)
        check(_lib.lib().ms_sweep_stats(self.h, out.ctypes.data_as(C.POINTER(C.c_longlong))))
        return dict(zip(SWEEP_STATS, (int(v) for v in out)))

    def close(self):
        if getattr(self, "h", None):
            check(_lib.lib().ms_sweep_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ServerOpts(C.Structure):
    _fields_ = [("hbm_budget", C.c_longlong), ("max_active", C.c_int), ("precond", C.c_int),
                ("coarse", C.c_int), ("round_iters", C.c_int), ("delta", C.c_double)]


class Server:
    def __init__(self, precond="dilu", coarse=0, max_active=64, round_iters=16, hbm_budget=0, delta=0.1):
        L = _lib.lib()
        check(L.ms_init(0))
        o = ServerOpts(int(hbm_budget), int(max_active), PRECOND[precond], int(coarse), int(round_iters),
                       float(delta))
        self.h = C.c_void_p()
        check(L.ms_server_create(C.byref(self.h), C.byref(o)))
        self._keep = {}
        self._n = {}

    def add_pattern(self, row_ptr, col_idx, xc=None):
        """1-based CSR pattern (+ optional cell centres (n,3) for the coarse space)."""
        rp = np.ascontiguousarray(row_ptr, np.int32)
        ci = np.ascontiguousarray(col_idx, np.int32)
        n = len(rp) - 1
        xcp = None
        if xc is not None:
            xc = np.ascontiguousarray(xc, np.float64).reshape(n, 3)
            xcp = dp(xc)
        pid = C.c_int(-1)
        check(_lib.lib().ms_server_add_pattern(self.h, n, ip(rp), ip(ci), xcp, C.byref(pid)))
        self._n[pid.value] = n
        return pid.value

    def submit(self, pattern, val, b, tol=1e-10, maxit=20000):
        """Queue a solve; val (pattern CSR order) and b are copied before
        this returns."""
        val = np.ascontiguousarray(val, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        t = C.c_longlong(0)
        check(_lib.lib().ms_server_submit(self.h, int(pattern), dp(val), dp(b), float(tol), int(maxit),
                                          C.byref(t)))
        self._keep[t.value] = (None, None, self._n[pattern])
        return t.value

    def result(self, ticket, timeout_ms=-1.0):
        """dict(x, it, res, ierr, t_admit, t_done) or None if not ready."""
        n = self._keep[ticket][2]
        x = np.zeros(n)
        it, ierr, ready = C.c_int(0), C.c_int(0), C.c_int(0)
        res, ta, td = C.c_double(0), C.c_double(0), C.c_double(0)
        check(_lib.lib().ms_server_result(self.h, C.c_longlong(ticket), float(timeout_ms), dp(x), C.byref(it),
                                          C.byref(res), C.byref(ierr), C.byref(ta), C.byref(td), C.byref(ready)))
        if not ready.value:
            return None
        del self._keep[ticket]
        return dict(x=x, it=it.value, res=res.value, ierr=ierr.value, t_admit=ta.value, t_done=td.value)

    def stats(self):
        o = np.zeros(8, np.int64)
        check(_lib.lib().ms_server_stats(self.h, o.ctypes.data_as(C.POINTER(C.c_longlong))))
        keys = ("submitted", "completed", "active", "pending", "rounds", "launches", "peak_budget_bytes",
                "patterns")
        return dict(zip(keys, (int(v) for v in o)))

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().ms_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --------------------------------------------------------------- workload
def _face_geometry(m):
    """Face centroids (node mean), Newell area vectors, cell centres (mean of
    the face centroids) of a meshgen.Mesh."""
    nf = m.nfaces
    xf = np.zeros((nf, 3))
    sf = np.zeros((nf, 3))
    cnt = np.diff(m.fptr)
    for k in np.unique(cnt):
        f = np.nonzero(cnt == k)[0]
        idx = m.fptr[f][:, None] + np.arange(k)[None, :]
        P = m.x[m.fnodes[idx] - 1]                       # (nfk, k, 3)
        xf[f] = P.mean(axis=1)
        Q = np.roll(P, -1, axis=1)
        sf[f] = 0.5 * np.cross(P, Q).sum(axis=1)
    nc = m.ncells
    c0 = m.c0 - 1
    inner = m.c1 > 0
    c1 = m.c1[inner] - 1
    xc = np.zeros((nc, 3))
    w = np.bincount(c0, minlength=nc) + np.bincount(c1, minlength=nc)
    for d in range(3):
        xc[:, d] = (np.bincount(c0, xf[:, d], minlength=nc) + np.bincount(c1, xf[inner, d], minlength=nc)) / w
    return xf, sf, xc


class LaplacianPattern:
    """SPD pressure-correction-like system family on one mesh (see module doc)."""

    def __init__(self, mesh):
        self.mesh = mesh
        xf, sf, xc = _face_geometry(mesh)
        self.xc = xc
        n = mesh.ncells
        self.n = n
        area = np.linalg.norm(sf, axis=1)
        c0 = mesh.c0 - 1
        inner = mesh.c1 > 0
        c1 = np.where(inner, mesh.c1 - 1, -1)
        dvec = np.where(inner[:, None], xc[np.maximum(c1, 0)] - xc[c0], xf - xc[c0])
        self.geo = area / np.linalg.norm(dvec, axis=1)     # |S_f| / |d_f|
        # CSR pattern: diagonal + both directions of every interior face
        fi = np.nonzero(inner)[0]
        rows = np.concatenate([np.arange(n), c0[fi], c1[fi]])
        cols = np.concatenate([np.arange(n), c1[fi], c0[fi]])
        key = rows.astype(np.int64) * n + cols
        order = np.argsort(key, kind="stable")
        ks = key[order]
        self.row_ptr = (np.searchsorted(ks // n, np.arange(n + 1)) + 1).astype(np.int32)
        self.col_idx = (ks % n + 1).astype(np.int32)
        pos = np.empty_like(order)
        pos[order] = np.arange(len(order))
        self.nnz = len(ks)
        self._pos_diag = pos[:n]
        self._pos01 = pos[n:n + len(fi)]
        self._pos10 = pos[n + len(fi):]
        self._fi = fi
        self._c0 = c0
        self._inner = inner

    def values(self, rng, sigma=1.0):
        """Matrix values (pattern order) for seeded log-normal face diffusivity."""
        a = np.exp(rng.normal(0.0, sigma, self.mesh.nfaces)) * self.geo
        diag = np.bincount(self._c0, a, minlength=self.n)
        ai = a[self._fi]
        diag += np.bincount(self.mesh.c1[self._fi] - 1, ai, minlength=self.n)
        val = np.zeros(self.nnz)
        val[self._pos_diag] = diag
        val[self._pos01] = -ai
        val[self._pos10] = -ai
        return val

    def rhs(self, rng):
        b = rng.normal(0.0, 1.0, self.n)
        return b - b.mean()

    def dense(self, val):
        A = np.zeros((self.n, self.n))
        for i in range(self.n):
            A[i, self.col_idx[self.row_ptr[i] - 1:self.row_ptr[i + 1] - 1] - 1] = \
                val[self.row_ptr[i] - 1:self.row_ptr[i + 1] - 1]
        return A


def request_mesh(ncells, kind, seed, jitter=0.10):
    """Jittered unit-box hex or prism mesh of about ncells cells."""
    return meshgen.perturbed_box(int(ncells), kind, jitter=jitter, seed=seed)
