"""Device linear solvers behind the reference's solver signatures.

cg_jacobi / cg_ic0 / bicgstab_ilu0 take a 1-based CSR (csr_t: row_ptr,
col_idx sorted per row, val), the right-hand side b and the initial guess x0,
and return (x, it, res, ierr) with the meaning of mod_uns_linsolver.f90:148,
:230, :405 (ierr 0 converged, 1 maxit, 2 breakdown, 11/12 factorization).
``solve`` exposes the extended device options (PCG/BiCGSTAB, Jacobi /
natural-order factorization, FP64 or FP32-inner mixed precision).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, dp, ip

PC = dict(jacobi=0, natural=1, ic0=1, ilu0=1, dilu=2)


def _args(row_ptr, col_idx, val, b, x0):
    rp = np.ascontiguousarray(row_ptr, np.int32)
    ci = np.ascontiguousarray(col_idx, np.int32)
    v = np.ascontiguousarray(val, np.float64)
    bb = np.ascontiguousarray(b, np.float64)
    x = np.array(x0, np.float64, copy=True)
    return rp, ci, v, bb, x


def _call(fn, row_ptr, col_idx, val, b, x0, tol, maxit):
    rp, ci, v, bb, x = _args(row_ptr, col_idx, val, b, x0)
    it, res, ierr = C.c_int(0), C.c_double(0.0), C.c_int(0)
    check(getattr(_lib.lib(), fn)(len(bb), ip(rp), ip(ci), dp(v), dp(bb), dp(x), float(tol), int(maxit),
                                  C.byref(it), C.byref(res), C.byref(ierr)))
    return x, it.value, res.value, ierr.value


def cg_jacobi(row_ptr, col_idx, val, b, x0, tol, maxit):
    return _call("ms_cg_jacobi", row_ptr, col_idx, val, b, x0, tol, maxit)


def cg_jacobi_noalloc(row_ptr, col_idx, val, b, x0, tol, maxit):
    """cg_jacobi through the entry point that allocates nothing: the system
    is staged in the library's fixed per-GPU scratch array (at most
    131072 rows and 1048576 nonzeros; larger systems raise)."""
    return _call("ms_cg_jacobi_noalloc", row_ptr, col_idx, val, b, x0, tol, maxit)


def cg_ic0(row_ptr, col_idx, val, b, x0, tol, maxit):
    return _call("ms_cg_ic0", row_ptr, col_idx, val, b, x0, tol, maxit)


def bicgstab_ilu0(row_ptr, col_idx, val, b, x0, tol, maxit):
    return _call("ms_bicgstab_ilu0", row_ptr, col_idx, val, b, x0, tol, maxit)


def solve(row_ptr, col_idx, val, b, x0, tol, maxit, solver="pcg", precond="jacobi", precision="fp64",
          inner_tol=1e-5):
    rp, ci, v, bb, x = _args(row_ptr, col_idx, val, b, x0)
    it, res, ierr = C.c_int(0), C.c_double(0.0), C.c_int(0)
    check(_lib.lib().ms_solve(len(bb), ip(rp), ip(ci), dp(v), dp(bb), dp(x), float(tol), int(maxit),
                              0 if solver == "pcg" else 1, PC[precond], 0 if precision == "fp64" else 1,
                              float(inner_tol), C.byref(it), C.byref(res), C.byref(ierr)))
    return x, it.value, res.value, ierr.value
