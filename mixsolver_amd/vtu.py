"""VTU output of the cell fields (the reference's vtk_write files).

``write_vtu`` calls the library's host writer ``ms_vtu_write``, a
restatement of vtk_write (mod_uns_output.f90:81-240, tetrahedral fan) with
vtu_write's inline-ASCII and appended-binary layouts (mod_vtu_writer.f90),
so the file is byte-identical to the one the reference writes for the same
mesh and fields, and the reference's case scripts (cylinder, Beavers-Joseph,
conduction checks) read it unchanged.  ``Case.write_vtu`` writes a device
system's current fields and decides the optional arrays as the reference
does: temperature_solid when LTNE exchange is active (max h_sf a_sf > 0) or a
solid zone exists, heat_source_solid when a solid zone exists.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, dp, ip


def write_vtu(path, mesh, xc, p, u, T, T_s=None, q_src=None, binary=False):
    """Write the cell fields of ``mesh`` (natural order) to ``path``."""
    nc = mesh.ncells
    x = np.ascontiguousarray(mesh.x, np.float64)
    fptr = np.ascontiguousarray(mesh.fptr, np.int64)
    xc, p, u, T = (np.ascontiguousarray(np.asarray(v, np.float64).reshape(-1)) for v in (xc, p, u, T))
    if xc.size != 3 * nc or p.size != nc or u.size != 3 * nc or T.size != nc:
        raise ValueError("write_vtu: field sizes do not match the mesh")
    ts = None if T_s is None else np.ascontiguousarray(np.asarray(T_s, np.float64).reshape(-1))
    qs = None if q_src is None else np.ascontiguousarray(np.asarray(q_src, np.float64).reshape(-1))
    for v in (ts, qs):
        if v is not None and v.size != nc:
            raise ValueError("write_vtu: field sizes do not match the mesh")
    check(_lib.lib().ms_vtu_write(str(path).encode(), 1 if binary else 0, mesh.nnodes, dp(x), mesh.nfaces,
                                  fptr.ctypes.data_as(C.POINTER(C.c_longlong)),
                                  ip(np.ascontiguousarray(mesh.fnodes, np.int32)),
                                  ip(np.ascontiguousarray(mesh.c0, np.int32)),
                                  ip(np.ascontiguousarray(mesh.c1, np.int32)), nc, dp(xc), dp(p), dp(u), dp(T),
                                  None if ts is None else dp(ts), None if qs is None else dp(qs)))


def optional_arrays(cztype, h_sf, a_sf):
    """(write temperature_solid, write heat_source_solid) as vtk_write decides."""
    has_solid = cztype is not None and bool(np.any(np.asarray(cztype) == 3))
    has_ts = bool(np.size(h_sf)) and float(np.max(np.asarray(h_sf) * np.asarray(a_sf))) > 0.0
    return has_ts or has_solid, has_solid
