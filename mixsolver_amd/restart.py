"""Field dump v1 (the reference's restart file), host side.

``write_dump`` / ``read_dump`` call the library's ``ms_dump_write`` /
``ms_dump_read``, restatements of write_field_dump / read_field_dump
(mod_uns_restart.f90:77-152, :166-334): an ASCII header (version, source
mesh, ncells, ts_order, generation time, field list), then u, p, T, u_old,
u_old_old, T_old, T_old_old, one cell per line in Fortran ES24.16, read back
list-directed with the reference's error codes (0 ok, 1 cannot open,
2 header, 3 version, 4 ncells, 5 data).  ``Case.write_dump`` /
``Case.read_dump`` move a device system's state through such a file; a
restored ts_order of 2 keeps the first PIMPLE step second order, as in the
reference (piso_run_mpi takes a first-order first step either way).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, dp

FIELDS = ("u", "p", "T", "u_old", "u_old_old", "T_old", "T_old_old")
_W = dict(u=3, u_old=3, u_old_old=3)


def write_dump(path, ncells, fields, ts_order=0, source_mesh="", generated=None):
    """fields: dict with the seven arrays (natural cell order)."""
    arrs = []
    for k in FIELDS:
        a = np.ascontiguousarray(np.asarray(fields[k], np.float64).reshape(-1))
        if a.size != _W.get(k, 1) * ncells:
            raise ValueError(f"write_dump: {k} has {a.size} values, expected {_W.get(k, 1) * ncells}")
        arrs.append(a)
    check(_lib.lib().ms_dump_write(str(path).encode(), int(ncells), int(ts_order), source_mesh.encode(),
                                   None if generated is None else generated.encode(), *[dp(a) for a in arrs]))


def read_dump(path, ncells):
    """(ier, fields dict, ts_order, source_mesh); fields are None unless ier == 0."""
    arrs = {k: np.zeros(_W.get(k, 1) * ncells) for k in FIELDS}
    ts, ier = C.c_int(0), C.c_int(0)
    src = C.create_string_buffer(1024)
    check(_lib.lib().ms_dump_read(str(path).encode(), int(ncells), *[dp(arrs[k]) for k in FIELDS], C.byref(ts), src,
                                  1024, C.byref(ier)))
    if ier.value != 0:
        return ier.value, None, 0, ""
    for k in _W:
        arrs[k] = arrs[k].reshape(ncells, 3)
    return 0, arrs, ts.value, src.value.decode()
