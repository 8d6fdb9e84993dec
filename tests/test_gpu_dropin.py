"""The drop-in Fortran modules on the device, called from Fortran.

tests/dropin/dropin_check is built by tests/dropin/build.sh from this
repository's sources only (its own kinds module, the ISO_C_BINDING module and
the drop-in mod_uns_linsolver / mod_uns_linsolver_mpi), so it is the kind of
caller a maintainer's build produces, without any reference object.

* Solves: the serial drop-in (bicgstab_ilu0, cg_jacobi, cg_ic0) and the
  one-partition _mpi drop-in (bicgstab_ilu0_mpi, cg_jacobi_mpi, cg_ic0_mpi)
  on the reference's solver fixtures (tests/golden/solvers_*.npz, written by
  the reference's own solvers): it / ierr equal, res and x within the bars of
  the device solver tests (the device reduces dot products in another order).
* Loop level: ms_piso_run through the Fortran binding on the reference's
  channel PISO fixture, rel-L2 <= 1e-10 on u, p, flux, T."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, load_golden, mesh_from

pytestmark = pytest.mark.gpu
CHECK = os.path.join(ROOT, "tests", "dropin", "dropin_check")
needs_check = pytest.mark.skipif(not os.path.exists(CHECK), reason="dropin_check not built (no Fortran compiler)")


def rel_l2(a, b):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _reader(path):
    b = np.fromfile(path, dtype=np.uint8)
    o = [0]

    def take(dt, n):
        a = np.frombuffer(b, dtype=dt, count=n, offset=o[0])
        o[0] += a.nbytes
        return a
    return take


@needs_check
@pytest.mark.parametrize("fixture", ["solvers_spd", "solvers_nonsym"])
@pytest.mark.parametrize("mpi", [0, 1])
@pytest.mark.parametrize("which,key", [("cgj", "cg_jacobi"), ("cgic", "cg_ic0"), ("bicg", "bicgstab_ilu0")])
def test_dropin_solvers_match_reference(fixture, mpi, which, key, tmp_path):
    d = load_golden(fixture)
    rp, ci = d["row_ptr"].astype(np.int32), d["col_idx"].astype(np.int32)
    n, nnz = len(rp) - 1, len(ci)
    with open(tmp_path / "in.bin", "wb") as f:
        np.array([n, nnz], np.int32).tofile(f)
        rp.tofile(f)
        ci.tofile(f)
        for a in (d["val"], d["b"], d["x0"], np.array([float(d["tol"])])):
            np.ascontiguousarray(a, np.float64).tofile(f)
        np.array([int(d["maxit"])], np.int32).tofile(f)
    print(f"dropin_check solve {which} {mpi} on {fixture}", flush=True)
    r = subprocess.run([CHECK, "solve", which, str(mpi), "in.bin", "out.bin"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    take = _reader(tmp_path / "out.bin")
    it, ierr = (int(v) for v in take(np.int32, 2))
    res = float(take(np.float64, 1)[0])
    x = take(np.float64, n)
    assert ierr == int(d[key + "_ierr"])
    assert abs(it - int(d[key + "_it"])) <= 1, (it, int(d[key + "_it"]))
    ref_res = float(d[key + "_res"])
    bar = 1e-10 if key.startswith("cg") else 1e-9
    assert abs(res - ref_res) <= bar * max(ref_res, 1e-300) + 1e-13, (res, ref_res)
    assert rel_l2(x, d[key + "_x"]) <= 1e-10


@needs_check
def test_dropin_loop_level_piso(tmp_path):
    from mixsolver_amd.control import groups_array, parse_control, with_interface_groups
    d = load_golden("channel_piso_jacobi")
    m = mesh_from(d)
    ctrl, groups = parse_control(str(d["control"]), warn=False)
    garr = groups_array(with_interface_groups(m, ctrl, groups))
    nsteps = int(d["nsteps"])
    with open(tmp_path / "in.bin", "wb") as f:
        np.array([m.nnodes, m.nfaces, m.ncells, len(garr), nsteps], np.int32).tofile(f)
        np.ascontiguousarray(m.x, np.float64).tofile(f)
        np.ascontiguousarray(m.fptr, np.int64).tofile(f)
        for a in (m.fnodes, m.c0, m.c1, m.zone):
            np.ascontiguousarray(a, np.int32).tofile(f)
        f.write(bytes(ctrl))
        f.write(bytes(garr))
        for a in (d["in_u"], d["in_p"], d["in_T"]):
            np.ascontiguousarray(a, np.float64).tofile(f)
    print("dropin_check piso on channel_piso_jacobi", flush=True)
    r = subprocess.run([CHECK, "piso", "in.bin", "out.bin"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    take = _reader(tmp_path / "out.bin")
    nc, nf = m.ncells, m.nfaces
    ier = int(take(np.int32, 1)[0])
    u, p, flux, T = take(np.float64, 3 * nc), take(np.float64, nc), take(np.float64, nf), take(np.float64, nc)
    assert ier == int(d["out_ier"])
    for k, v in (("u", u), ("p", p), ("flux", flux), ("T", T)):
        assert rel_l2(v, d["out_" + k]) <= 1e-10, (k, rel_l2(v, d["out_" + k]))
