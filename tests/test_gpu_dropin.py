"""The drop-in module on the reference's own drivers.

oracle/_ref/ref_dropin is the reference's serial module chain built with its
mod_uns_linsolver.f90 replaced by mixsolver_amd/fortran/mod_uns_linsolver.f90
(oracle/build_ref.sh), so the reference's simple_run / piso_run call the
device solvers through the C-ABI exactly as a maintainer's build would.  The
runs must reproduce the reference's own results (committed fixtures) to the
drivers' bar, rel-L2 <= 1e-10 (the device solvers reduce dot products in
another order).  Skipped where the binary was not built (no reference
sources when the tree was built)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import ROOT, load_golden, mesh_from

pytestmark = pytest.mark.gpu
DROPIN = os.path.join(ROOT, "oracle", "_ref", "ref_dropin")


def rel_l2(a, b):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.skipif(not os.path.exists(DROPIN), reason="drop-in binary not built (needs the reference sources)")
@pytest.mark.parametrize("name,mode,zones", [
    ("cavity16_simple_jacobi", "simple", {4: "symmetry", 5: "wall"}),
    ("cavity16_simple_ic0_blend", "simple", {4: "symmetry", 5: "wall"}),
    ("channel_piso_jacobi", "piso", {4: "symmetry", 5: "wall", 6: "velocity-inlet", 7: "outflow"}),
])
def test_reference_drivers_with_device_solvers(name, mode, zones):
    from mixsolver_amd import cas
    d = load_golden(name)
    m = mesh_from(d)
    nc, nf = m.ncells, m.nfaces
    with tempfile.TemporaryDirectory() as td:
        cas.write_cas(os.path.join(td, "m.cas"), m, zones)
        open(os.path.join(td, "c.control"), "w").write(str(d["control"]))
        inf = ""
        if "in_u" in d:
            inf = os.path.join(td, "in.bin")
            with open(inf, "wb") as f:
                for a in (d["in_u"], d["in_p"], d["in_T"]):
                    np.ascontiguousarray(a, np.float64).tofile(f)
        r = subprocess.run([DROPIN, "m.cas", "c.control", mode, "out.bin", inf, str(int(d["nsteps"]))], cwd=td,
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        b = np.fromfile(os.path.join(td, "out.bin"), dtype=np.uint8)
    o = 0

    def take(dt, n):
        nonlocal o
        a = np.frombuffer(b, dtype=dt, count=n, offset=o)
        o += a.nbytes
        return a
    ier = int(take(np.int32, 1)[0])
    u, p, flux, T = take(np.float64, 3 * nc), take(np.float64, nc), take(np.float64, nf), take(np.float64, nc)
    assert ier == int(d["out_ier"])
    for k, v in (("u", u), ("p", p), ("flux", flux), ("T", T)):
        assert rel_l2(v, d["out_" + k]) <= 1e-10, (k, rel_l2(v, d["out_" + k]))
