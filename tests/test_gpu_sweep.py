"""Parameter-sweep server (ms_sweep_*): whole transient porous LTNE cases
forked from one resident system, differing in the inlet speed.

Checks that a case's fields and per-step history are the same bit for bit
alone (one case in flight) and in a mix (several cases in flight on their own
streams), and the same as a system created directly with that inlet speed;
that the base system is left as it was; that the device-memory budget bounds
the cases in flight; and that the base's zone properties are locked while a
sweep exists.  The per-case physics is the partitioned transient driver at
one partition (piso_run_mpi, mod_uns_simple_mpi.f90:298-578) with the
reduced-system mixed-precision p' solve."""
import numpy as np
import pytest

from mixsolver_amd import meshgen, zones
from mixsolver_amd._lib import MixSolverError
from mixsolver_amd.control import parse_control
from mixsolver_amd.serve import Sweep
from mixsolver_amd.solver import Case

pytestmark = pytest.mark.gpu

CTL = """rho = 1.0
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
bc = 6 velocity-inlet-parabolic MEAN 2.0 2 -1.0 1.0
bc = 7 outflow
bc = 5 wall
bc = 4 symmetry
tbc = 5 1 0.5
"""
FAST_SCHUR = dict(gpu_ppe_precond=0, gpu_ppe_solver=2, gpu_mom_precond=2, gpu_coarse=16, gpu_precision=1,
                  gpu_inner_tol=0.1)
ZONE = "cell_zone = 8 porous perm=0.01 perm_yy=0.02 inertial=0.5 porosity=0.4 disp_l=0.01 disp_t=0.002 " \
       "bj_alpha=0.3 k_s=2.0 cp_s=0.8 rho_s=3.0 h_sf=5.0 a_sf=10.0\n"
NSTEPS = 3


def _mesh():
    nx, ny, nz = 40, 16, 8
    m = meshgen.hex_box(np.linspace(0.0, 4.0, nx + 1), np.linspace(-1.0, 1.0, ny + 1), np.linspace(0.0, 0.5, nz + 1),
                        zone_ids=dict(xmin=6, xmax=7, ymin=5, ymax=5, zmin=4, zmax=4))
    return meshgen.with_cell_zones(m, np.full(m.ncells, 8), {8: ("porous", "bed")})


def _case(m, mean):
    """A porous LTNE channel with a seeded porosity field in [0.4, 0.9] and
    the parabolic inlet's mean speed `mean`."""
    text = CTL.replace("MEAN", repr(float(mean))) + ZONE
    ctrl, _ = parse_control(text, warn=False)
    cz = zones.resolve_cell_zones(m, ctrl.cell_zones)
    f = zones.setup_porous_fields(m, ctrl.cell_zones, cz)
    f["porosity"] = 0.4 + 0.5 * np.random.default_rng(3).random(m.ncells)
    c2, g2 = parse_control(text, warn=False)
    c2.cell_zones = []
    for k, v in FAST_SCHUR.items():
        setattr(c2, k, v)
    cs = Case(m, c2, g2)
    cs.set_cell_types(cz)
    for k, v in f.items():
        cs.set_field(k, v)
    u0 = np.zeros((m.ncells, 3))
    u0[:, 0] = 0.05
    cs.set_field("u", u0)
    return cs, u0


SPEEDS = [float(v) for v in np.random.default_rng(11).uniform(0.01, 0.1, 6)]


def _run_sweep(base, max_active, budget=0):
    sw = Sweep(base, budget_bytes=budget, max_active=max_active)
    ids = [sw.submit(NSTEPS, zone=6, speed=s, keep_fields=True) for s in SPEEDS]
    out = [sw.result(i, wait_s=120.0) for i in ids]
    st = sw.stats()
    sw.close()
    return out, st


@pytest.fixture(scope="module")
def sweeps():
    m = _mesh()
    base, u0 = _case(m, 0.05)
    alone, st1 = _run_sweep(base, 1)
    mix, st4 = _run_sweep(base, 4)
    print("sweep stats alone", st1, "mix", st4, flush=True)
    yield m, base, u0, alone, mix, st1, st4
    base.close()


def test_sweep_case_same_alone_and_in_mix(sweeps):
    m, base, u0, alone, mix, st1, st4 = sweeps
    assert st1["peak_active"] == 1 and st4["peak_active"] > 1, (st1, st4)
    assert st1["done"] == st4["done"] == len(SPEEDS) and st1["failed"] == st4["failed"] == 0
    for a, b in zip(alone, mix):
        assert a["state"] == b["state"] == "done"
        assert a["ier"] == b["ier"] == 0 and a["steps"] == NSTEPS
        assert a["stats"]["ppe_fp64_solves"] == 0 and a["stats"]["ppe_iters_fp32"] > 0
        for k in ("u", "p", "T", "T_s"):
            assert np.isfinite(a[k]).all()
            assert np.array_equal(a[k], b[k]), k
        assert np.array_equal(a["history"], b["history"])
        assert (a["history"][:, 7] <= 1e-10).all()
    # the cases differ: the inlet speed reaches the fields
    assert not np.array_equal(alone[0]["u"], alone[1]["u"])


def test_sweep_case_equals_direct_system(sweeps):
    """A sweep case equals a system created with the case's inlet speed and
    stepped by ms_piso_run + ms_piso_continue, bit for bit; the base keeps its
    initial fields."""
    m, base, u0, alone, mix, st1, st4 = sweeps
    cs, _ = _case(m, SPEEDS[2])
    assert cs.piso_run(1) == 0
    assert cs.piso_continue(NSTEPS - 1) == 0
    for k in ("u", "p", "T", "T_s"):
        got = np.ravel(cs.get(k))
        ref = np.ravel(alone[2][k])
        assert np.array_equal(got, ref), (k, float(np.abs(got - ref).max()))
    cs.close()
    assert np.array_equal(base.get("u").reshape(-1, 3), u0)
    assert not np.any(base.get("p"))


def test_sweep_budget_bounds_cases_in_flight(sweeps):
    m, base, u0, alone, mix, st1, st4 = sweeps
    fp = st4["footprint_bytes"]
    assert fp > 0 and fp < base.info["device_bytes"] * 4
    out, st = _run_sweep(base, 8, budget=int(2.5 * fp))
    print("budget 2.5 footprints:", st, flush=True)
    assert st["peak_active"] == 2 and st["done"] == len(SPEEDS)
    for a, b in zip(alone, out):
        assert np.array_equal(a["u"], b["u"])


def test_sweep_locks_base_properties(sweeps):
    m, base, u0, alone, mix, st1, st4 = sweeps
    sw = Sweep(base, max_active=1)
    with pytest.raises(MixSolverError):
        base.set_field("porosity", np.full(m.ncells, 0.5))
    with pytest.raises(MixSolverError):
        sw.submit(1, zone=7, speed=1.0)   # an outflow group has no inlet speed
    sw.close()
    base.set_field("u", u0)   # fields stay writable once the sweep is gone
