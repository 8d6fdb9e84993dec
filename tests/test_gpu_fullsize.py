"""The full-size workloads on the device, through size-independent checks.

At 2M hex cells (periodic channel, 400x100x50, tanh-stretched), 1.2M
jittered prisms (cylinder O-grid, 600k triangles extruded 2 layers) and
16M hex cells (porous LTNE channel, 800x200x100) the oracle is too slow to
run, so each case runs one PISO step twice on the
device: the fast path (FP32 inner iterations with FP64 refinement, DILU,
coarse space) and the FP64 path with the same preconditioner.  Both p'
solves must end at an FP64 true relative residual below 1e-10 in every
corrector, the fields must stay finite, and the two paths must agree to the
accuracy the 1e-10 solves fix.
"""
import numpy as np
import pytest

from mixsolver_amd import meshgen
from mixsolver_amd.solver import Case

pytestmark = pytest.mark.gpu

FAST = dict(gpu_ppe_precond=2, gpu_mom_precond=2, gpu_coarse=512, gpu_precision=1, gpu_inner_tol=0.1)
FP64 = dict(gpu_ppe_precond=2, gpu_mom_precond=2, gpu_precision=0)

PRISM_CTL = """rho = 1.0
mu = 0.005
transient = true
dt = 0.001
n_time_max = 1
n_correct = 2
time_scheme = 2
lin_tol = 1e-10
lin_max = 20000
conv_blend = 0.0
nonorth_corr = 1.0
ppe_precond = jacobi
bc = 5 wall
bc = 4 symmetry
bc = 6 pressure-far-field 0.0 1.0 0.0 0.0
"""


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def _two_paths(mesh, ctl, u0):
    out = {}
    for name, over in (("fast", FAST), ("fp64", FP64)):
        cs = Case.from_control(mesh, ctl, **over)
        cs.set_field("u", u0)
        ier = cs.piso_run(1)
        h = cs.history()[-1]
        out[name] = (ier, cs.get("u").reshape(-1, 3), cs.get("p"), h)
        cs.close()
    return out


def _check(out, tag, bar_u, bar_p):
    (ia, ua, pa, ha), (ib, ub, pb, hb) = out["fast"], out["fp64"]
    eu, ep = rel_l2(ua, ub), rel_l2(pa - pa.mean(), pb - pb.mean())
    print(tag, "history fast", ha.tolist(), "fp64", hb.tolist(), "rel-L2 u", eu, "p", ep, flush=True)
    assert ia == 0 and ib == 0
    assert ha[7] <= 1e-10 and hb[7] <= 1e-10
    for a in (ua, pa, ub, pb):
        assert np.isfinite(a).all()
    assert eu <= bar_u and ep <= bar_p


def test_channel_2m_fast_path_equals_fp64():
    import bench
    nx, ny, nz = 400, 100, 50
    mesh = meshgen.channel(nx, ny, nz, periodic=True)
    assert mesh.ncells == 2_000_000
    out = _two_paths(mesh, bench.control_text(nx), bench.initial_field(mesh, ny, nz, nx))
    # measured: u 1.7e-10, p 4.3e-9
    _check(out, "channel 2M", 1e-8, 1e-7)


def test_prism_cylinder_1p2m_fast_path_equals_fp64():
    mesh = meshgen.cylinder_prisms(ntheta=1000, nr=300, layers=2, jitter=0.15, seed=200)
    assert mesh.ncells == 1_200_000
    u0 = np.zeros((mesh.ncells, 3))
    u0[:, 0] = 1.0
    out = _two_paths(mesh, PRISM_CTL, u0)
    # measured: u 2.5e-8, p 1.4e-8
    _check(out, "prisms 1.2M", 5e-7, 5e-7)


POROUS_CTL = """rho = 1.0
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
bc = 6 velocity-inlet-parabolic 1.0 2.0 2 -1.0 1.0
bc = 7 outflow
bc = 5 wall
bc = 4 symmetry
tbc = 5 1 0.5
cell_zone = 8 porous perm=0.01 perm_yy=0.02 inertial=0.5 porosity=0.4 disp_l=0.01 disp_t=0.002 bj_alpha=0.3 k_s=2.0 cp_s=0.8 rho_s=3.0 h_sf=5.0 a_sf=10.0
"""


def test_porous_ltne_16m_fast_path_equals_fp64():
    """16M hex cells (800x200x100) in one porous zone with a seeded porosity
    field in [0.4, 0.9] (smoothed Gaussian noise, correlation length 10
    cells) handed over per cell, Darcy-Forchheimer source, fluid and solid
    energy equations coupled by h_sf a_sf (LTNE)."""
    from scipy.ndimage import gaussian_filter
    from mixsolver_amd import zones
    from mixsolver_amd.control import parse_control
    nx, ny, nz = 800, 200, 100
    m = meshgen.hex_box(np.linspace(0.0, 8.0, nx + 1), np.linspace(-1.0, 1.0, ny + 1), np.linspace(0.0, 1.0, nz + 1),
                        zone_ids=dict(xmin=6, xmax=7, ymin=5, ymax=5, zmin=4, zmax=4))
    m = meshgen.with_cell_zones(m, np.full(m.ncells, 8), {8: ("porous", "bed")})
    assert m.ncells == 16_000_000
    ctrl, groups = parse_control(POROUS_CTL, warn=False)
    cz = zones.resolve_cell_zones(m, ctrl.cell_zones)
    f = zones.setup_porous_fields(m, ctrl.cell_zones, cz)
    g = gaussian_filter(np.random.default_rng(5).standard_normal((nz, ny, nx)), 10.0, mode="wrap").ravel()
    f["porosity"] = 0.4 + 0.5 * (g - g.min()) / (g.max() - g.min())
    ctrl.cell_zones = []
    u0 = np.zeros((m.ncells, 3))
    u0[:, 0] = 1.0
    out = {}
    for name, over in (("fast", FAST), ("fp64", FP64)):
        c2, g2 = parse_control(POROUS_CTL, warn=False)
        c2.cell_zones = []
        for k, v in over.items():
            setattr(c2, k, v)
        cs = Case(m, c2, g2)
        cs.set_cell_types(cz)
        for k, v in f.items():
            cs.set_field(k, v)
        cs.set_field("u", u0)
        ier = cs.piso_run(1)
        h = cs.history()[-1]
        out[name] = (ier, cs.get("u").reshape(-1, 3), cs.get("p"), h, cs.get("T"), cs.get("T_s"))
        cs.close()
    # measured: u 4.8e-11, p 1.1e-10, T 1.5e-10, T_s 1.1e-9
    _check({k: v[:4] for k, v in out.items()}, "porous LTNE 16M", 1e-8, 1e-8)
    for i in (4, 5):
        a, b = out["fast"][i], out["fp64"][i]
        assert np.isfinite(a).all()
        print("rel-L2", ("T", "T_s")[i - 4], rel_l2(a, b), flush=True)
        assert rel_l2(a, b) <= 1e-7
