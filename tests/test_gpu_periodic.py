"""Periodic boundaries on the device path.

The reference has no periodic boundary, so there is no oracle; these tests
pin the joined faces by what the mathematics fixes:
  * a periodic face is an interior face whose far cell sits one period away:
    an n-cell periodic box and the 2n-cell box made of two exact copies of it
    (the copies meet at interior faces) give the same fields after transient
    PISO steps on a jittered, non-orthogonal mesh with the limited
    deferred-correction convection (every shifted-geometry site is used);
  * translation invariance: on a mesh uniform in x and z, rolling the initial
    field by whole cells rolls the result;
  * laminar Poiseuille flow driven by a body force converges to the exact
    solution of the 1-D finite-volume problem across the channel (diffusion
    between cell centres, half-cell wall distance), and to the analytic
    parabola as the mesh is refined;
  * the mixed-precision DILU + coarse-space solvers reach the FP64 Jacobi
    result on the periodic channel of the benchmark (its reduced size).
"""
import numpy as np
import pytest

from mixsolver_amd import meshgen
from mixsolver_amd.solver import Case

pytestmark = pytest.mark.gpu

PER = "periodic = 7 6\nperiodic = 9 8\n"


def piso_ctl(extra=""):
    return f"""rho = 1.0
mu = 0.05
transient = true
dt = 0.02
n_time_max = 3
n_correct = 2
time_scheme = 2
lin_tol = 1e-13
lin_max = 5000
conv_blend = 0.5
nonorth_corr = 1.0
ppe_precond = jacobi
gpu_mom_precond = jacobi
body_force = 1.0 0.0 0.0
bc = 5 wall
{PER}{extra}"""


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300))


def test_periodic_seam_equals_interior_face():
    nx, ny, nz, lx = 4, 6, 3, 2.0
    ys = meshgen.tanh_stretch(ny, 1.1)
    zs = np.linspace(0.0, 1.5, nz + 1)
    zid = dict(xmin=6, xmax=7, ymin=5, ymax=5, zmin=8, zmax=9)
    one = meshgen.hex_box(np.linspace(0.0, lx, nx + 1), ys, zs, zone_ids=zid, jitter=0.15, seed=7,
                          periodic=("x", "z"), tile_x=nx)
    two = meshgen.hex_box(np.linspace(0.0, 2 * lx, 2 * nx + 1), ys, zs, zone_ids=zid, jitter=0.15, seed=7,
                          periodic=("x", "z"), tile_x=nx)
    rng = np.random.default_rng(11)
    u1 = rng.uniform(-0.5, 0.5, size=(nz, ny, nx, 3))
    u1[..., 0] += 1.0
    u2 = np.concatenate([u1, u1], axis=2)
    res = []
    for m, u in ((one, u1), (two, u2)):
        cs = Case.from_control(m, piso_ctl())
        cs.set_field("u", u.reshape(-1, 3))
        assert cs.piso_run(3) == 0
        res.append((cs.get("u").reshape(nz, ny, -1, 3), cs.get("p").reshape(nz, ny, -1)))
    (ua, pa), (ub, pb) = res
    assert np.abs(ua - u1).max() > 1e-3          # the flow evolved
    for half in (slice(0, nx), slice(nx, 2 * nx)):
        assert rel(ub[:, :, half], ua) < 1e-9
        assert np.abs(pb[:, :, half] - pa).max() < 1e-9 * np.abs(pa).max()


def test_periodic_translation_invariance():
    nx, ny, nz = 6, 5, 4
    m = meshgen.channel(nx, ny, nz, lx=3.0, lz=2.0, periodic=True)
    rng = np.random.default_rng(5)
    u0 = rng.uniform(-0.5, 0.5, size=(nz, ny, nx, 3))
    u0[..., 0] += 1.0
    out = []
    for sx, sz in ((0, 0), (2, 1)):
        cs = Case.from_control(m, piso_ctl())
        cs.set_field("u", np.roll(u0, (sz, sx), axis=(0, 2)).reshape(-1, 3))
        assert cs.piso_run(3) == 0
        out.append(cs.get("u").reshape(nz, ny, nx, 3))
    assert rel(out[1], np.roll(out[0], (1, 2), axis=(0, 2))) < 1e-9


def test_periodic_shadow_flux_mirrors_joined_face():
    m = meshgen.channel(5, 4, 3, lx=2.0, lz=1.0, periodic=True, jitter=0.1, seed=2)
    cs = Case.from_control(m, piso_ctl())
    rng = np.random.default_rng(1)
    cs.set_field("u", rng.uniform(0.5, 1.5, size=(m.ncells, 3)))
    assert cs.piso_run(2) == 0
    F = cs.get("flux")
    z = m.zone
    for za, zb in ((7, 6), (9, 8)):
        fa, fb = np.nonzero(z == za)[0], np.nonzero(z == zb)[0]
        assert np.abs(F[fa]).max() > 0
        assert np.array_equal(F[fb], -F[fa])
    # every cell balances its fluxes, the periodic ones included
    c0, c1 = m.c0 - 1, m.c1 - 1
    imb = np.bincount(c0, F, minlength=m.ncells)
    inner = c1 >= 0
    imb -= np.bincount(c1[inner], F[inner], minlength=m.ncells)
    assert np.abs(imb).max() < 1e-8 * np.abs(F).max()


def fv_poiseuille(yc, yw, mu, f):
    """Exact solution of the 1-D finite-volume problem across the channel:
    mu * sum over faces of (u_P - u_N) / |y_N - y_P| = f * dy_P, wall value 0
    at half a cell."""
    n = len(yc)
    A = np.zeros((n, n))
    b = f * np.diff(yw)
    for j in range(n):
        for nb in (j - 1, j + 1):
            if 0 <= nb < n:
                d = mu / abs(yc[nb] - yc[j])
                A[j, j] += d
                A[j, nb] -= d
            else:
                A[j, j] += mu / abs(yc[j] - (yw[0] if nb < 0 else yw[-1]))
    return np.linalg.solve(A, b)


@pytest.mark.parametrize("ny", [8, 16])
def test_periodic_poiseuille(ny):
    mu = 0.1
    m = meshgen.channel(4, ny, 3, lx=1.0, lz=1.0, periodic=True)
    ctl = f"""rho = 1.0
mu = {mu!r}
alpha_u = 0.7
alpha_p = 0.3
outer_max = 3000
outer_tol = 1e-12
lin_tol = 1e-13
lin_max = 2000
ppe_precond = jacobi
gpu_mom_precond = jacobi
body_force = 1.0 0.0 0.0
bc = 5 wall
{PER}"""
    cs = Case.from_control(m, ctl)
    ier, it = cs.simple_run()
    assert ier == 0 and it < 3000
    u = cs.get("u").reshape(3, ny, 4, 3)
    yw = meshgen.tanh_stretch(ny, 1.1)
    yc = cs.get("xc").reshape(3, ny, 4, 3)[0, :, 0, 1]
    ufv = fv_poiseuille(yc, yw, mu, 1.0)
    ux = u[..., 0]
    assert np.abs(ux - ufv[None, :, None]).max() < 1e-8 * ufv.max()
    assert np.abs(u[..., 1:]).max() < 1e-8 * ufv.max()
    exact = (1.0 - yc ** 2) / (2.0 * mu)
    err = np.abs(ufv - exact).max() / exact.max()
    assert err < (0.02 if ny == 8 else 0.006)


def test_periodic_mixed_dilu_coarse_matches_fp64():
    nx, ny, nz = 16, 12, 8
    m = meshgen.channel(nx, ny, nz, lx=4 * np.pi, lz=2 * np.pi, periodic=True)
    rng = np.random.default_rng(180)
    yc = 0.5 * (meshgen.tanh_stretch(ny, 1.1)[1:] + meshgen.tanh_stretch(ny, 1.1)[:-1])
    u0 = np.zeros((nz, ny, nx, 3))
    u0[..., 0] = 1.5 * 15.63 * (1.0 - yc[None, :, None] ** 2)
    u0 += rng.uniform(-0.05 * 15.63, 0.05 * 15.63, size=u0.shape)
    ctl = piso_ctl().replace("mu = 0.05", f"mu = {1 / 180!r}").replace("conv_blend = 0.5", "conv_blend = 0.0")
    ctl = ctl.replace("lin_tol = 1e-13", "lin_tol = 1e-10")
    out = []
    for over in (dict(), dict(gpu_ppe_precond=2, gpu_mom_precond=2, gpu_precision=1, gpu_coarse=16)):
        cs = Case.from_control(m, ctl, **over)
        cs.set_field("u", u0.reshape(-1, 3))
        assert cs.piso_run(2) == 0
        out.append(cs.get("u"))
        assert cs.history()[:, 7].max() <= 1e-10 * 1.0001   # p' final relative residual
    assert rel(out[1], out[0]) < 1e-7
