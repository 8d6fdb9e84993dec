"""The red-black reduced p' solve (gpu_ppe_solver = schur) on the device.

Eliminating colour 0 of a two-colour cell graph leaves S = D1 - B^T D0^-1 B
on colour 1; PCG on S with the Jacobi preconditioner runs the iterates of
colour-blocked DILU-PCG on the full system.  The device solve must end every
p' system at the FP64 1e-10 residual of the FULL system (recomputed after
the back substitution), leave the state of the reference path (PCG-Jacobi
p') to the solver tolerance, and take no more iterations than DILU-PCG with
the same coarse space.
"""
import numpy as np
import pytest

from mixsolver_amd import meshgen
from mixsolver_amd._lib import MixSolverError
from mixsolver_amd.solver import Case

pytestmark = pytest.mark.gpu

UB = 15.63


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def channel(nx, ny, nz, periodic, **over):
    dt = 0.5 * (4 * np.pi / nx) / (1.5 * UB * 1.05)
    bcs = ("bc = 5 wall\nperiodic = 7 6\nperiodic = 9 8\n" if periodic else
           f"bc = 6 velocity-inlet-parabolic {UB!r} 2.0 2 -1.0\nbc = 7 outflow\nbc = 5 wall\nbc = 4 symmetry\n")
    ctl = f"""rho = 1.0
mu = {1.0 / 180.0!r}
transient = true
dt = {dt!r}
n_time_max = 2
n_correct = 2
time_scheme = 2
lin_tol = 1e-10
lin_max = 20000
ppe_precond = jacobi
body_force = 1.0 0.0 0.0
{bcs}"""
    m = meshgen.channel(nx, ny, nz, periodic=periodic)
    cs = Case.from_control(m, ctl, **over)
    rng = np.random.default_rng(180)
    ys = meshgen.tanh_stretch(ny, 1.1)
    yc = 0.5 * (ys[1:] + ys[:-1])
    Y = np.broadcast_to(yc[None, :, None], (nz, ny, nx)).ravel()
    u0 = np.zeros((m.ncells, 3))
    u0[:, 0] = 1.5 * UB * (1.0 - Y ** 2)
    u0 += rng.uniform(-0.05 * UB, 0.05 * UB, size=u0.shape)
    cs.set_field("u", u0)
    return cs


@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("precision", [0, 1])
def test_schur_ppe_reaches_fp64_state(periodic, precision):
    dims = (48, 24, 12)
    ref = channel(*dims, periodic, gpu_ppe_precond=0, gpu_mom_precond=2)
    dilu = channel(*dims, periodic, gpu_ppe_precond=2, gpu_mom_precond=2, gpu_precision=precision, gpu_coarse=32)
    sch = channel(*dims, periodic, gpu_ppe_precond=0, gpu_ppe_solver=2, gpu_mom_precond=2, gpu_precision=precision,
                  gpu_coarse=32)
    for cs in (ref, dilu, sch):
        assert cs.piso_run(2) == 0
    hd, hs = dilu.history(), sch.history()
    print("periodic", periodic, "precision", precision, "p' its DILU", hd[:, 6].tolist(), "Schur", hs[:, 6].tolist(),
          "res", hs[:, 7].tolist(), flush=True)
    assert np.all(hs[:, 7] <= 1e-10)
    assert hs[:, 6].sum() <= 1.1 * hd[:, 6].sum()
    for k in ("u", "p"):
        a, b = sch.get(k), ref.get(k)
        if k == "p" and periodic:
            a, b = a - a.mean(), b - b.mean()
        assert rel_l2(a, b) <= 1e-7, (k, rel_l2(a, b))
    st = sch.solver_stats()
    assert st["ppe_solves"] == 4
    if precision == 0:
        assert st["ppe_iters_fp32"] == 0 and st["ppe_iters_fp64"] == int(hs[:, 6].sum())
    else:
        assert st["ppe_iters_fp32"] > 0


def test_schur_rejects_three_colour_graphs():
    """A periodic ring of odd length is not two-colourable: the reduction
    refuses the system with a clear error."""
    with pytest.raises(MixSolverError, match="two-colour"):
        channel(41, 8, 4, True, gpu_ppe_precond=0, gpu_ppe_solver=2, gpu_mom_precond=2)


@pytest.mark.parametrize("periodic", [False, True])
def test_schur_matfree_equals_formed_s(periodic):
    """The FP32 iterations of the reduced solve apply S = D1 - B^T D0^-1 B
    matrix-free (two launches over A's rows) or as the formed S: both end
    every p' solve at the FP64 1e-10 residual, take about the same number of
    iterations (the FP32 rounding differs) and leave the same state to the
    solver tolerance."""
    from mixsolver_amd.solver import set_schur_matfree
    dims = (48, 24, 12)
    out = {}
    try:
        for mf in (False, True):
            set_schur_matfree(mf)
            cs = channel(*dims, periodic, gpu_ppe_precond=0, gpu_ppe_solver=2, gpu_mom_precond=2, gpu_precision=1,
                         gpu_coarse=32)
            assert cs.piso_run(2) == 0
            out[mf] = (cs.history(), cs.get("u"), cs.get("p"), cs.solver_stats(), cs.time_kernel(reduced=True))
            cs.close()
    finally:
        set_schur_matfree(True)
    (h0, u0, p0, s0, k0), (h1, u1, p1, s1, k1) = out[False], out[True]
    print("periodic", periodic, "its formed", h0[:, 6].tolist(), "matfree", h1[:, 6].tolist(), "kernel formed", k0,
          "matfree", k1, flush=True)
    assert np.all(h0[:, 7] <= 1e-10) and np.all(h1[:, 7] <= 1e-10)
    assert abs(h1[:, 6].sum() - h0[:, 6].sum()) <= 0.05 * h0[:, 6].sum()
    assert s1["ppe_fp64_solves"] == 0 and s1["ppe_iters_fp32"] > 0
    assert rel_l2(u1, u0) <= 1e-8
    if periodic:
        p0, p1 = p0 - p0.mean(), p1 - p1.mean()
    assert rel_l2(p1, p0) <= 1e-7
    assert k0[0] > 0 and k1[0] > 0 and k1[1] > 0


@pytest.mark.parametrize("periodic", [False, True])
def test_fused_p_update_equals_separate_launch(periodic):
    """The reduced PCG's p update applied inside the next matrix-free product
    gives the same iterates as its own launch: histories and fields equal bit
    for bit, including solves with reliable updates (tight inner tolerance)."""
    from mixsolver_amd.solver import set_fused_p_update
    dims = (48, 24, 12)
    out = {}
    try:
        for on in (False, True):
            set_fused_p_update(on)
            cs = channel(*dims, periodic, gpu_ppe_precond=0, gpu_ppe_solver=2, gpu_mom_precond=2, gpu_precision=1,
                         gpu_coarse=32, gpu_inner_tol=0.01)
            assert cs.piso_run(2) == 0
            out[on] = (cs.history(), cs.get("u"), cs.get("p"), cs.solver_stats())
            cs.close()
    finally:
        set_fused_p_update(False)
    (h0, u0, p0, s0), (h1, u1, p1, s1) = out[False], out[True]
    print("periodic", periodic, "its", h0[:, 6].tolist(), h1[:, 6].tolist(), s1, flush=True)
    assert np.array_equal(h0, h1)
    assert np.array_equal(u0, u1) and np.array_equal(p0, p1)
    assert s1["ppe_iters_fp32"] > 0 and s1["ppe_fp64_solves"] == 0
