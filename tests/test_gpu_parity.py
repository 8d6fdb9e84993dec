"""Device path against the reference (golden fixtures) and the CPU oracle.

Bars (see the module docstring of each block):
  * geometry, topology, assembly, gradients, Rhie-Chow, correction:
    bit-exact (the kernels keep the reference's per-cell accumulation order
    and evaluation order, built without FMA contraction);
  * outflow rescale: rtol 1e-14 (global sums use a tree reduction);
  * linear solvers and drivers: relative L2 <= 1e-10 (reduction order of
    the dot products is the only difference).
"""
import numpy as np
import pytest

from conftest import check_printed_history, load_golden, mesh_from
from mixsolver_amd import linsolver
from mixsolver_amd.control import parse_control
from mixsolver_amd.solver import Case
from oracle import pyoracle

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def make(d, **over):
    ctrl, groups = parse_control(str(d["control"]), warn=False)
    for k, v in over.items():
        setattr(ctrl, k, v)
    return Case(mesh_from(d), ctrl, groups)


@pytest.mark.parametrize("name", ["cavity16_simple_jacobi", "channel_piso_jacobi", "prism_piso_ic0"])
def test_geometry_topology_bitexact(name):
    d = load_golden(name)
    cs = make(d)
    for k in ("xf", "sf", "area", "vol", "xc", "lf"):
        assert np.array_equal(cs.get(k).reshape(d["geom_" + k].shape), d["geom_" + k]), k
    t = cs.topology()
    for k in ("row_ptr", "col_idx", "fgrp"):
        assert np.array_equal(t[k], d["geom_" + k]), k


@pytest.mark.parametrize("name", ["funcs_prism", "funcs_outlet", "funcs_channel"])
@pytest.mark.parametrize("reorder", [None, "dilu"])
def test_per_function_bitexact(name, reorder):
    """Every assembled value equals the reference's; with the colour-blocked
    renumbering (reorder='dilu') the values are the same, only stored in
    another row order, which the C-ABI maps back."""
    d = load_golden(name)
    over = {}
    if reorder:
        over = dict(gpu_ppe_precond=2, gpu_mom_precond=2)
    cs = make(d, **over)
    if reorder:
        assert cs.info["renumbered"]
    for k in ("u", "p", "T", "flux", "u_old", "u_old_old", "T_old", "T_old_old", "apc"):
        cs.set_field(k, d["in_" + k])
    cs.set_field("T_s", d["in_T"])
    cs.set_ts_order(int(d["in_ts_order"]))
    cs.compute_gradients()
    assert np.array_equal(cs.get("gp"), d["out_gp"])
    assert np.array_equal(cs.get("gu"), d["out_gu"])
    assert np.array_equal(cs.get("gt"), d["out_gt"])
    for comp in (1, 2, 3):
        val, rhs, ap = cs.momentum_assembly(comp)
        assert np.array_equal(val, d[f"out_mom{comp}_val"]), comp
        assert np.array_equal(rhs, d[f"out_mom{comp}_rhs"]), comp
        assert np.array_equal(ap, d[f"out_mom{comp}_ap"]), comp
    val, rhs, ap = cs.temperature_assembly()
    assert np.array_equal(val, d["out_tem_val"])
    assert np.array_equal(rhs, d["out_tem_rhs"])
    assert np.array_equal(ap, d["out_tem_ap"])
    cs.flux_rhiechow()
    assert np.array_equal(cs.get("flux"), d["out_flux_rc"])
    cs.outflow_mass_rescale()
    np.testing.assert_allclose(cs.get("flux"), d["out_flux_out"], rtol=1e-14, atol=0)
    cs.set_field("flux", d["out_flux_out"])
    dpp = cs.correct_fields(d["in_pp"])
    assert np.array_equal(dpp, d["out_cor_dpp"])
    assert np.array_equal(cs.get("u"), d["out_cor_u"])
    assert np.array_equal(cs.get("p"), d["out_cor_p"])
    assert np.array_equal(cs.get("flux"), d["out_cor_flux"])


@pytest.mark.parametrize("name", ["funcs_prism", "funcs_outlet", "funcs_channel"])
def test_ppe_assembly_bitexact_vs_oracle(name):
    """ppe_assembly is private in the reference; the oracle restatement is
    pinned through the SIMPLE/PISO fixtures, and the device PPE must equal it."""
    d = load_golden(name)
    cs = make(d)
    ctrl, groups = parse_control(str(d["control"]), warn=False)
    oc = pyoracle.OracleCase(mesh_from(d), ctrl, groups)
    for k in ("flux", "apc"):
        cs.set_field(k, d["in_" + k])
        oc.set_field(k, d["in_" + k])
    v1, r1 = cs.ppe_assembly()
    v2, r2 = oc.ppe_assembly()
    assert np.array_equal(v1, v2)
    assert np.array_equal(r1, r2)


@pytest.mark.parametrize("name", ["solvers_spd", "solvers_nonsym"])
def test_solver_histories(name):
    d = load_golden(name)
    rp, ci, val, b, x0 = d["row_ptr"], d["col_idx"], d["val"], d["b"], d["x0"]
    tol = float(d["tol"])
    fns = dict(cg_jacobi=linsolver.cg_jacobi, cg_ic0=linsolver.cg_ic0, bicgstab_ilu0=linsolver.bicgstab_ilu0)
    if name == "solvers_nonsym":
        # CG on a nonsymmetric matrix diverges (residual grows), which
        # amplifies reduction-order noise without bound; only BiCGSTAB is a
        # meaningful parity case there.
        fns = dict(bicgstab_ilu0=linsolver.bicgstab_ilu0)
    for s, fn in fns.items():
        for k in (1, 2, 3, 5, 10, 25):
            if k > int(d["kmax"]):
                continue
            x, it, res, ierr = fn(rp, ci, val, b, x0, tol, k)
            assert it == d[f"{s}_hist_it"][k - 1], (s, k)
            assert ierr == d[f"{s}_hist_ierr"][k - 1], (s, k)
            # CG: 1e-10; BiCGSTAB's rho/omega recurrences amplify the
            # reduction-order noise of the dot products, bar 1e-9
            # (plus an absolute floor: a converged FP64 residual ~1e-14 is noise)
            bar = 1e-10 if s.startswith("cg") else 1e-9
            ref = d[f"{s}_hist_res"][k - 1]
            assert abs(res - ref) <= bar * ref + 1e-13, (s, k, res, ref)
        x, it, res, ierr = fn(rp, ci, val, b, x0, tol, int(d["maxit"]))
        assert ierr == d[f"{s}_ierr"], s
        assert abs(it - int(d[f"{s}_it"])) <= 1, s
        assert rel_l2(x, d[f"{s}_x"]) <= 1e-10, s


def test_mixed_precision_reaches_fp64_tolerance():
    d = load_golden("solvers_spd")
    rp, ci, val, b = d["row_ptr"], d["col_idx"], d["val"], d["b"]
    x64, _, _, _ = linsolver.cg_jacobi(rp, ci, val, b, np.zeros_like(b), 1e-12, 5000)
    for pc in ("jacobi", "natural"):
        x, it, res, ierr = linsolver.solve(rp, ci, val, b, np.zeros_like(b), 1e-10, 5000, precond=pc,
                                           precision="mixed", inner_tol=0.1)
        assert ierr == 0 and res <= 1e-10, (pc, res, ierr)
        # true FP64 residual recomputed on the host
        n = len(b)
        A = np.zeros((n, n))
        for i in range(n):
            A[i, ci[rp[i] - 1:rp[i + 1] - 1] - 1] = val[rp[i] - 1:rp[i + 1] - 1]
        assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) <= 1e-10
        assert rel_l2(x, x64) <= 1e-9


@pytest.mark.parametrize("name,mode", [("cavity16_simple_jacobi", "simple"),
                                       ("cavity16_simple_ic0_blend", "simple"),
                                       ("outlet_simple_jacobi", "simple"),
                                       ("channel_piso_jacobi", "piso"),
                                       ("prism_piso_ic0", "piso"),
                                       ("channel_pimple_ic0", "pimple")])
def test_drivers_match_reference(name, mode):
    d = load_golden(name)
    cs = make(d)
    if "in_u" in d:
        cs.set_field("u", d["in_u"])
        cs.set_field("p", d["in_p"])
        cs.set_field("T", d["in_T"])
        cs.set_field("T_s", d["in_T"])
    if mode == "simple":
        ier, _ = cs.simple_run(int(d["nsteps"]))
    elif mode == "pimple":
        ier = cs.pimple_run(int(d["nsteps"]))
    else:
        ier = cs.piso_run(int(d["nsteps"]))
    assert ier == int(d["out_ier"])
    for k in ("u", "p", "flux", "T"):
        assert rel_l2(cs.get(k), d["out_" + k]) <= 1e-10, (k, rel_l2(cs.get(k), d["out_" + k]))
    # per-iteration / per-step diagnostics against the reference's printed
    # lines, to the printed digits (lin-it within one); the mass imbalance
    # after the corrections is rounding noise (1e-13 .. 1e-18) and is checked
    # against the solver tolerance instead
    h = cs.history()
    check_printed_history(h, d["out_log"], exact=False, skip=("imbal",))
    assert np.all(h[:, 0] <= 1e-8)


def test_piso_restart_follows_partitioned_rule():
    """ms_piso_run (piso_run_mpi) from a restored BDF2 state takes a
    first-order first step, as mod_uns_simple_mpi.f90:393-397 does; the
    oracle's partitioned driver, pinned through its pinned routines, gives
    the same state."""
    d = load_golden("channel_piso_restart")
    cs = make(d)
    oc = pyoracle.OracleCase(mesh_from(d), *parse_control(str(d["control"]), warn=False))
    for case in (cs, oc):
        for k in ("u", "p", "T", "u_old", "u_old_old", "T_old", "T_old_old"):
            case.set_field(k, d["in_" + k])
        case.set_field("T_s", d["in_T"])
    cs.set_ts_order(2)
    oc.ts_order = 2
    assert cs.piso_run(2) == 0 and oc.piso_run(2, mpi=True) == 0
    for k in ("u", "p", "flux", "T"):
        assert rel_l2(cs.get(k), oc.get(k)) <= 1e-10, (k, rel_l2(cs.get(k), oc.get(k)))


def test_piso_fast_path_converges_to_same_state():
    """Multicolour DILU + mixed precision: the p' solves reach the same 1e-10
    FP64 tolerance, so the transient state agrees with the FP64 reference
    path to the solver tolerance."""
    d = load_golden("channel_piso_jacobi")
    ref = make(d)
    fast = make(d, gpu_ppe_precond=2, gpu_mom_precond=2, gpu_precision=1, gpu_inner_tol=0.1)
    for cs in (ref, fast):
        cs.set_field("u", d["in_u"])
        cs.set_field("p", d["in_p"])
        cs.set_field("T", d["in_T"])
        assert cs.piso_run(3) == 0
    for k in ("u", "p"):
        assert rel_l2(fast.get(k), ref.get(k)) <= 1e-7, k
    h = fast.history()
    assert np.all(h[:, 7] <= 1e-10)


def _channel_case(nx, ny, nz, **over):
    from mixsolver_amd import meshgen
    ub = 15.63
    dt = 0.5 * (4 * np.pi / nx) / (1.5 * ub * 1.05)
    ctl = f"""rho = 1.0
mu = {1.0 / 180.0!r}
transient = true
dt = {dt!r}
n_time_max = 1
n_correct = 2
time_scheme = 2
lin_tol = 1e-10
lin_max = 5000
ppe_precond = jacobi
body_force = 1.0 0.0 0.0
bc = 6 velocity-inlet-parabolic {ub!r} 2.0 2 -1.0
bc = 7 outflow
bc = 5 wall
bc = 4 symmetry
"""
    m = meshgen.channel(nx, ny, nz)
    cs = Case.from_control(m, ctl, **over)
    rng = np.random.default_rng(180)
    ys = meshgen.tanh_stretch(ny, 1.1)
    yc = 0.5 * (ys[1:] + ys[:-1])
    Y = np.broadcast_to(yc[None, :, None], (nz, ny, nx)).ravel()
    u0 = np.zeros((m.ncells, 3))
    u0[:, 0] = 1.5 * ub * (1.0 - Y ** 2)
    u0 += rng.uniform(-0.05 * ub, 0.05 * ub, size=u0.shape)
    cs.set_field("u", u0)
    return cs


@pytest.mark.parametrize("pc,prec", [(2, 0), (2, 1), (0, 0), (0, 1), (1, 0)])
def test_two_level_ppe_matches_one_level(pc, prec):
    """p' PCG with the aggregate coarse correction (gpu_coarse) reaches the
    same 1e-10 FP64 residual as the one-level solve, so the PISO state agrees
    to the solver tolerance, with far fewer iterations."""
    from mixsolver_amd.solver import coarse_aggregates
    ref = _channel_case(48, 24, 12, gpu_ppe_precond=pc, gpu_mom_precond=2)
    two = _channel_case(48, 24, 12, gpu_ppe_precond=pc, gpu_mom_precond=2, gpu_precision=prec, gpu_coarse=32)
    k, nch, agg = two.coarse_info()
    assert k == 32 and nch >= k
    assert np.array_equal(agg, coarse_aggregates(two.get("xc"), 32))
    for cs in (ref, two):
        assert cs.piso_run(2) == 0
    h1, h2 = ref.history(), two.history()
    assert np.all(h1[:, 7] <= 1e-10) and np.all(h2[:, 7] <= 1e-10)
    assert h2[:, 6].sum() < 0.7 * h1[:, 6].sum(), (h1[:, 6], h2[:, 6])
    for f in ("u", "p"):
        assert rel_l2(two.get(f), ref.get(f)) <= 1e-7, f


def test_cavity_re1000_published_answer():
    """End-to-end known answer: the reference's shipped Re=1000 cavity
    (80x80x1, its control: SIMPLE, alpha 0.7/0.3, limited 2nd-order upwind,
    1000 outer iterations) solved on the device fast path (multicolour DILU,
    mixed precision, coarse space) reproduces the published max|u| 9.11512E-01
    and Ghia L2 2.89706E-02 (steady state: the p' solver only changes the
    path, bar 2e-5)."""
    from mixsolver_amd.post import ghia_compare
    d = load_golden("cavity_re1000_known_answer")
    cs = make(d, gpu_ppe_precond=2, gpu_mom_precond=2, gpu_precision=1, gpu_coarse=16)
    ier, it = cs.simple_run(int(d["nsteps"]))
    assert ier == 0
    u = cs.get("u").reshape(-1, 3)
    umax = float(np.linalg.norm(u, axis=1).max())
    l2, _ = ghia_compare(cs.get("xc").reshape(-1, 3), u, 1000)
    assert abs(umax - float(d["readme_umax"])) <= 2e-5, umax
    assert abs(l2 - float(d["readme_ghia_l2"])) <= 2e-5, l2


def _grid_matrix(nx, ny, nz, skew, seed):
    """1-based CSR of a 7-point operator on an nx x ny x nz grid in natural
    order: diagonally dominant, symmetric for skew = 0 (else the upper
    couplings are scaled by 1 - skew, the lower by 1 + skew)."""
    rng = np.random.default_rng(seed)
    n = nx * ny * nz
    idx = np.arange(n).reshape(nz, ny, nx)
    rows, cols, vals = [], [], []
    for sl_a, sl_b in (((slice(None), slice(None), slice(1, None)), (slice(None), slice(None), slice(None, -1))),
                       ((slice(None), slice(1, None), slice(None)), (slice(None), slice(None, -1), slice(None))),
                       ((slice(1, None), slice(None), slice(None)), (slice(None, -1), slice(None), slice(None)))):
        a, b = idx[sl_a].ravel(), idx[sl_b].ravel()      # a > b: b is the lower neighbour of a
        w = rng.uniform(0.5, 1.5, a.size)
        rows += [a, b]
        cols += [b, a]
        vals += [-w * (1 + skew), -w * (1 - skew)]
    rows, cols, vals = np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)
    diag = np.zeros(n)
    np.add.at(diag, rows, np.abs(vals))
    rows = np.concatenate([rows, np.arange(n)])
    cols = np.concatenate([cols, np.arange(n)])
    vals = np.concatenate([vals, diag * rng.uniform(1.01, 1.2, n)])
    o = np.lexsort((cols, rows))
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, rows + 1, 1)
    return np.cumsum(rp) + 1, cols[o] + 1, vals[o]


@pytest.mark.parametrize("kind", ["cg_ic0", "bicgstab_ilu0"])
def test_natural_sweeps_wide_and_narrow_levels(kind):
    """Natural-order IC(0) / ILU(0) on a 64^3 grid: the level schedule has
    diagonal planes of up to ~3000 rows (one launch per level) and narrow
    levels at both corners (runs of levels in one workgroup), so factor and
    sweeps take both launch forms.  Residual histories against the oracle
    (bar as the reference's solver histories: CG 1e-10, BiCGSTAB 1e-9)."""
    from oracle import pyoracle
    rp, ci, val = _grid_matrix(64, 64, 64, 0.0 if kind == "cg_ic0" else 0.3, 7)
    n = len(rp) - 1
    rng = np.random.default_rng(8)
    b, x0 = rng.normal(size=n), np.zeros(n)
    fn = dict(cg_ic0=linsolver.cg_ic0, bicgstab_ilu0=linsolver.bicgstab_ilu0)[kind]
    bar = 1e-10 if kind == "cg_ic0" else 1e-9
    for k in (1, 2, 3, 5):
        x, it, res, ierr = fn(rp, ci, val, b, x0, 1e-30, k)
        xr, itr, resr, ierrr = pyoracle.solve(kind, rp, ci, val, b, x0, 1e-30, k)
        assert (it, ierr) == (itr, ierrr), k
        assert abs(res - resr) <= bar * resr, (k, res, resr)
        assert rel_l2(x, xr) <= bar, k


@pytest.mark.parametrize("pc,coarse", [(2, 32), (2, 0), (0, 32), (1, 0)])
def test_graph_chunks_equal_eager(pc, coarse):
    """The mixed PCG's 16-iteration chunks replayed from a captured HIP graph
    give the same PISO run bit for bit as launching the kernels one by one:
    three steps (momentum solves, two p' correctors each, reliable updates
    and set-up between the replays), every field and the whole history.
    The eager run comes first in a fresh system, then the graph run, then
    eager again (the graph cache must not leak into eager runs)."""
    from mixsolver_amd.solver import set_graphs
    over = dict(gpu_ppe_precond=pc, gpu_mom_precond=2, gpu_precision=1, gpu_coarse=coarse)
    out = []
    try:
        for on in (False, True, False):
            set_graphs(on)
            cs = _channel_case(48, 24, 12, **over)
            assert cs.piso_run(3) == 0
            out.append((cs.get("u"), cs.get("p"), cs.get("flux"), cs.history()))
    finally:
        set_graphs(True)
    h = out[1][3]
    assert h[:, 6].min() > 32, h[:, 6]          # every p' solve ran full chunks
    for a, b in ((out[0], out[1]), (out[0], out[2])):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("name,drv", [("channel_pimple_ic0", "pimple"), ("channel_piso_jacobi", "piso")])
def test_restart_through_field_dump(name, drv, tmp_path):
    """A run continued from a field dump equals the uninterrupted run bit
    for bit: two steps, dump (u, p, T, time history, ts_order), a fresh
    system reads it, then two more steps on both.  The dump's 17
    significant digits restore every double exactly, and each driver call
    recomputes the fluxes and gradients on entry, as the reference does."""
    d = load_golden(name)
    a = make(d)
    for k, f in (("u", "in_u"), ("p", "in_p"), ("T", "in_T")):
        a.set_field(k, d[f])
    run = (lambda cs: cs.pimple_run(2)) if drv == "pimple" else (lambda cs: cs.piso_run(2))
    assert run(a) == 0
    a.write_dump(tmp_path / "d.txt", source_mesh="channel.cas")
    b = make(d)
    ier, src = b.read_dump(tmp_path / "d.txt")
    assert ier == 0 and src == "channel.cas"
    assert b.get_ts_order() == a.get_ts_order()
    assert run(a) == 0 and run(b) == 0
    for k in ("u", "p", "T", "flux", "u_old", "u_old_old"):
        assert np.array_equal(a.get(k), b.get(k)), k


@pytest.mark.parametrize("reorder", [None, "dilu"])
def test_nan_guard_groups_device(reorder):
    """ms_nan_scan on the device fields (caller's numbering under the
    colour-blocked renumbering): the reference NaN-guard unit test's groups;
    and simple_run returns 60 when only T_s is non-finite."""
    from conftest import nan_guard_groups
    d = load_golden("cavity16_simple_jacobi")
    over = dict(gpu_ppe_precond=2, gpu_mom_precond=2) if reorder else {}
    cs = make(d, **over)
    nan_guard_groups(cs, cs.nc)
    cs2 = make(d, **over)
    ts = np.zeros(cs2.nc)
    ts[5] = np.nan
    cs2.set_field("T_s", ts)
    assert cs2.simple_run(2)[0] == 60


def test_mixed_stall_guard_hands_over_to_fp64():
    """The mixed PCG's stall guard, forced with a 16-iteration window on the
    Jacobi-preconditioned p' system of a stretched channel (whose FP32
    residual does not halve every 16 iterations): the FP32 loop stops, its
    correction is folded into x, the FP64 PCG finishes within the remaining
    budget, and every iteration is counted once (FP32 + FP64 = the step's p'
    iterations).  The switch is sticky: the next step's solves run in FP64
    only.  The state agrees with the FP64 path to the solver tolerance."""
    from mixsolver_amd.solver import set_stall_window
    set_stall_window(16)
    try:
        cs = _channel_case(48, 24, 12, gpu_ppe_precond=0, gpu_mom_precond=2, gpu_precision=1)
        assert cs.solver_stats()["stall_window"] == 16
        assert cs.piso_run(1) == 0
        st1 = cs.solver_stats(reset=True)
        h1 = cs.history()[-1]
        print("step 1", st1, h1.tolist(), flush=True)
        assert st1["fp32_stalled"] == 1 and st1["ppe_fp64_solves"] >= 1
        assert st1["ppe_solves"] == 2
        assert st1["ppe_iters_fp32"] > 0 and st1["ppe_iters_fp64"] > 0
        assert st1["ppe_iters_fp32"] + st1["ppe_iters_fp64"] == int(h1[6])
        assert h1[7] <= 1e-10
        # the second step: FP64 only
        assert cs.piso_run(1) == 0
        st2 = cs.solver_stats()
        h2 = cs.history()[-1]
        print("step 2", st2, h2.tolist(), flush=True)
        assert st2["ppe_iters_fp32"] == 0 and st2["ppe_fp64_solves"] == st2["ppe_solves"] == 2
        assert st2["ppe_iters_fp64"] == int(h2[6]) and h2[7] <= 1e-10
    finally:
        set_stall_window(0)
    # first-step state against the FP64 path
    one = _channel_case(48, 24, 12, gpu_ppe_precond=0, gpu_mom_precond=2)
    assert one.piso_run(1) == 0
    set_stall_window(16)
    try:
        two = _channel_case(48, 24, 12, gpu_ppe_precond=0, gpu_mom_precond=2, gpu_precision=1)
        assert two.piso_run(1) == 0
    finally:
        set_stall_window(0)
    for k in ("u", "p"):
        a, b = two.get(k), one.get(k)
        if k == "p":
            a, b = a - a.mean(), b - b.mean()
        assert rel_l2(a, b) <= 1e-7, (k, rel_l2(a, b))


@pytest.mark.parametrize("mode", ["piso", "pimple"])
def test_continued_run_equals_one_run(mode):
    """ms_piso_continue / ms_pimple_continue cut a run where a caller syncs
    fields (snapshots): 1 + 2 continued steps equal 3 steps in one call, bit
    for bit, and the history rows of the two calls are the rows of the one
    run."""
    name = "channel_piso_jacobi" if mode == "piso" else "channel_pimple_ic0"
    d = load_golden(name)
    one, cut = make(d), make(d)
    for cs in (one, cut):
        cs.set_field("u", d["in_u"])
        cs.set_field("p", d["in_p"])
        cs.set_field("T", d["in_T"])
    run = getattr(one, mode + "_run")
    assert run(3) == 0
    assert getattr(cut, mode + "_run")(1) == 0
    h1 = cut.history()
    assert getattr(cut, mode + "_continue")(2) == 0
    h2 = cut.history()
    for k in ("u", "p", "flux", "T"):
        assert np.array_equal(one.get(k), cut.get(k)), k
    assert np.array_equal(np.vstack([h1, h2]), one.history())


def test_cavity_128_c1_matches_reference():
    """Config 1: the Re = 1000 cavity on 128 x 128 x 1 cells, plain FP64 steady
    SIMPLE (PCG-Jacobi p', BiCGSTAB-ILU(0) momentum) run to outer_tol = 1e-8
    on the device, against the reference's run of the same case to
    convergence: the same iteration count (within one), the printed
    diagnostics to their digits, and the converged fields."""
    import os
    from conftest import GOLDEN
    if not os.path.exists(os.path.join(GOLDEN, "cavity128_simple_c1.npz")):
        pytest.skip("fixture cavity128_simple_c1 not generated")
    d = load_golden("cavity128_simple_c1")
    cs = make(d)
    ier, it = cs.simple_run(0)
    print("C1 iterations", it, "reference", int(d["iters"]), flush=True)
    assert ier == int(d["out_ier"]) == 0
    assert abs(it - int(d["iters"])) <= 1
    for k in ("u", "p"):
        print("C1", k, rel_l2(cs.get(k), d["out_" + k]), flush=True)
        assert rel_l2(cs.get(k), d["out_" + k]) <= 1e-8, (k, rel_l2(cs.get(k), d["out_" + k]))
    check_printed_history(cs.history(), d["out_log"], exact=False, skip=("imbal",))


@pytest.mark.parametrize("name,mode", [("cavity16_simple_ic0_blend", "simple"), ("prism_piso_ic0", "piso")])
def test_rcm_numbering_matches_reference(name, mode):
    """gpu_reorder = rcm: the device numbering is the reverse Cuthill-McKee
    order of the cell graph (natural-order IC0/ILU0 then run in that order),
    and the drivers still reproduce the reference's run to the solver
    tolerance (the preconditioner's elimination order
    differs from the reference's, so only the converged state is shared:
    bar 1e-6 for SIMPLE, which stops at its outer tolerance, and 1e-8 for
    PISO, whose solves all end at 1e-10)."""
    d = load_golden(name)
    cs = make(d, gpu_reorder=1)
    if "in_u" in d:
        cs.set_field("u", d["in_u"])
        cs.set_field("p", d["in_p"])
        cs.set_field("T", d["in_T"])
        cs.set_field("T_s", d["in_T"])
    bar = 1e-6 if mode == "simple" else 1e-8
    if mode == "simple":
        ier, _ = cs.simple_run(int(d["nsteps"]))
    else:
        ier = cs.piso_run(int(d["nsteps"]))
    assert ier == int(d["out_ier"])
    for k in ("u", "p", "flux"):
        e = rel_l2(cs.get(k), d["out_" + k])
        print(name, "rcm", k, e, flush=True)
        assert e <= bar, (k, e)


def _laplacian(nx, seed):
    """1-based CSR of a 7-point finite-volume Laplacian on an nx^3 grid with
    log-normal face coefficients and Dirichlet walls (SPD), and a seeded
    right-hand side."""
    rng = np.random.default_rng(seed)
    n = nx ** 3
    idx = np.arange(n).reshape(nx, nx, nx)
    rows, cols, vals = [], [], []
    diag = np.zeros(n)
    for ax in range(3):
        a = np.take(idx, np.arange(nx - 1), axis=ax).ravel()
        b = np.take(idx, np.arange(1, nx), axis=ax).ravel()
        k = np.exp(rng.standard_normal(a.size))
        rows += [a, b]
        cols += [b, a]
        vals += [-k, -k]
        np.add.at(diag, a, k)
        np.add.at(diag, b, k)
        for face in (0, nx - 1):
            w = np.take(idx, [face], axis=ax).ravel()
            np.add.at(diag, w, np.exp(rng.standard_normal(w.size)))
    rows.append(np.arange(n))
    cols.append(np.arange(n))
    vals.append(diag)
    r, c, v = np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)
    o = np.lexsort((c, r))
    r, c, v = r[o], c[o], v[o]
    rp = np.zeros(n + 1, np.int32)
    np.add.at(rp, r + 1, 1)
    return (np.cumsum(rp).astype(np.int32) + 1), (c + 1).astype(np.int32), v, rng.standard_normal(n)


@pytest.mark.parametrize("case", ["solvers_spd", "solvers_nonsym", "laplace40"])
def test_mixed_bicgstab_reaches_fp64_tolerance(case):
    """BiCGSTAB with natural-order ILU(0) in mixed precision (FP32 inner
    solves inside FP64 refinement) reaches the FP64 tolerance, with the
    FP64 true residual recomputed on the host, in at most 3x the FP64
    BiCGSTAB's iterations."""
    import scipy.sparse as sp
    if case == "laplace40":
        rp, ci, val, b = _laplacian(40, 7)
    else:
        d = load_golden(case)
        rp, ci, val, b = d["row_ptr"], d["col_idx"], d["val"], d["b"]
    n = len(b)
    A = sp.csr_matrix((val, ci - 1, rp - 1), shape=(n, n))
    out = {}
    for prec in ("fp64", "mixed"):
        x, it, res, ierr = linsolver.solve(rp, ci, val, b, np.zeros_like(b), 1e-10, 20000, solver="bicgstab",
                                           precond="natural", precision=prec)
        out[prec] = (x, it, res, ierr)
        print(case, prec, "it", it, "res", res, "ierr", ierr, flush=True)
    x, it, res, ierr = out["mixed"]
    assert ierr == 0 and res <= 1e-10
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) <= 1e-10
    assert it <= 3 * max(out["fp64"][1], 10), (it, out["fp64"][1])
