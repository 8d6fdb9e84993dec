"""The CPU oracle against the reference's own FP64 Fortran outputs.

Every fixture under tests/golden was written by the reference code itself
(tests/golden/make_golden.py drives the reference modules through
oracle/ref_harness.f90).  The oracle restates the same arithmetic in the same
order, so every comparison here is bit-exact (np.array_equal).
"""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import IFACE_CASES, POROUS_CASES, check_printed_history, load_golden, mesh_from, run_funcs2, \
    seed_interface, iface_mesh
from mixsolver_amd.control import parse_control
from oracle import pyoracle


def make_case(d):
    ctrl, groups = parse_control(str(d["control"]))
    return pyoracle.OracleCase(mesh_from(d), ctrl, groups)


@pytest.mark.parametrize("name", ["cavity16_simple_jacobi", "channel_piso_jacobi", "prism_piso_ic0"])
def test_geometry_topology_bitexact(name):
    d = load_golden(name)
    oc = make_case(d)
    g, t = oc.geometry(), oc.topology()
    for k in ("xf", "sf", "area", "vol", "xc", "lf"):
        assert np.array_equal(g[k], d["geom_" + k]), k
    for k in ("cf_ptr", "cf", "row_ptr", "col_idx", "fgrp"):
        assert np.array_equal(t[k], d["geom_" + k]), k


@pytest.mark.parametrize("name,mode", [("cavity16_simple_jacobi", "simple"),
                                       ("cavity16_simple_ic0_blend", "simple"),
                                       ("outlet_simple_jacobi", "simple"),
                                       ("channel_piso_jacobi", "piso"),
                                       ("prism_piso_ic0", "piso"),
                                       ("channel_pimple_ic0", "pimple")])
def test_driver_bitexact(name, mode):
    d = load_golden(name)
    oc = make_case(d)
    if "in_u" in d:
        oc.set_field("u", d["in_u"])
        oc.set_field("p", d["in_p"])
        oc.set_field("T", d["in_T"])
        oc.set_field("T_s", d["in_T"])
    if mode == "simple":
        ier, _ = oc.simple_run(int(d["nsteps"]), mpi=False)
    elif mode == "pimple":
        ier = oc.pimple_run(int(d["nsteps"]), mpi=False)
    else:
        ier = oc.piso_run(int(d["nsteps"]), mpi=False)
    assert ier == int(d["out_ier"])
    for k in ("u", "p", "flux", "T", "gp", "gu", "apc", "gt"):
        assert np.array_equal(oc.get(k).reshape(d["out_" + k].shape), d["out_" + k]), k
    # the per-iteration / per-step diagnostics the reference printed, digit for digit
    assert check_printed_history(oc.history(), d["out_log"], exact=True) >= 1


def _seed_restart(case, d):
    for k in ("u", "p", "T", "u_old", "u_old_old", "T_old", "T_old_old"):
        case.set_field(k, d["in_" + k])
    case.set_field("T_s", d["in_T"])
    case.set_field("T_s_old", d["in_T_old"])
    case.set_field("T_s_old_old", d["in_T_old_old"])


def test_piso_restart_keeps_bdf2_serial_and_bdf1_partitioned():
    """A restored BDF2 state (ts_order 2): the serial piso_run keeps the
    first step second order (mod_uns_simple.f90:380-414), bit-exact against
    the reference run of that state; piso_run_mpi takes a first-order first
    step either way (mod_uns_simple_mpi.f90:393-397), so its first step
    equals a cold first step of the same state, where the history levels are
    not read."""
    d = load_golden("channel_piso_restart")
    oc = make_case(d)
    _seed_restart(oc, d)
    oc.ts_order = 2
    assert oc.piso_run(int(d["nsteps"]), mpi=False) == int(d["out_ier"])
    for k in ("u", "p", "flux", "T", "gp", "gu", "apc", "gt"):
        assert np.array_equal(oc.get(k).reshape(d["out_" + k].shape), d["out_" + k]), k
    check_printed_history(oc.history(), d["out_log"], exact=True)
    # the partitioned driver: restored and cold first steps are identical,
    # and differ from the serial restored step
    a = make_case(d)
    _seed_restart(a, d)
    a.ts_order = 2
    assert a.piso_run(1, mpi=True) == 0
    b = make_case(d)
    for k in ("u", "p", "T"):
        b.set_field(k, d["in_" + k])
    b.set_field("T_s", d["in_T"])
    assert b.piso_run(1, mpi=True) == 0
    c = make_case(d)
    _seed_restart(c, d)
    c.ts_order = 2
    assert c.piso_run(1, mpi=False) == 0
    for k in ("u", "p", "flux", "T"):
        assert np.array_equal(a.get(k), b.get(k)), k
    assert not np.array_equal(a.get("u"), c.get("u"))


@pytest.mark.parametrize("name", IFACE_CASES)
def test_ppe_assembly_mpi_reduces_to_pinned_serial(name):
    """ppe_assembly_mpi (mod_uns_simple_mpi.f90:882-968) tests the pressure-
    Dirichlet type of a face only (POUTLET / FARFIELD), where the serial
    ppe_assembly (mod_uns_simple.f90:1552-1644) also counts interface faces
    under the interface pressure model (bc_face_p_dirichlet,
    mod_uns_bc.f90:879-893).  So the partitioned matrix with the model on
    equals the serial matrix with the model off (interface faces Neumann,
    cell 1 pinned when nothing else is Dirichlet), bit for bit; the serial
    assembly is pinned by the reference's SIMPLE runs of these cases."""
    d = load_golden(name)
    ctrl, groups = parse_control(str(d["control"]), warn=False)
    cases = []
    for p_dir, mpi in ((None, True), (False, False), (None, False)):
        oc = pyoracle.OracleCase(iface_mesh(d), ctrl, groups)
        seed_interface(oc, d)
        for k in ("u", "p", "T", "flux", "apc"):
            oc.set_field(k, d["ts0_in_" + k])
        if p_dir is not None:
            sw = [bool(v) for v in d["switches"]]
            oc.set_interface_modes(p_dir=p_dir, slip=sw[1], zhang=sw[2], T_ltne=sw[3])
        cases.append(oc.ppe_assembly(mpi=mpi))
    (vm, rm), (vs, rs), (v0, r0) = cases
    assert np.array_equal(vm, vs) and np.array_equal(rm, rs)
    if bool(d["switches"][0]):       # the model on: the serial variant differs
        assert not np.array_equal(vm, v0)
    else:
        assert np.array_equal(vm, v0) and np.array_equal(rm, r0)


@pytest.mark.parametrize("name", ["solvers_spd", "solvers_nonsym"])
@pytest.mark.parametrize("kind", ["cg_jacobi", "cg_ic0", "bicgstab_ilu0"])
def test_partitioned_solvers_follow_serial_histories(name, kind):
    """The _mpi solvers at one partition (mod_uns_linsolver_mpi.f90:155-363)
    run the serial recurrences with sqrt(dot_product) in place of norm2 and a
    breakdown test on the IC0 diagonal: their residual histories equal the
    reference's serial histories (maxit = k re-calls) to rounding, with the
    same iteration counts and codes."""
    d = load_golden(name)
    key = kind
    kmax = int(d["kmax"])
    for k in range(1, kmax + 1):
        x, it, res, ierr = pyoracle.solve(kind, d["row_ptr"], d["col_idx"], d["val"], d["b"], d["x0"],
                                          float(d["tol"]), k, mpi=True)
        ref = d[f"{key}_hist_res"][k - 1]
        assert abs(res - ref) <= 1e-12 * abs(ref) + 1e-300, (k, res, ref)
        assert it == int(d[f"{key}_hist_it"][k - 1]) and ierr == int(d[f"{key}_hist_ierr"][k - 1]), k
    x, it, res, ierr = pyoracle.solve(kind, d["row_ptr"], d["col_idx"], d["val"], d["b"], d["x0"],
                                      float(d["tol"]), int(d["maxit"]), mpi=True)
    assert it == int(d[f"{key}_it"]) and ierr == int(d[f"{key}_ierr"])
    assert np.linalg.norm(x - d[f"{key}_x"]) <= 1e-12 * np.linalg.norm(d[f"{key}_x"])


def _piso_mpi_composed(oc, ctrl, nsteps):
    """piso_run_mpi at one partition (mod_uns_simple_mpi.f90:376-497) written
    out as the sequence of the per-function routines it calls, each pinned
    bit for bit against the reference (momentum_assembly, the _mpi solvers,
    compute_gradients, flux_rhiechow, outflow_mass_rescale, ppe_assembly_mpi,
    correct_fields, temperature_assembly): apc = 1, gradients and Rhie-Chow
    once; per step the history shift, a first-order first step (:393-397),
    momentum solves with apc = the last ap, gradients, Rhie-Chow, outflow
    rescale, n_correct PPE solves and corrections, the temperature solve, and
    the solid phase under LTNE only (:483-495).  The driver has no
    zero_solid_velocity call."""
    t = oc.topology()
    rp, ci = t["row_ptr"], t["col_idx"]
    tol, mx = ctrl.lin_tol, ctrl.lin_max
    oc.set_field("apc", np.ones(oc.nc))
    oc.compute_gradients()
    oc.flux_rhiechow()
    for it in range(1, nsteps + 1):
        for a, b in (("u_old_old", "u_old"), ("u_old", "u"), ("T_old_old", "T_old"), ("T_old", "T"),
                     ("T_s_old_old", "T_s_old"), ("T_s_old", "T_s")):
            oc.set_field(a, oc.get(b))
        oc.ts_order = 1 if (it == 1 or ctrl.time_scheme == 1) else 2
        for comp in (1, 2, 3):
            val, rhs, ap = oc.momentum_assembly(comp)
            u = oc.get("u").reshape(-1, 3).copy()
            x, _, _, ierr = pyoracle.solve("bicgstab_ilu0", rp, ci, val, rhs, u[:, comp - 1], tol, mx, mpi=True)
            assert ierr != 2
            u[:, comp - 1] = x
            oc.set_field("u", u)
        oc.set_field("apc", ap)
        oc.compute_gradients()
        oc.flux_rhiechow()
        oc.outflow_mass_rescale()
        for _ in range(ctrl.n_correct):
            val, rhs = oc.ppe_assembly(mpi=True)
            kind = "cg_ic0" if ctrl.ppe_precond == 1 else "cg_jacobi"
            pp, _, _, ierr = pyoracle.solve(kind, rp, ci, val, rhs, np.zeros(oc.nc), tol, mx, mpi=True)
            assert ierr != 2
            oc.correct_fields(pp)
        val, rhs, ap = oc.temperature_assembly()
        T, _, _, ierr = pyoracle.solve("bicgstab_ilu0", rp, ci, val, rhs, oc.get("T"), tol, mx, mpi=True)
        oc.set_field("T", T)
        if ctrl.thermal_model != 0:
            val, rhs, ap = oc.solid_temperature_assembly()
            Ts, _, _, ierr = pyoracle.solve("bicgstab_ilu0", rp, ci, val, rhs, oc.get("T_s"), tol, mx, mpi=True)
            oc.set_field("T_s", Ts)


@pytest.mark.parametrize("name", ["porous_solid_cht", "porous_plug_ltne"])
def test_partitioned_piso_is_the_composition_of_pinned_routines(name):
    """The oracle's piso_run_mpi equals, bit for bit, the sequence of pinned
    per-function routines that mod_uns_simple_mpi.f90:376-497 calls.  On
    porous_solid_cht (a solid block, no LTNE) this fixes the two lines where
    the partitioned driver departs from the serial one: no
    zero_solid_velocity, so the solid cells keep their momentum-solve
    velocities until correct_fields, and no solid-phase solve, so T_s stays
    as it was; the serial step (pinned by the reference run of this case)
    does both.  porous_plug_ltne runs the LTNE solid phase."""
    from mixsolver_amd import zones
    d = load_golden(name)
    ctrl, groups = parse_control(str(d["control"]), warn=False)

    def fresh():
        oc = pyoracle.OracleCase(mesh_from(d), ctrl, groups)
        for k, f in (("u", "in_u"), ("p", "in_p"), ("T", "in_T"), ("T_s", "in_T")):
            oc.set_field(k, d[f])
        return oc

    drv, comp = fresh(), fresh()
    assert drv.piso_run(2, mpi=True) == 0
    _piso_mpi_composed(comp, ctrl, 2)
    for k in ("u", "p", "flux", "T", "T_s", "gp", "apc"):
        assert np.array_equal(drv.get(k), comp.get(k)), k
    if name == "porous_solid_cht":
        assert ctrl.thermal_model == 0
        solid = zones.resolve_cell_zones(mesh_from(d), ctrl.cell_zones) == 3
        assert solid.any()
        assert np.array_equal(drv.get("T_s"), d["in_T"])
        ser = fresh()
        assert ser.piso_run(2, mpi=False) == int(d["piso_ier"])
        assert np.array_equal(ser.get("u").reshape(d["piso_u"].shape), d["piso_u"])
        assert np.array_equal(ser.get("T_s"), d["piso_T_s"])
        assert not np.array_equal(ser.get("T_s"), d["in_T"])


@pytest.mark.parametrize("name", ["funcs_prism", "funcs_outlet", "funcs_channel"])
def test_per_function_bitexact(name):
    d = load_golden(name)
    oc = make_case(d)
    for k in ("u", "p", "T", "flux", "u_old", "u_old_old", "T_old", "T_old_old", "apc"):
        oc.set_field(k, d["in_" + k])
    oc.set_field("T_s", d["in_T"])
    oc.ts_order = int(d["in_ts_order"])
    oc.compute_gradients()
    assert np.array_equal(oc.get("gp"), d["out_gp"])
    assert np.array_equal(oc.get("gu"), d["out_gu"])
    assert np.array_equal(oc.get("gt"), d["out_gt"])
    for comp in (1, 2, 3):
        val, rhs, ap = oc.momentum_assembly(comp)
        assert np.array_equal(val, d[f"out_mom{comp}_val"]), comp
        assert np.array_equal(rhs, d[f"out_mom{comp}_rhs"]), comp
        assert np.array_equal(ap, d[f"out_mom{comp}_ap"]), comp
    val, rhs, ap = oc.temperature_assembly()
    assert np.array_equal(val, d["out_tem_val"])
    assert np.array_equal(rhs, d["out_tem_rhs"])
    assert np.array_equal(ap, d["out_tem_ap"])
    oc.flux_rhiechow()
    assert np.array_equal(oc.get("flux"), d["out_flux_rc"])
    oc.outflow_mass_rescale()
    assert np.array_equal(oc.get("flux"), d["out_flux_out"])
    dpp = oc.correct_fields(d["in_pp"])
    assert np.array_equal(dpp, d["out_cor_dpp"])
    assert np.array_equal(oc.get("u"), d["out_cor_u"])
    assert np.array_equal(oc.get("p"), d["out_cor_p"])
    assert np.array_equal(oc.get("flux"), d["out_cor_flux"])


@pytest.mark.parametrize("name", ["solvers_spd", "solvers_nonsym"])
def test_solver_histories_bitexact(name):
    d = load_golden(name)
    rp, ci, val, b, x0 = d["row_ptr"], d["col_idx"], d["val"], d["b"], d["x0"]
    L, e = pyoracle.factor("ic0", rp, ci, val)
    assert e == int(d["ic0_ierr"])
    if e == 0:
        assert np.array_equal(L, d["ic0_L"])
    LU, e = pyoracle.factor("ilu0", rp, ci, val)
    assert e == int(d["ilu0_ierr"])
    assert np.array_equal(LU, d["ilu0_LU"])
    tol = float(d["tol"])
    for s in ("cg_jacobi", "cg_ic0", "bicgstab_ilu0"):
        for k in range(1, int(d["kmax"]) + 1):
            x, it, res, ierr = pyoracle.solve(s, rp, ci, val, b, x0, tol, k, mpi=False)
            assert it == d[f"{s}_hist_it"][k - 1], (s, k)
            assert ierr == d[f"{s}_hist_ierr"][k - 1], (s, k)
            assert res == d[f"{s}_hist_res"][k - 1], (s, k)
        x, it, res, ierr = pyoracle.solve(s, rp, ci, val, b, x0, tol, int(d["maxit"]), mpi=False)
        assert (it, ierr, res) == (d[f"{s}_it"], d[f"{s}_ierr"], d[f"{s}_res"]), s
        assert np.array_equal(x, d[f"{s}_x"]), s


def test_norm2_matches_flang_runtime_algorithm():
    # closed forms: 3-4-5 triangle, scale invariance and zeros
    assert pyoracle.norm2(np.array([3.0, 4.0, 0.0])) == 5.0
    assert pyoracle.norm2(np.zeros(5)) == 0.0
    v = np.array([1e-300, 3e-300, 4e-300])
    assert abs(pyoracle.norm2(v) - np.sqrt(26.0) * 1e-300) < 1e-312


def test_cavity_known_answer_fixture_matches_published():
    """The reference's own published Re=1000 cavity answer (README of its
    cavity case: max|u| 9.11512E-01, Ghia centreline L2 2.89706E-02) is what
    its solver gives for the shipped mesh and control after 1000 SIMPLE
    iterations; the fixture keeps both."""
    d = load_golden("cavity_re1000_known_answer")
    assert int(d["out_ier"]) == 0
    assert abs(float(d["ref_umax"]) - float(d["readme_umax"])) <= 5e-7
    assert abs(float(d["ref_ghia_l2"]) - float(d["readme_ghia_l2"])) <= 5e-8


def test_ghia_probe_closed_form():
    """probe_velocity of a linear field on a uniform grid of centroids at a
    centroid returns that centroid's value (weight 1e12 dominates)."""
    from mixsolver_amd.post import ghia_compare, probe_velocity
    g = (np.arange(10) + 0.5) / 10
    X, Y = np.meshgrid(g, g, indexing="ij")
    xc = np.stack([X.ravel(), Y.ravel(), np.full(X.size, 0.5)], 1)
    u = np.stack([2 * xc[:, 1] - 1, xc[:, 0], 0 * xc[:, 0]], 1)
    assert np.allclose(probe_velocity(xc, u, xc[37]), u[37], atol=1e-9)
    l2, uq = ghia_compare(xc, u, 1000)
    assert np.isfinite(l2) and uq.shape == (17,)


# ---------------------------------------------------------------- porous zones
@pytest.mark.parametrize("name", POROUS_CASES)
def test_zone_resolution_bitexact(name):
    """Host restatement of resolve_cell_zones + setup_porous_fields
    (mixsolver_amd/zones.py) against the reference's resolved block types and
    per-cell coefficient arrays: explicit types, a VC tag in a zone name,
    name-matched coefficient lines, solid zones."""
    from mixsolver_amd import zones
    d = load_golden(name)
    m = mesh_from(d)
    ctrl, _ = parse_control(str(d["control"]), warn=False)
    cz = zones.resolve_cell_zones(m, ctrl.cell_zones)
    f = zones.setup_porous_fields(m, ctrl.cell_zones, cz)
    assert np.array_equal(cz, d["ts0_out_cztype"])
    for k in zones.FIELDS:
        assert np.array_equal(f[k], d["ts0_out_" + k]), k


@pytest.mark.parametrize("name", POROUS_CASES)
@pytest.mark.parametrize("tso", [0, 2])
def test_porous_per_function_bitexact(name, tso):
    """Kink face pressure (gradients, pressure force), Beavers-Joseph slip,
    unified harmonic viscosity, Darcy-Forchheimer, solid momentum source,
    dispersion, LTNE fluid / solid phases, conjugate heat transfer, solid
    pins and zero_solid_velocity: bit-exact against the reference."""
    d = load_golden(name)
    ctrl, groups = parse_control(str(d["control"]), warn=False)
    oc = pyoracle.OracleCase(mesh_from(d), ctrl, groups)
    out = run_funcs2(oc, d, tso)
    for k, v in out.items():
        assert np.array_equal(np.ravel(v), np.ravel(d[f"ts{tso}_out_{k}"])), k


@pytest.mark.parametrize("name", POROUS_CASES)
@pytest.mark.parametrize("drv", ["simple", "piso"])
def test_porous_driver_bitexact(name, drv):
    """Serial simple_run / piso_run with porous and solid zones, including the
    solid-phase temperature solve."""
    d = load_golden(name)
    ctrl, groups = parse_control(str(d["control"]), warn=False)
    oc = pyoracle.OracleCase(mesh_from(d), ctrl, groups)
    for k, f in (("u", "in_u"), ("p", "in_p"), ("T", "in_T"), ("T_s", "in_T")):
        oc.set_field(k, d[f])
    ier = oc.simple_run(3, mpi=False)[0] if drv == "simple" else oc.piso_run(2, mpi=False)
    assert ier == int(d[f"{drv}_ier"]) == 0
    for k in ("u", "p", "flux", "T", "T_s"):
        assert np.array_equal(oc.get(k).ravel(), d[f"{drv}_{k}"].ravel()), k


# ------------------------------------------- Beavers-Joseph published answers
BJ_GEOM = dict(lx=0.10, ly=0.010, nx=100, ny=40, K=1.0e-6)
BJ_PHYS = dict(H=0.005, K=1.0e-6, mu=1.846e-5, fb=0.05)


def _bj_xc(d):
    ctrl, groups = parse_control(str(d["control_a1"]), warn=False)
    return pyoracle.OracleCase(mesh_from(d), ctrl, groups).geometry()["xc"].reshape(-1, 3)


def _printed(v, fmt="%.4e"):
    return float(fmt % v)


def test_bj_two_layer_closed_form_published():
    """The two-layer closed form (fluid over Brinkman bed, stress jump or
    continuity at the interface) reproduces the reference case's published
    uF1_ref, uD_ref and classical-slip columns to the printed digits, and
    satisfies its own ODEs and interface conditions: wall no-slip on both
    sides, equal shear stress, the alpha-weighted velocity jump."""
    from casecheck import bj_classical_slip, bj_two_layer
    d = load_golden("bj_porous_known_answer")
    H, K, mu, fb = (BJ_PHYS[k] for k in ("H", "K", "mu", "fb"))
    lam, dy = np.sqrt(K), 0.010 / 40
    assert _printed(K * fb / mu) == float(d["readme_uD_ref"])
    for i, a in enumerate(d["readme_alpha"]):
        uf, up, uB = bj_two_layer(a, H, K, mu, fb)
        assert _printed(uf(dy / 2)) == d["readme_uF1_ref"][i]
        assert _printed(bj_classical_slip(a, H, K, mu, fb)) == d["readme_uB_class"][i]
        h = 1e-5
        # fluid: mu u'' + fb = 0; bed: mu u'' - mu/K u + fb = 0 (central differences)
        y = 0.3 * H
        assert abs(mu * (uf(y + h) - 2 * uf(y) + uf(y - h)) / h ** 2 + fb) <= 1e-6 * fb
        y = -0.3 * lam
        assert abs(mu * (up(y + h) - 2 * up(y) + up(y - h)) / h ** 2 - mu / K * up(y) + fb) <= 1e-4 * fb
        assert abs(uf(H)) <= 1e-15 and abs(up(-H)) <= 1e-15
        h = 1e-7
        duf = (uf(h) - uf(-h)) / (2 * h)
        dup = (up(h) - up(-h)) / (2 * h)
        assert abs(duf - dup) <= 1e-6 * abs(duf)
        assert uB == pytest.approx(float(uf(0.0)), rel=1e-14)
        if a > 0:
            assert duf == pytest.approx(a / lam * (uf(0.0) - up(0.0)), rel=1e-6)
        else:
            assert float(uf(0.0)) == pytest.approx(float(up(0.0)), rel=1e-12)


def test_bj_known_answer_fixture_matches_published():
    """The reference's shipped Beavers-Joseph channel, re-run to convergence
    by the reference's own simple_run, gives its README's numbers: iteration
    counts 754 / 882 / 809, the it=1 log line of alpha = 1, and per alpha the
    uF1 / uD columns and the profile RMS error against the closed form."""
    from casecheck import bj_profile, bj_two_layer
    d = load_golden("bj_porous_known_answer")
    xc = _bj_xc(d)
    assert np.array_equal(d["readme_it1_a1"], d["ref_it1_a1"])
    dy = 0.010 / 40
    yc = (np.arange(40) + 0.5) * dy
    for i in range(3):
        assert int(d[f"ref_ier_a{i}"]) == 0
        assert int(d[f"ref_iters_a{i}"]) == int(d["readme_iters"][i])
        col, uF1, uD = bj_profile(xc, d[f"ref_u_a{i}"], **BJ_GEOM)
        assert _printed(uF1) == d["readme_uF1"][i]
        assert _printed(uD) == d["readme_uD"][i]
        uf, up, _ = bj_two_layer(d["readme_alpha"][i], **BJ_PHYS)
        yrel = yc - BJ_PHYS["H"]
        uref = np.where(yrel >= 0.0, uf(np.maximum(yrel, 0.0)), up(np.minimum(yrel, 0.0)))
        rms = np.sqrt(np.mean((col - uref) ** 2)) / uref.max() * 100
        assert float("%.2f" % rms) == d["readme_rms_pct"][i]


def test_bj_oracle_reproduces_reference_run():
    """End to end on the shipped mesh: the oracle's serial SIMPLE driver with
    the stress-jump control (alpha = 1, IC0 PCG, limited second-order upwind,
    zone typed by its 'VC:porous' tag) stops at the same iteration as the
    reference and leaves the same velocity field, bit for bit."""
    d = load_golden("bj_porous_known_answer")
    ctrl, groups = parse_control(str(d["control_a1"]), warn=False)
    oc = pyoracle.OracleCase(mesh_from(d), ctrl, groups)
    ier, it = oc.simple_run(2000, mpi=False)
    assert ier == 0 and it == int(d["ref_iters_a1"])
    assert np.array_equal(oc.get("u").reshape(-1, 3), d["ref_u_a1"])
    h = oc.history()
    assert "%.3E" % h[0, 0] == "%.3E" % d["ref_it1_a1"][0]
    assert "%.2E" % h[0, 5] == "%.2E" % d["ref_it1_a1"][2]


# ------------------------------------------- pure conduction published answers
HC_KIND = dict(c1="linear", c2="uniform", c3="annulus", c4="uniform", c5="source")


def _hc_mesh(d, c):
    return mesh_from({k[len(c) + 1:]: v for k, v in d.items() if k.startswith(c + "_")})


def test_conduction_closed_forms():
    """The closed-form profiles satisfy what the README states of them:
    T(a) = 400 K and T(b) = 480 K for the source bar, the energy balance
    -k T'(b) + k T'(a) = q (b - a) = 4e4 W/m^2, T'' = -q/k inside the source,
    end temperatures 300 / 500 K, and T(r = 0.02) = 426.19 K on the annulus."""
    from casecheck import conduction_exact
    f = lambda x: conduction_exact("source", x, 0.0)
    assert float(f(0.03)) == pytest.approx(400.0, abs=1e-9)
    assert float(f(0.07)) == pytest.approx(480.0, abs=1e-9)
    assert float(f(0.0)) == 300.0 and float(f(0.1)) == pytest.approx(500.0, abs=1e-9)
    for x in (0.03, 0.07):   # value continuity across the source edges (slopes ~5e3 K/m)
        assert abs(float(f(x - 1e-9)) - float(f(x + 1e-9))) <= 2e-5
    h = 1e-6
    d1 = lambda x: (float(f(x + h)) - float(f(x - h))) / (2 * h)
    assert -15.0 * d1(0.08) + 15.0 * d1(0.02) == pytest.approx(4.0e4, rel=1e-6)
    h2 = 1e-4
    assert (float(f(0.05 + h2)) - 2 * float(f(0.05)) + float(f(0.05 - h2))) / h2 ** 2 == pytest.approx(-1e6 / 15.0,
                                                                                                         rel=1e-6)
    assert float(conduction_exact("annulus", 0.02, 0.0)) == pytest.approx(426.19, abs=5e-3)
    assert float(conduction_exact("annulus", 0.0, 0.03)) == pytest.approx(500.0, abs=1e-9)
    assert float(conduction_exact("linear", 0.05, 0.0)) == 400.0


@pytest.mark.parametrize("c", list(HC_KIND))
def test_conduction_known_answer(c):
    """The reference's pure-conduction family (all cells solid, fixed and
    adiabatic walls, a source zone in c5): the reference's own solid
    temperature has the README's maximum / RMS error against the closed form
    (printed digits; c1, c2, c4 within the 1e-6 K tolerance), and the
    oracle's simple_run reproduces it bit for bit."""
    from casecheck import conduction_exact
    d = load_golden("heat_conduction_known_answer")
    m = _hc_mesh(d, c)
    ctrl, groups = parse_control(str(d[f"{c}_control"]), warn=False)
    oc = pyoracle.OracleCase(m, ctrl, groups)
    xc = oc.geometry()["xc"].reshape(-1, 3)
    ts = d[f"{c}_ref_T_s"]
    err = ts - conduction_exact(HC_KIND[c], xc[:, 0], xc[:, 1])
    emax, rms, tol = d[f"{c}_readme"]
    assert int(d[f"{c}_ref_ier"]) == 0 and int(d[f"{c}_ref_iters"]) == 1
    if emax > 0:
        assert float("%.3e" % np.abs(err).max()) == emax
        assert float("%.3e" % np.sqrt(np.mean(err ** 2))) == rms
    else:
        assert np.abs(err).max() <= tol
    ier, it = oc.simple_run(3000, mpi=False)
    assert ier == 0 and it == 1
    assert np.array_equal(oc.get("T_s"), ts)


# ------------------------------------------- Re=40 cylinder: forces and indicators
def _cyl_case(d):
    ctrl, groups = parse_control(str(d["control"]), warn=False)
    oc = pyoracle.OracleCase(mesh_from(d), ctrl, groups)
    oc.set_field("u", d["ref_u"])
    oc.set_field("p", d["ref_p"])
    return oc


def test_cylinder_known_answer_fixture_matches_published():
    """The reference's shipped Re=40 cylinder after its 8000 SIMPLE
    iterations gives the README's regression indicators (its own check
    indicators computed by casecheck.cylinder_check): Cp at the front stagnation
    point, the base and the suction peak, the separation angles, the
    recirculation length, max|u|, the strongest reverse flow, the ring's
    pressure drag and the Cp deviation from the archived wall baseline."""
    from casecheck import cylinder_check
    d = load_golden("cylinder_re40_known_answer")
    assert int(d["ref_ier"]) == 0
    xc = _cyl_case(d).geometry()["xc"].reshape(-1, 3)
    a = d["archived_cp"]
    c = cylinder_check(xc, d["ref_u"], d["ref_p"], cp_ref=(a[:, 0], a[:, 1]))
    f4 = lambda v: float("%.4f" % v)
    assert c["ring_cells"] == 80
    assert f4(c["cp_front"]) == d["readme_cp_front"]
    assert f4(c["cp_base"]) == d["readme_cp_base"]
    assert f4(c["cp_min"]) == d["readme_cp_min"]
    assert float("%.1f" % c["phi_cp_min"]) == d["readme_phi_cp_min"]
    assert [float("%.1f" % v) for v in c["separation_from_wake_axis"]] == list(d["readme_sep"])
    assert f4(c["wake_length_D"]) == d["readme_wake_D"]
    assert float("%.6f" % c["umax"]) == d["readme_umax"]
    assert float("%.5f" % c["max_reverse_ux"]) == d["readme_max_reverse"]
    assert f4(c["cd_pressure"]) == d["readme_cd_pressure"]
    assert f4(c["dcp_mean"]) == d["readme_dcp_mean"] and f4(c["dcp_max"]) == d["readme_dcp_max"]


def test_cylinder_forces_oracle_matches_reference():
    """compute_forces on the final fields: the oracle's Fx Fy Fz CD CL u_inf
    A_ref p_inf equal the reference's bit for bit, its pressure / viscous
    parts print as the reference printed them; the Cp rows equal the
    reference's table to its printed digits (f10.3, es12.4) and the
    separation angles and wake length to theirs (f8.2, f8.3)."""
    d = load_golden("cylinder_re40_known_answer")
    oc = _cyl_case(d)
    f = oc.compute_forces()
    ref = d["ref_forces"]
    for i, k in enumerate(pyoracle.FORCE_KEYS[:8]):
        assert f[k] == ref[i], k
    q_a = 0.5 * 1.0 * f["u_inf"] ** 2 * f["A_ref"]
    printed = ["%.6E" % v for v in d["ref_forces_printed"]]
    assert ["%.6E" % v for v in (f["Fx_p"], f["Fx_v"], f["Fx_p"] / q_a, f["Fx_v"] / q_a)] == printed
    cp, rt = oc.cp_distribution(), d["ref_cp_table"]
    assert cp.shape == rt.shape == (80, 6)
    assert ["%10.3f" % v for v in cp[:, 0]] == ["%10.3f" % v for v in rt[:, 0]]
    for j in (1, 2):
        assert ["%.4E" % v for v in cp[:, j]] == ["%.4E" % v for v in rt[:, j]]
    s = oc.separation()
    assert s["found_top"] and s["found_bot"]
    got = [float("%.2f" % np.degrees(s["theta_top"])), float("%.2f" % np.degrees(s["theta_bot"])),
           float("%.3f" % s["wake_length"])]
    assert got == list(d["ref_separation_printed"])


def test_forces_closed_forms():
    """compute_forces on a box with one wall zone (the y = 0 side of a
    channel): a uniform gauge pressure p0 gives F = p0 * sum of the wall
    face vectors (they point out of the fluid, -y: Fy = -p0 * area), and a
    uniform tangential velocity U at the wall cells gives the wall shear
    mu * U / dn * area along t = n x z; the reference area is the x extent
    of the wall face centroids times the z extent of the nodes."""
    from mixsolver_amd import meshgen
    from mixsolver_amd.control import parse_control as pc
    m = meshgen.hex_box(np.linspace(0.0, 3.0, 7), np.linspace(0.0, 2.0, 5), np.linspace(0.0, 1.0, 3),
                        zone_ids=dict(ymin=6))
    ctl = "rho = 1.0\nmu = 0.1\nbc = 6 wall\nbc = 5 symmetry\nbc = 4 symmetry\n"
    ctrl, groups = pc(ctl, warn=False)
    oc = pyoracle.OracleCase(m, ctrl, groups)
    nc = m.ncells
    oc.set_field("p", np.full(nc, 2.5))
    u = np.zeros((nc, 3))
    u[:, 0] = 0.4
    oc.set_field("u", u)
    f = oc.compute_forces()
    wall_area = 3.0 * 1.0
    assert f["u_inf"] == 0.0 and f["CD"] == 0.0
    assert f["Fy_p"] == pytest.approx(-2.5 * wall_area, rel=1e-14)
    assert abs(f["Fx_p"]) <= 1e-14 and abs(f["Fz_p"]) <= 1e-14
    # wall-adjacent cell centres sit dn = (2.0 / 4) / 2 above the wall; n = -y, t = n x z = (-1, 0, 0)
    assert f["Fx_v"] == pytest.approx(0.1 * 0.4 * -1.0 / 0.25 * -1.0 * wall_area, rel=1e-13)
    # A_ref = x extent of the wall face centroids (0.25 .. 2.75) times the node z extent
    assert f["A_ref"] == pytest.approx(2.5 * 1.0, rel=1e-14)


# ------------------------------------------------------ coupling interface
from conftest import IFACE_CASES, iface_mesh, seed_interface  # noqa: E402


@pytest.mark.parametrize("name", IFACE_CASES)
@pytest.mark.parametrize("tso", [0, 2])
def test_interface_per_function_bitexact(name, tso):
    """Coupling-interface faces (an interface-named zone without a bc line):
    velocity-Dirichlet, Zhang viscous closure, pressure-Dirichlet with the
    tangential stress jump, Dirichlet / Neumann temperature with the LTNE
    phase split -- every assembled value, gradient, flux and correction
    bit-exact against the reference's own routines fed through
    set_interface_* and the set_iface_* switches."""
    d = load_golden(name)
    ctrl, groups = parse_control(str(d["control"]), warn=False)
    oc = pyoracle.OracleCase(iface_mesh(d), ctrl, groups)
    seed_interface(oc, d)
    out = run_funcs2(oc, d, tso)
    for k, v in out.items():
        assert np.array_equal(np.ravel(v), np.ravel(d[f"ts{tso}_out_{k}"])), k


@pytest.mark.parametrize("name", IFACE_CASES)
def test_interface_simple_bitexact(name):
    """Serial simple_run (3 iterations) with the coupling interface: the
    serial PPE treats interface faces as pressure-Dirichlet under the
    interface pressure model (no pinned cell)."""
    d = load_golden(name)
    ctrl, groups = parse_control(str(d["control"]), warn=False)
    oc = pyoracle.OracleCase(iface_mesh(d), ctrl, groups)
    seed_interface(oc, d)
    for k in ("u", "p", "T"):
        oc.set_field(k, d[f"ts0_in_{k}"])
    ier, it = oc.simple_run(3, mpi=False)
    assert ier == int(d["simple_ier"]) == 0
    for k in ("u", "p", "flux", "T", "T_s"):
        assert np.array_equal(oc.get(k).ravel(), d[f"simple_{k}"].ravel()), k


# ------------------------------------------------------------ NaN guard
from conftest import nan_guard_groups  # noqa: E402


def test_nan_guard_groups_oracle():
    """nan_scan restated: the reference unit test's groups (clean, NaN / Inf
    detection with field, cell and kind, first offender by cell index, T_s
    scanned, stateless), and the drivers' NAN_IER = 60 when T_s alone is
    non-finite (the guard scans T_s too)."""
    d = load_golden("cavity16_simple_jacobi")
    ctrl, groups = parse_control(str(d["control"]), warn=False)
    oc = pyoracle.OracleCase(mesh_from(d), ctrl, groups)
    nan_guard_groups(oc, oc.nc)
    oc2 = pyoracle.OracleCase(mesh_from(d), ctrl, groups)
    ts = np.zeros(oc2.nc)
    ts[5] = np.nan
    oc2.set_field("T_s", ts)
    assert oc2.simple_run(2, mpi=False)[0] == 60


def test_channel_reference_run_matches_readme():
    """The reference's low-speed channel re-run by the fixture generator
    (its own SIMPLE, IC0 PCG) and measured by the restated channel check
    gives the README's table to its printed digits, converges after the
    printed 773 iterations and prints the same first-iteration line; the
    rebuilt tensor grid holds 101 y levels clustered at both walls."""
    d = load_golden("channel_lowspeed_known_answer")
    assert int(d["ref_ier"]) == 0
    assert int(d["ref_iterations"]) == int(d["readme_iterations"])
    assert np.array_equal(d["ref_it1"], d["readme_it1"]) and int(d["ref_it1_lin"]) == int(d["readme_it1_lin"])

    def sig(a, n):
        return float(f"{float(a):.{n - 1}e}")

    for k, n in (("u_bulk_50", 7), ("u_bulk_out", 7), ("umag_max", 7), ("dpdx_fit", 7), ("p_min", 7),
                 ("p_rise", 5)):
        assert sig(d["ref_" + k], n) == float(d["readme_" + k]), k
    assert round(float(d["ref_umax_ratio_out"]), 4) == float(d["readme_umax_ratio_out"])
    for k in ("profile_err_max", "profile_err_l1", "mass_drift"):
        assert round(100 * float(d["ref_" + k]), 3) == pytest.approx(100 * float(d["readme_" + k]), abs=1e-9), k
    assert float(d["ref_entrance_length"]) == float(d["readme_entrance_length"])
    assert np.allclose(np.round(d["ref_umag_max_at"], 3), d["readme_umag_max_at"], atol=1e-9)
    ys = d["ys"]
    assert len(ys) == 101 and ys[0] == 0.0 and ys[-1] == 10.0
    assert np.diff(ys).min() == pytest.approx(0.01, rel=1e-9)
    assert np.allclose(ys + ys[::-1], 10.0, atol=1e-12)
    # the outlet profile is the developed Poiseuille profile within the printed max error
    eta = d["ref_out_y"] / 10.0
    ub = float(d["ref_u_bulk_target"])
    assert np.abs(d["ref_out_u"] - 6 * ub * eta * (1 - eta)).max() / ub == pytest.approx(
        float(d["ref_profile_err_max"]), rel=1e-12)


# ------------------------------------------------ loop-level Fortran drop-in
CASE_DUMP = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "case_dump")


@pytest.mark.skipif(not os.path.exists(CASE_DUMP), reason="the reference build (oracle/build_ref.sh) is absent")
@pytest.mark.parametrize("name,zk", [
    ("channel_piso_jacobi", {4: "symmetry", 5: "wall", 6: "velocity-inlet", 7: "outflow"}),
    ("porous_solid_cht", {4: "symmetry", 5: "wall", 6: "velocity-inlet", 7: "outflow"}),
    ("iface_pdir_slip_ltne", {4: "symmetry", 5: "wall", 6: "interface", 7: "outflow"}),
    ("cavity16_simple_ic0_blend", {4: "symmetry", 5: "wall"})])
def test_loop_dropin_case_conversion(name, zk, tmp_path):
    """mixsolver_case.f90 (the loop-level drop-in's adapter) turns the
    reference's own mesh_t / ctrl_t / bc_t / fields_t, set up by the
    reference's read_cas, read_control, resolve_cell_zones, build_bc,
    init_fields and setup_porous_fields, into the same ms_ctrl_t, boundary
    groups, mesh arrays, cell types and porous coefficients the Python
    front end builds from the same files (the device then recomputes
    uspeed, lid_tol and tbc_tol from the mesh)."""
    import subprocess
    from mixsolver_amd import cas, zones
    from mixsolver_amd.control import BC_VINLET_PARAB, MsBcGroup, MsCtrl, groups_array, with_interface_groups
    d = load_golden(name)
    m0 = mesh_from(d)
    if name.startswith("iface"):
        m0 = iface_mesh(d)
    cas.write_cas(str(tmp_path / "m.cas"), m0, zk)
    ctl = str(d["control"])
    (tmp_path / "c.control").write_text(ctl)
    r = subprocess.run([CASE_DUMP, "m.cas", "c.control", "out.bin"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = (tmp_path / "out.bin").read_bytes()
    nn, nf, nc, nb = np.frombuffer(raw, np.int32, 4)
    o = 16
    fc = MsCtrl.from_buffer_copy(raw[o:o + C.sizeof(MsCtrl)])
    o += C.sizeof(MsCtrl)
    fg = []
    for _ in range(nb):
        fg.append(MsBcGroup.from_buffer_copy(raw[o:o + C.sizeof(MsBcGroup)]))
        o += C.sizeof(MsBcGroup)

    def take(dt, n):
        nonlocal o
        a = np.frombuffer(raw, dt, n, o)
        o += a.nbytes
        return a

    x, fptr = take(np.float64, 3 * nn), take(np.int64, nf + 1)
    fnodes, c0, c1, zone = take(np.int32, int(fptr[-1])), take(np.int32, nf), take(np.int32, nf), take(np.int32, nf)
    has_cz = take(np.int32, 1)[0]
    cz = take(np.int32, nc) if has_cz else None
    por, perm, inert, ks, qs = (take(np.float64, nc), take(np.float64, 3 * nc), take(np.float64, nc),
                                take(np.float64, nc), take(np.float64, nc))
    # the Python front end on the same files
    m = cas.read_cas(str(tmp_path / "m.cas"))
    ctrl, groups = parse_control(ctl, warn=False)
    groups = with_interface_groups(m, ctrl, groups)
    assert nb == len(groups) == ctrl.nbc == fc.nbc
    for fname, _ in MsCtrl._fields_:
        a, b = getattr(fc, fname), getattr(ctrl, fname)
        if hasattr(a, "__len__"):
            assert np.array_equal(np.ctypeslib.as_array(a), np.ctypeslib.as_array(b)), fname
        else:
            assert a == b, fname
    garr = groups_array(groups)
    for i in range(nb):
        for fname, _ in MsBcGroup._fields_:
            if fname in ("uspeed", "lid_tol", "tbc_tol"):
                continue
            if fname.startswith("prfl_") and garr[i].btype != BC_VINLET_PARAB:
                continue      # read only for parabolic inlets
            a, b = getattr(fg[i], fname), getattr(garr[i], fname)
            if hasattr(a, "__len__"):
                assert np.array_equal(np.ctypeslib.as_array(a), np.ctypeslib.as_array(b)), (i, fname)
            else:
                assert a == b, (i, fname)
    assert np.array_equal(x, np.ravel(m.x)) and np.array_equal(fptr, m.fptr)
    for a, b in ((fnodes, m.fnodes), (c0, m.c0), (c1, m.c1), (zone, m.zone)):
        assert np.array_equal(a, b)
    if ctrl.cell_zones:
        czp = zones.resolve_cell_zones(m, ctrl.cell_zones)
        f = zones.setup_porous_fields(m, ctrl.cell_zones, czp)
        assert has_cz and np.array_equal(np.maximum(cz, 1), czp)
        assert np.array_equal(por, f["porosity"]) and np.array_equal(perm, np.ravel(f["perm_dir"]))
        assert np.array_equal(inert, f["inertial"]) and np.array_equal(ks, f["k_s"])
        assert np.array_equal(qs, f["q_src"])
