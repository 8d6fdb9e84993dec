"""The CPU oracle against the reference's own FP64 Fortran outputs.

Every fixture under tests/golden was written by the reference code itself
(tests/golden/make_golden.py drives the reference modules through
oracle/ref_harness.f90).  The oracle restates the same arithmetic in the same
order, so every comparison here is bit-exact (np.array_equal).
"""
import numpy as np
import pytest

from conftest import POROUS_CASES, load_golden, mesh_from, run_funcs2
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
