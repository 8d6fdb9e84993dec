"""Coupling-interface faces on the device path against the reference.

Fixtures tests/golden/iface_*.npz come from the reference's own routines
(oracle/ref_harness.f90 modes ifuncs / isimple): the inlet zone of the zoned
LTNE channel is an interface-named zone without a bc line, its per-face
velocity / pressure / temperature / heat flux / slip data are applied through
set_interface_* and the coupling switches (velocity-Dirichlet legacy, Zhang
viscous closure, pressure-Dirichlet with the tangential stress jump and the
LTNE phase split).  Bars: every assembled value bit-exact (also under the
colour-blocked renumbering); simple_run rel-L2 <= 1e-10; PISO (the
partitioned driver, whose PPE keeps interface faces non-Dirichlet) against
the oracle's piso_run_mpi to 1e-10.
"""
import numpy as np
import pytest

from conftest import FUNCS2_FIELDS, IFACE_CASES, iface_mesh, load_golden, seed_interface
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
    cs = Case(iface_mesh(d), ctrl, groups)
    seed_interface(cs, d)
    return cs


@pytest.mark.parametrize("name", IFACE_CASES)
@pytest.mark.parametrize("tso", [0, 2])
@pytest.mark.parametrize("reorder", [None, "dilu"])
def test_interface_per_function_bitexact(name, tso, reorder):
    """Gradients, momentum (3 components), fluid and solid temperature,
    Rhie-Chow flux and correction bit-exact; the outflow rescale to rtol
    1e-14 (its global sums are tree-ordered), after which the reference's
    rescaled flux is loaded so the correction is compared bit for bit."""
    d = load_golden(name)
    cs = make(d, **(dict(gpu_ppe_precond=2, gpu_mom_precond=2) if reorder else {}))
    for k in FUNCS2_FIELDS:
        cs.set_field(k, d[f"ts{tso}_in_{k}"])
    cs.set_ts_order(tso)
    o = lambda k: d[f"ts{tso}_out_{k}"]
    cs.compute_gradients()
    for k in ("gp", "gu", "gt", "gts"):
        assert np.array_equal(cs.get(k).ravel(), o(k).ravel()), k
    for comp in (1, 2, 3):
        val, rhs, ap = cs.momentum_assembly(comp)
        assert np.array_equal(val, o(f"mom{comp}_val")), comp
        assert np.array_equal(rhs, o(f"mom{comp}_rhs")), comp
        assert np.array_equal(ap, o(f"mom{comp}_ap")), comp
    for t, fn in (("tem", cs.temperature_assembly), ("sol", cs.solid_temperature_assembly)):
        val, rhs, ap = fn()
        assert np.array_equal(val, o(f"{t}_val")), t
        assert np.array_equal(rhs, o(f"{t}_rhs")), t
        assert np.array_equal(ap, o(f"{t}_ap")), t
    cs.flux_rhiechow()
    assert np.array_equal(cs.get("flux"), o("flux_rc"))
    cs.outflow_mass_rescale()
    np.testing.assert_allclose(cs.get("flux"), o("flux_out"), rtol=1e-14, atol=0)
    cs.set_field("flux", o("flux_out"))
    dpp = cs.correct_fields(d[f"ts{tso}_in_pp"])
    assert np.array_equal(dpp.ravel(), o("cor_dpp").ravel())
    assert np.array_equal(cs.get("u").ravel(), o("cor_u").ravel())
    assert np.array_equal(cs.get("p"), o("cor_p"))
    assert np.array_equal(cs.get("flux"), o("cor_flux"))
    assert np.array_equal(cs.get("u_old").ravel(), o("cor_u_old").ravel())


@pytest.mark.parametrize("name", IFACE_CASES)
def test_interface_simple_matches_reference(name):
    d = load_golden(name)
    cs = make(d)
    for k in ("u", "p", "T"):
        cs.set_field(k, d[f"ts0_in_{k}"])
    ier, it = cs.simple_run(3)
    assert ier == int(d["simple_ier"]) == 0
    for k in ("u", "p", "flux", "T", "T_s"):
        assert rel_l2(cs.get(k), d[f"simple_{k}"]) <= 1e-10, k


@pytest.mark.parametrize("name", IFACE_CASES)
def test_interface_piso_matches_oracle(name):
    """The partitioned PISO driver keeps interface faces out of the PPE's
    Dirichlet set (ppe_assembly_mpi tests the types) while correct_fields
    applies the interface pressure: the device run equals the oracle's
    piso_run_mpi restatement to 1e-10."""
    d = load_golden(name)
    ctrl, groups = parse_control(str(d["control"]), warn=False)
    oc = pyoracle.OracleCase(iface_mesh(d), ctrl, groups)
    seed_interface(oc, d)
    cs = make(d)
    for case in (oc, cs):
        for k in ("u", "p", "T", "T_s"):
            case.set_field(k, d[f"ts0_in_{k}"])
    assert oc.piso_run(2, mpi=True) == 0
    assert cs.piso_run(2) == 0
    for k in ("u", "p", "flux", "T", "T_s"):
        assert rel_l2(cs.get(k), oc.get(k)) <= 1e-10, k
