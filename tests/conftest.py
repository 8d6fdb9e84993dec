import os
import re
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def load_golden(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))


def mesh_from(d):
    from mixsolver_amd.meshgen import Mesh
    m = Mesh(x=d["x"], fptr=d["fptr"], fnodes=d["fnodes"], c0=d["c0"], c1=d["c1"],
             zone=d["zone"], ncells=int(d["ncells"]), ctype=int(d["ctype"]))
    if "czone" in d:
        m.czone = d["czone"]
        m.czone_names = {int(i): (str(c), str(n))
                         for i, c, n in zip(d["czone_ids"], d["czone_cond"], d["czone_name"])}
    return m


POROUS_CASES = ("porous_plug_ltne", "porous_plug_unified", "porous_solid_cht")
IFACE_CASES = ("iface_legacy", "iface_zhang", "iface_pdir_slip_ltne")
IFACE_KEYS = ("vel", "p", "T", "Ts", "q", "slip_C", "peer_vel", "T_mode")


def iface_mesh(d):
    """Mesh of an interface fixture: zone 6 is the named coupling interface."""
    m = mesh_from(d)
    m.zones = {6: ("interface", "interface-6")}
    return m


def seed_interface(case, d):
    """Apply an interface fixture's per-face data and switches to a device
    Case or an OracleCase."""
    for k in IFACE_KEYS:
        case.set_interface(k, d["iface_" + k])
    p_dir, slip, zhang, t_ltne = (bool(v) for v in d["switches"])
    case.set_interface_modes(p_dir=p_dir, slip=slip, zhang=zhang, T_ltne=t_ltne)
FUNCS2_FIELDS = ("u", "p", "T", "flux", "u_old", "u_old_old", "T_old", "T_old_old", "apc", "T_s", "T_s_old",
                 "T_s_old_old")


def run_funcs2(case, d, tso):
    """Seed case (device Case or OracleCase) with the inputs of a funcs2
    fixture and return its per-function outputs under the fixture's keys."""
    for k in FUNCS2_FIELDS:
        case.set_field(k, d[f"ts{tso}_in_{k}"])
    if hasattr(case, "set_ts_order"):
        case.set_ts_order(tso)
    else:
        case.ts_order = tso
    out = {}
    case.compute_gradients()
    for k in ("gp", "gu", "gt", "gts"):
        out[k] = case.get(k)
    for comp in (1, 2, 3):
        out[f"mom{comp}_val"], out[f"mom{comp}_rhs"], out[f"mom{comp}_ap"] = case.momentum_assembly(comp)
    out["tem_val"], out["tem_rhs"], out["tem_ap"] = case.temperature_assembly()
    out["sol_val"], out["sol_rhs"], out["sol_ap"] = case.solid_temperature_assembly()
    case.flux_rhiechow()
    out["flux_rc"] = case.get("flux")
    case.outflow_mass_rescale()
    out["flux_out"] = case.get("flux")
    out["cor_dpp"] = case.correct_fields(d[f"ts{tso}_in_pp"])
    for k, f in (("cor_u", "u"), ("cor_p", "p"), ("cor_flux", "flux"), ("cor_u_old", "u_old")):
        out[k] = case.get(f)
    return out


@pytest.fixture
def golden():
    return load_golden


def nan_guard_groups(case, nc):
    """The six groups of the reference's NaN-guard unit test
    (src/unstructured/test_nan_guard.f90:72-121) on a Case or OracleCase:
    kc1 = nc/2 + 1, kc2 = max(1, nc/4) (1-based)."""
    kc1, kc2 = nc // 2 + 1, max(1, nc // 4)
    if kc1 == kc2:
        kc1 += 1
    fields = {k: np.zeros(3 * nc if k == "u" else nc) for k in ("u", "p", "T", "T_s")}

    def put(kv=None):
        for k, v in fields.items():
            case.set_field(k, v)
        for (k, i), v in (kv or {}).items():
            a = fields[k].copy()
            a[i] = v
            case.set_field(k, a)
        return case.nan_scan()

    assert put()[0] == 0                                                    # group 1: clean
    ifin, f, c, v, kind = put({("u", 3 * (kc1 - 1) + 1): np.nan})        # group 2: NaN in u_y
    assert (ifin, f, c, kind) == (1, "u_y", kc1, "NaN") and np.isnan(v)
    ifin, f, c, v, kind = put({("p", kc2 - 1): np.inf})                   # group 3: Inf in p
    assert (ifin, f, c, kind) == (1, "p", kc2, "Inf") and v == np.inf
    ifin, f, c, v, kind = put({("T", kc2 - 1): np.nan, ("u", 3 * (kc1 - 1)): np.inf})
    assert (ifin, f, c) == (1, "T", kc2)                                    # group 4: lowest cell first
    ifin, f, c, v, kind = put({("T_s", kc1 - 1): np.nan})                 # group 5: T_s scanned
    assert (ifin, f, c) == (1, "T_s", kc1)
    assert put()[0] == 0                                                    # group 6: stateless


# The per-step / per-iteration lines the reference's drivers print:
# piso_run / pimple_run (mod_uns_simple.f90:519-523, 771-775) and
# simple_run (:249-265).
_STEP_RE = re.compile(r"step=\s*(\d+)\s+t=\s*(\S+)\s+mass-imbal=\s*(\S+)\s+du/dt=\s*(\S+)\s+max\|u\|=\s*(\S+)"
                      r"\s+CFL~=\s*(\S+)\s+lin-it=\s*(\d+)\s+lin-res=\s*(\S+)")
_IT_RE = re.compile(r"\bit=\s*(\d+)\s+mass-imbal=\s*(\S+)\s+du_max=\s*(\S+)\s+dT_max=\s*(\S+)"
                    r"\s+lin-it=\s*(\d+)\s+lin-res=\s*(\S+)")


def printed_history(log):
    """Rows of the reference's printed diagnostics as (iteration, dict of
    printed strings keyed by the history columns imbal, du, col2 (max|u|
    for time steps, dT_max for SIMPLE), cfl, lin_it, lin_res, with their
    Fortran edit descriptors)."""
    rows = []
    for ln in str(log).splitlines():
        m = _STEP_RE.search(ln)
        if m:
            g = m.groups()
            rows.append((int(g[0]), dict(imbal=(g[2], 9, 2), du=(g[3], 9, 2), col2=(g[4], 9, 2),
                                         cfl=(g[5], 9, 2), lin_it=int(g[6]), lin_res=(g[7], 9, 2))))
            continue
        m = _IT_RE.search(ln)
        if m:
            g = m.groups()
            rows.append((int(g[0]), dict(imbal=(g[1], 10, 3), du=(g[2], 10, 3), col2=(g[3], 10, 3),
                                         lin_it=int(g[4]), lin_res=(g[5], 9, 2))))
    return rows


def fortran_es(v, w, d):
    """Fortran ESw.d output of v (exponent of two digits, as flang writes for
    |exponent| < 100)."""
    return f"{v:.{d}E}".rjust(w).strip()


HIST_COLS = dict(imbal=0, du=1, col2=2, cfl=3, lin_it=4, lin_res=5)


def check_printed_history(hist, log, exact=True, skip=()):
    """Compare a driver history (rows of imbal du col2 cfl lin_it lin_res ...)
    with the printed lines: the same strings when exact, else within the
    printed precision (one unit in the last printed digit) and lin-it within
    two (the largest iteration count of the step's solves: BiCGSTAB's
    residual is not monotone, so reduction-order differences near the
    tolerance move its stopping iteration by up to two).  Returns the number
    of printed rows checked."""
    rows = printed_history(log)
    assert rows, "no printed diagnostics in the fixture"
    for it, vals in rows:
        h = hist[it - 1]
        for key, pv in vals.items():
            if key in skip:
                continue
            got = h[HIST_COLS[key]]
            if key == "lin_it":
                assert (int(got) == pv) if exact else abs(int(got) - pv) <= 2, (it, key, got, pv)
                continue
            text, w, d = pv
            if exact:
                assert fortran_es(got, w, d) == text, (it, key, got, text)
            else:
                ref = float(text)
                ulp = 10.0 ** (int(f"{ref:.{d}E}".split("E")[1]) - d) if ref != 0.0 else 0.0
                assert abs(got - ref) <= 1.01 * ulp, (it, key, got, text)
    return len(rows)
