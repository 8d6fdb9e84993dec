"""Cell zones: block types and per-cell porous-medium coefficients.

Host-side restatement of the reference's zone set-up, which runs once before
the solver loop:

  resolve_cell_zones   mod_uns_control.f90:575-733 (+ vc_tag_type :755-799)
  setup_porous_fields  mod_uns_fields.f90:418-504

The result is what the device system consumes: the block type of every cell
(fluid / porous / solid) and the per-cell coefficient arrays (porosity,
diagonal permeability, Forchheimer coefficient, thermal dispersivities,
Beavers-Joseph coefficient, solid properties, fluid-solid exchange and solid
heat source).  Per-cell fields given directly (``Case.set_field``) take the
same path, so a uniform per-cell field and the equivalent zone entry give
identical systems.
"""
from __future__ import annotations

import numpy as np

from .control import CZ_AUTO, CZ_FLUID, CZ_POROUS, CZ_SOLID, ControlError

FIELDS = ("porosity", "perm_dir", "inertial", "disp_l", "disp_t", "bj_alpha", "k_s", "cp_s", "rho_s",
          "h_sf", "a_sf", "q_src")


def vc_tag_type(name1: str, name2: str) -> int:
    """Block type named by a 'VC: fluid|porous|solid' tag in a zone's
    condition or user name (mod_uns_control.f90:755-799); 0 without a tag."""
    lc = (name1.strip().lower() + " " + name2.strip().lower()).rstrip()
    p = lc.find("vc")
    while p >= 0:
        if p == 0 or not ("a" <= lc[p - 1] <= "z"):
            q = p + 2
            while q < len(lc) and lc[q] in ": ":
                q += 1
            if q < len(lc):
                e = q
                while e < len(lc) and lc[e] != " ":
                    e += 1
                tok = lc[q:e][:24].rstrip()
                while tok and tok[-1] in ",;:=.":
                    tok = tok[:-1]
                if tok in ("porous", "fluid", "solid"):
                    return {"porous": CZ_POROUS, "fluid": CZ_FLUID, "solid": CZ_SOLID}[tok]
        p = lc.find("vc", p + 1)
    return 0


def zone_table(mesh):
    """Cell-zone table in the reader's order (ascending zone id,
    mod_uns_cas_reader.f90:626-632): [(id, cond_name, user_name)]."""
    cz = mesh.cell_zones()
    out = []
    for z in sorted(set(int(v) for v in cz)):
        if z < 1:
            continue
        cond, name = mesh.czone_names.get(z, ("", ""))
        out.append((z, cond, name))
    return out


def _find(table, cz, fold_case):
    for iz, (zid, cond, user) in enumerate(table):
        if cz["id"] != 0:
            if zid == cz["id"]:
                return iz
        else:
            want = cz["name"].strip()
            c, u = cond.strip(), user.strip()
            if fold_case:
                want, c, u = want.lower(), c.lower(), u.lower()
            if c == want or u == want:
                return iz
    return -1


def resolve_cell_zones(mesh, cell_zones):
    """Block type (CZ_FLUID / CZ_POROUS / CZ_SOLID) of every cell: explicit
    cell_zone types first, then VC tags in zone names, fluid by default."""
    czone = mesh.cell_zones()
    table = zone_table(mesh)
    cztype = np.full(mesh.ncells, CZ_FLUID, np.int32)
    explicit = np.zeros(mesh.ncells, bool)
    for cz in cell_zones:
        iz = _find(table, cz, fold_case=True)
        if iz < 0:
            what = f"id {cz['id']}" if cz["id"] != 0 else f"name: {cz['name']}"
            raise ControlError(11, f"cell_zone refers to unknown zone {what}")
        if cz["ztype"] != CZ_AUTO:
            sel = czone == table[iz][0]
            cztype[sel] = cz["ztype"]
            explicit[sel] = True
    for zid, cond, user in table:
        vt = vc_tag_type(cond, user)
        if vt == 0:
            continue
        sel = (czone == zid) & ~explicit
        cztype[sel] = vt
    return cztype


def setup_porous_fields(mesh, cell_zones, cztype):
    """Per-cell coefficient arrays (natural cell order): fluid defaults
    (porosity 1, everything else 0) overwritten from the matching cell_zone
    entry for porous cells (entry type porous or coefficients-only) and
    solid cells (entry type solid)."""
    n = mesh.ncells
    f = {k: np.zeros(n) for k in FIELDS}
    f["porosity"][:] = 1.0
    f["perm_dir"] = np.zeros((n, 3))
    czone = mesh.cell_zones()
    table = zone_table(mesh)
    for iz, (zid, _, _) in enumerate(table):
        j = -1
        for jj, cz in enumerate(cell_zones):
            if cz["id"] != 0:
                hit = cz["id"] == zid
            else:
                hit = _find([table[iz]], cz, fold_case=False) == 0
            if hit:
                j = jj
                break
        if j < 0:
            continue
        cz = cell_zones[j]
        cells = np.nonzero(czone == zid)[0]
        por = cells[cztype[cells] == CZ_POROUS]
        if cz["ztype"] in (CZ_POROUS, CZ_AUTO) and len(por):
            for d, key in enumerate(("perm_xx", "perm_yy", "perm_zz")):
                f["perm_dir"][por, d] = cz[key] if cz[key] > 0.0 else cz["perm"]
            for key in ("inertial", "porosity", "disp_l", "disp_t", "bj_alpha", "k_s", "cp_s", "rho_s",
                        "h_sf", "a_sf"):
                f[key][por] = cz[key]
        sol = cells[cztype[cells] == CZ_SOLID]
        if cz["ztype"] == CZ_SOLID and len(sol):
            f["porosity"][sol] = 0.0
            for key in ("k_s", "cp_s", "rho_s", "q_src"):
                f[key][sol] = cz[key]
    return f


def apply_zones(case, mesh, ctrl):
    """Resolve the control's cell_zone lines (and VC tags of the mesh's zone
    names) and hand the block types and coefficient arrays to ``case``
    (anything with ``set_cell_types`` and ``set_field``).  A mesh without
    zone information and a control without cell_zone lines keep the fluid
    defaults and change nothing."""
    cell_zones = list(getattr(ctrl, "cell_zones", []) or [])
    tagged = any(vc_tag_type(c, u) for _, c, u in zone_table(mesh))
    if not cell_zones and not tagged:
        return None
    cztype = resolve_cell_zones(mesh, cell_zones)
    fields = setup_porous_fields(mesh, cell_zones, cztype)
    case.set_cell_types(cztype)
    for k in FIELDS:
        case.set_field(k, fields[k])
    return cztype, fields


# ------------------------------------------------------------- packed beds
# Per-cell coefficient fields for a packed bed of spheres of diameter dp at
# local porosity eps (the reference takes per-zone constants; these closures
# produce the per-cell arrays of a graded bed):
#   Kozeny-Carman permeability  K   = eps^3 dp^2 / (180 (1 - eps)^2)
#   Ergun inertial coefficient  C_F = 1.75 (1 - eps) / (eps^3 dp)   [1/m]
#     (momentum sink rho C_F |u| u, the reference's 'inertial')
#   specific fluid-solid area   a_sf = 6 (1 - eps) / dp              [1/m]
def kozeny_carman_permeability(eps, dp):
    eps = np.asarray(eps, np.float64)
    return eps ** 3 * dp ** 2 / (180.0 * (1.0 - eps) ** 2)


def ergun_inertial(eps, dp):
    eps = np.asarray(eps, np.float64)
    return 1.75 * (1.0 - eps) / (eps ** 3 * dp)


def specific_surface(eps, dp):
    eps = np.asarray(eps, np.float64)
    return 6.0 * (1.0 - eps) / dp


def packed_bed_fields(eps, dp, k_s=0.0, cp_s=0.0, rho_s=0.0, h_sf=0.0):
    """Per-cell coefficient arrays (names of ``Case.set_field``) of a packed
    bed with local porosity eps (n,) and particle diameter dp."""
    eps = np.asarray(eps, np.float64)
    n = eps.size
    K = kozeny_carman_permeability(eps, dp)
    return dict(porosity=eps.copy(), perm_dir=np.repeat(K[:, None], 3, axis=1), inertial=ergun_inertial(eps, dp),
                k_s=np.full(n, float(k_s)), cp_s=np.full(n, float(cp_s)), rho_s=np.full(n, float(rho_s)),
                h_sf=np.full(n, float(h_sf)), a_sf=specific_surface(eps, dp))
