"""Checks of the reference's shipped validation cases (test infrastructure).

``bj_two_layer`` and ``bj_profile`` restate the reference's check of its
Beavers-Joseph fluid/porous channel (cluster_cases/bj_porous/plot_bj.py:
reference() :62-92, extraction :128-146): a fluid layer over a Brinkman
porous layer of equal height H, driven by a body force fb, has the closed
form

  fluid  (0 <= y <= H):  u_f = -fb y^2 / (2 mu) + a1 y + a2,   u_f(H)  = 0
  porous (-H <= y <= 0): u_p = u_D + E e^(y/lam) + F e^(-y/lam), u_p(-H) = 0

with lam = sqrt(K), u_D = K fb / mu, equal shear stress at the interface
(a1 = (E - F) / lam) and, for alpha > 0, the stress-jump condition
a1 = (alpha / lam)(u_f(0) - u_p(0)), else continuity u_f(0) = u_p(0).
The numerical check takes the u_x column at the middle x station: the first
fluid cell above the interface (uF1) and the mean over the porous rows at
least lam away from the bed wall and the interface (uD).

``cylinder_check`` computes the indicators the reference reports for its Re=40
cylinder (cluster_cases/cylinder_Re40/check_cylinder.py:70-190) on cell values: the
wall ring (innermost cell of each of 80 angular bins), Cp = p / (rho U^2 / 2)
at the front stagnation point, at the base and its minimum, the near-wall
u_x sign changes (separation, measured from the front stagnation point), the
recirculation length on the wake axis (farthest x with u_x < 0 over 0.05-wide
x bins, cell nearest y = 0), the pressure drag of the ring, and the mean /
max |Cp - Cp_ref| against a tabulated Cp(theta) interpolated periodically.

``conduction_exact`` holds the closed-form solid temperatures of the
reference's pure-conduction family (cases/heat_conduction/
check_heat_conduction.py:50-90): a bar of length L between fixed ends
(linear), one fixed wall with every other face adiabatic (uniform), an
annulus between fixed radii (logarithmic) and a bar with a uniform source
q on [a, b) (two lines joined by a parabola; slope jumps q (x - a)/k inside,
ends fixed).

``channel_check`` restates the reference's check of its low-speed planar
channel (cluster_cases/channel_lowspeed/plot_channel.py: layer_weights()
:60-72, main() :74-139): per x layer of cell centres the bulk velocity and
mean pressure are width-weighted sums over the rows (midpoint edges, end
rows clipped to the walls), the developed profile at x = 299.5 is compared
with u = 6 U eta (1 - eta), U = Re_H mu / (rho H), the entrance length is the
first layer whose centreline u is within 1 % of 1.5 U, and dp/dx is the
least-squares slope of the layer pressure over x in [200, 295]
(exact -12 mu U / H^2).
"""
from __future__ import annotations

import numpy as np


def bj_two_layer(alpha, H, K, mu, fb):
    """(u_f(y), u_p(y), u_f(0)) of the two-layer closed form, y measured
    from the interface."""
    lam = np.sqrt(K)
    uD = K * fb / mu
    A = np.array([[H, 1.0, 0.0, 0.0],
                  [0.0, 0.0, np.exp(-H / lam), np.exp(H / lam)],
                  [1.0, 0.0, -1.0 / lam, 1.0 / lam],
                  [1.0, -alpha / lam, alpha / lam, alpha / lam] if alpha > 0.0 else [0.0, 1.0, -1.0, -1.0]])
    b = np.array([fb * H * H / (2.0 * mu), -uD, 0.0, -alpha / lam * uD if alpha > 0.0 else uD])
    a1, a2, E, F = np.linalg.solve(A, b)

    def uf(y):
        y = np.asarray(y, np.float64)
        return -fb * y * y / (2.0 * mu) + a1 * y + a2

    def up(y):
        y = np.asarray(y, np.float64)
        return uD + E * np.exp(y / lam) + F * np.exp(-y / lam)

    return uf, up, a2


def bj_classical_slip(alpha, H, K, mu, fb):
    """Interface speed of the classical Beavers-Joseph slip condition over a
    pure-Darcy porous side (for comparison only)."""
    lam = np.sqrt(K)
    uD = K * fb / mu
    return (fb * H * H / (2.0 * mu) + alpha * (H / lam) * uD) / (1.0 + alpha * H / lam)


def bj_profile(xc, u, lx, ly, nx, ny, K):
    """u_x column at x station nx // 2 of an nx x ny channel of size lx x ly
    (bed in the lower half): (column, uF1, uD)."""
    xc = np.asarray(xc).reshape(-1, 3)
    ux = np.asarray(u).reshape(-1, 3)[:, 0]
    dx, dy = lx / nx, ly / ny
    ix = np.round((xc[:, 0] - dx / 2) / dx).astype(int)
    iy = np.round((xc[:, 1] - dy / 2) / dy).astype(int)
    sel = ix == nx // 2
    col = np.array([ux[sel & (iy == j)].mean() for j in range(ny)])
    lam = np.sqrt(K)
    r0, r1 = int(np.ceil(lam / dy)), ny // 2 - int(np.ceil(lam / dy))
    return col, float(col[ny // 2]), float(col[r0:r1].mean())


def conduction_exact(kind, x, y, t_cold=300.0, t_hot=500.0, length=0.1, k=15.0, q=1.0e6, a=0.03, b=0.07,
                     ri=0.01, ro=0.03, t_uniform=400.0):
    """Closed-form steady solid temperature at (x, y): kind 'linear',
    'uniform', 'annulus' or 'source'."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    if kind == "linear":
        return t_cold + (t_hot - t_cold) * x / length
    if kind == "uniform":
        return np.full_like(x, t_uniform)
    if kind == "annulus":
        return t_cold + (t_hot - t_cold) * np.log(np.hypot(x, y) / ri) / np.log(ro / ri)
    qk = q / k
    # T = Al x + Bl on [0, a), -qk x^2 / 2 + Am x + Bm on [a, b), Ar x + Br on [b, L]:
    # value and slope continuous at a and b, T(0) = t_cold, T(L) = t_hot
    bl = t_cold
    bm = bl - 0.5 * qk * a * a
    br = bm + 0.5 * qk * b * b
    ar = (t_hot - br) / length
    am = ar + qk * b
    al = am - qk * a
    return np.where(x < a, al * x + bl, np.where(x < b, -0.5 * qk * x * x + am * x + bm, ar * x + br))


def _first_per_group(group, key):
    """Index of the smallest key in every group (lowest index on ties),
    groups in ascending order."""
    order = np.lexsort((np.arange(len(key)), key, group))
    g = group[order]
    return order[np.concatenate(([True], g[1:] != g[:-1]))]


def cylinder_check(xc, u, p, r_cyl=0.5, u_inf=1.0, rho=1.0, nbins=80, cp_ref=None):
    """Regression indicators of a 2-D cylinder in x-flow (dict).

    Wall ring: the angle theta = atan2(y, x) in degrees splits the cells into
    nbins equal bins over [-180, 180); the cell of smallest radius in each
    bin is the ring cell of that bin, and the ring is ordered by theta.
    phi = theta mod 360 - 180 is the angle from the front stagnation point
    (the free stream is +x).  Separation: consecutive ring cells (ordered by
    phi) whose u_x changes sign, reported as 180 - |phi| of the first of the
    two.  Wake axis: stations x0 = 0.5, 0.55, ... < 6; at each, the cell of
    smallest |y| among those with |x - x0| < 0.025; the recirculation length
    is the largest station x with u_x < 0, in diameters.  Pressure drag:
    -sum Cp cos(theta) R dtheta / D with dtheta the angle to the next ring
    cell (wrapping)."""
    xc = np.asarray(xc).reshape(-1, 3)
    u = np.asarray(u).reshape(-1, 3)
    p = np.asarray(p)
    x, y, ux = xc[:, 0], xc[:, 1], u[:, 0]
    qdyn = 0.5 * rho * u_inf ** 2
    theta = np.degrees(np.arctan2(y, x))
    radius = np.hypot(x, y)
    sector = np.clip(((theta + 180.0) / 360.0 * nbins).astype(int), 0, nbins - 1)
    ring = _first_per_group(sector, radius)
    ring = ring[np.argsort(theta[ring])]
    th_r, ux_r, cp_r = theta[ring], ux[ring], p[ring] / qdyn
    phi = np.mod(th_r, 360.0) - 180.0
    front, base = int(np.argmin(np.abs(phi))), int(np.argmax(np.abs(phi)))
    res = dict(ring_cells=len(ring), ring_r_mean=float(radius[ring].mean()), cp_front=float(cp_r[front]),
               cp_base=float(cp_r[base]), cp_min=float(cp_r.min()), phi_cp_min=float(phi[int(np.argmin(cp_r))]),
               umax=float(np.hypot(u[:, 0], u[:, 1]).max() / u_inf))
    by_phi = np.argsort(phi)
    s = np.sign(ux_r[by_phi])
    flips = np.nonzero(s[:-1] * s[1:] < 0)[0]
    res["separation_from_wake_axis"] = [float(180.0 - abs(phi[by_phi][k])) for k in flips]
    axis_x, axis_u = [], []
    for x0 in np.arange(0.5, 6.0, 0.05):
        near = np.nonzero((x > x0 - 0.025) & (x < x0 + 0.025))[0]
        if near.size:
            k = near[np.argmin(np.abs(y[near]))]
            axis_x.append(x[k])
            axis_u.append(ux[k])
    axis_x, axis_u = np.array(axis_x), np.array(axis_u)
    back = axis_u < 0.0
    res["wake_length_D"] = float(axis_x[back].max() / (2 * r_cyl)) if back.any() else 0.0
    res["max_reverse_ux"] = float(axis_u.min())
    dtheta = np.radians(np.diff(np.append(th_r, th_r[0] + 360.0)))
    res["cd_pressure"] = float(-(cp_r * np.cos(np.radians(th_r)) * r_cyl * dtheta).sum() / (2 * r_cyl))
    if cp_ref is not None:
        ang, cpv = np.asarray(cp_ref[0], float), np.asarray(cp_ref[1], float)
        ang = np.where(ang > 180.0, ang - 360.0, ang)
        k = np.argsort(ang)
        ang, cpv = ang[k], cpv[k]
        dev = np.abs(cp_r - np.interp(th_r, np.concatenate((ang - 360.0, ang, ang + 360.0)), np.tile(cpv, 3)))
        res["dcp_mean"] = float(dev.mean())
        res["dcp_max"] = float(dev.max())
    return res

def _layer_widths(y, height):
    """Control widths of cell-centre rows of one x layer: the midpoints
    between neighbouring centres bound each row, the outermost rows extend
    half their neighbour spacing and stop at the walls y = 0 and y = H."""
    o = np.argsort(y)
    ys = y[o]
    e = np.concatenate(([max(ys[0] - 0.5 * (ys[1] - ys[0]), 0.0)], 0.5 * (ys[1:] + ys[:-1]),
                        [min(ys[-1] + 0.5 * (ys[-1] - ys[-2]), height)]))
    w = np.empty_like(ys)
    w[o] = np.diff(e)
    return w


def channel_check(xc, u, p, height=10.0, re_h=100.0, rho=1.177, mu=1.846e-5, x_fit=(200.0, 295.0)):
    """Regression indicators of a planar Poiseuille channel (dict)."""
    xc = np.asarray(xc).reshape(-1, 3)
    u = np.asarray(u).reshape(-1, 3)
    p = np.asarray(p)
    x, y, ux = xc[:, 0], xc[:, 1], u[:, 0]
    ub = re_h * mu / (rho * height)
    xr = np.round(x, 9)
    lv = np.unique(xr)
    bulk, pm, ucl = np.empty(len(lv)), np.empty(len(lv)), np.empty(len(lv))
    rows = []
    for k, x0 in enumerate(lv):
        m = np.nonzero(np.abs(x - x0) < 1e-6)[0]
        w = _layer_widths(y[m], height)
        bulk[k] = np.dot(ux[m], w) / height
        pm[k] = np.dot(p[m], w) / height
        ucl[k] = ux[m][np.argmin(np.abs(y[m] - 0.5 * height))]
        rows.append(m)
    m = rows[int(np.argmin(np.abs(lv - 299.5)))]
    eta = y[m] / height
    err = np.abs(ux[m] - 6.0 * ub * eta * (1.0 - eta)) / ub
    band = np.abs(ucl - 1.5 * ub) / ub < 0.01
    f = (lv >= x_fit[0]) & (lv <= x_fit[1])
    slope = np.polyfit(lv[f], pm[f], 1)[0]
    mag = np.sqrt((u * u).sum(axis=1))
    jm = int(np.argmax(mag))
    return dict(u_bulk_target=ub, dpdx_exact=-12.0 * mu * ub / height ** 2, n_layers=len(lv),
                u_bulk_in=float(bulk[0]), u_bulk_50=float(bulk[50]), u_bulk_out=float(bulk[-1]),
                mass_drift=float((bulk[-1] - bulk[0]) / bulk[0]),
                profile_err_max=float(err.max()), profile_err_l1=float(err.mean()),
                umax_ratio_out=float(ux[m].max() / ub),
                entrance_length=float(lv[int(np.argmax(band))]) if band.any() else float("nan"),
                dpdx_fit=float(slope), p_rise=float(pm[0] - pm[-1]),
                p_min=float(p.min()), p_max=float(p.max()),
                umag_max=float(mag[jm]), umag_max_at=xc[jm].tolist())
