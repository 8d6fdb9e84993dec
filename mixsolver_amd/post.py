"""Post-processing on host copies of the fields.

``probe_velocity`` and ``ghia_compare`` restate the reference's centreline
check of the lid-driven cavity (mod_uns_output.f90:246-284 and :291-350):
velocity at a point = inverse-squared-distance weighting of the 8 nearest
cell centroids (weight 1/(d^2 + 1e-12)); the u-velocity on x = 0.5, z = 0.5
at the 17 stations of Ghia, Ghia & Shin (1982, JCP 48:387-411, Tables I/II)
is compared with the published values, and the relative L2 error over the
15 interior stations is returned.
"""
from __future__ import annotations

import numpy as np

GHIA_Y = np.array([1.0000, 0.9766, 0.9688, 0.9609, 0.9531, 0.8516, 0.7344, 0.6172, 0.5000, 0.4531,
                   0.2813, 0.1719, 0.1016, 0.0703, 0.0625, 0.0547, 0.0000])
GHIA_U = {
    100: np.array([1.00000, 0.84123, 0.78871, 0.68722, 0.60751, 0.23151, -0.03393, -0.14320, -0.20533,
                   -0.21090, -0.15662, -0.10150, -0.06434, -0.04775, -0.04192, -0.03717, 0.00000]),
    1000: np.array([1.00000, 0.65928, 0.57492, 0.51117, 0.46604, 0.33304, 0.18719, 0.05702, -0.06080,
                    -0.10648, -0.27805, -0.38289, -0.29730, -0.22220, -0.20196, -0.18109, 0.00000]),
}


def probe_velocity(xc, u, xq, knn=8):
    """Velocity at xq from the knn nearest cell centroids (ties by cell index)."""
    d2 = ((np.asarray(xc) - np.asarray(xq)) ** 2).sum(axis=1)
    idx = np.argsort(d2, kind="stable")[:knn]
    w = 1.0 / (d2[idx] + 1.0e-12)
    return (w[:, None] * np.asarray(u)[idx]).sum(axis=0) / w.sum()


def ghia_compare(xc, u, re):
    """Relative L2 error of the centreline u_x against Ghia et al. (interior
    stations), and the probed values at all 17 stations."""
    uref = GHIA_U[int(round(re))]
    uq = np.array([probe_velocity(xc, u, (0.5, y, 0.5))[0] for y in GHIA_Y])
    err = uq - uref
    l2 = np.sqrt((err[1:-1] ** 2).sum() / max((uref[1:-1] ** 2).sum(), np.finfo(float).tiny))
    return l2, uq
