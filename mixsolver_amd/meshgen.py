"""Synthetic unstructured meshes in the owner/neighbour face-list layout.

The layout is the one ``read_cas`` fills in ``mesh_t`` (mod_uns_mesh.f90:17-54):
node coordinates, one record per face holding its node list, its owner cell
``c0`` (1-based), its neighbour ``c1`` (1-based, 0 on a boundary) and its face
zone id.  Face order is the file order of a Fluent ``.cas`` (interior zone
first, then one zone per boundary patch), which is the order every reference
face loop walks (mod_uns_cas_reader.f90:712-727).

Generators:
  * ``hex_box``         -- structured-as-unstructured hexahedra (cavity,
                           channel, serving boxes), optional node jitter;
  * ``cavity``          -- unit-square lid-driven cavity extruded one cell;
  * ``channel``         -- wall-normal tanh-stretched channel;
  * ``cylinder_prisms`` -- O-grid triangles around a cylinder, extruded into
                           wedge (prism) cells, optional seeded jitter;
  * ``perturbed_box``   -- seeded jittered hex or prism box used for
                           pressure-solve streams.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

HEX, WEDGE = 4, 6          # Fluent element type codes (section 12)


@dataclass
class Mesh:
    x: np.ndarray                      # (nnodes, 3) float64
    fptr: np.ndarray                   # (nfaces+1,) int64 offsets into fnodes
    fnodes: np.ndarray                 # (sum nn,) int32, 1-based node ids
    c0: np.ndarray                     # (nfaces,) int32, 1-based owner
    c1: np.ndarray                     # (nfaces,) int32, 1-based neighbour / 0
    zone: np.ndarray                   # (nfaces,) int32 face zone id
    ncells: int
    ctype: int = HEX
    zones: dict = field(default_factory=dict)   # id -> (kind, name)
    cell_zone: int = 2
    czone: np.ndarray | None = None    # (ncells,) cell zone id (None: all cell_zone)
    czone_names: dict = field(default_factory=dict)   # cell zone id -> (condition, name)

    @property
    def nnodes(self) -> int:
        return int(self.x.shape[0])

    @property
    def nfaces(self) -> int:
        return int(self.c0.shape[0])

    @property
    def n_interior(self) -> int:
        return int(np.count_nonzero(self.c1 > 0))

    def face_nn(self) -> np.ndarray:
        return np.diff(self.fptr).astype(np.int32)


    def cell_zones(self) -> np.ndarray:
        if self.czone is None:
            return np.full(self.ncells, self.cell_zone, np.int32)
        return np.asarray(self.czone, np.int32)


def with_cell_zones(m: Mesh, czone, names=None) -> Mesh:
    """Copy of m with cell zones: cells are renumbered so that every zone is
    one contiguous id range (zones in ascending id order, the previous order
    inside each), as the cell records of a .cas file require.  czone gives
    the zone id of every cell of m; names maps zone ids to (condition, name).
    The copy's ``cell_perm`` holds the old 0-based id of every new cell."""
    czone = np.asarray(czone, np.int32)
    order = np.argsort(czone, kind="stable")          # new -> old
    o2n = np.empty_like(order)
    o2n[order] = np.arange(len(order))
    c0 = (o2n[m.c0 - 1] + 1).astype(np.int32)
    c1 = np.where(m.c1 > 0, o2n[np.maximum(m.c1, 1) - 1] + 1, 0).astype(np.int32)
    out = Mesh(x=m.x, fptr=m.fptr, fnodes=m.fnodes, c0=c0, c1=c1, zone=m.zone, ncells=m.ncells,
               ctype=m.ctype, zones=dict(m.zones), cell_zone=int(czone[order][0]), czone=czone[order],
               czone_names=dict(names or {}))
    out.cell_perm = order
    return out


# ---------------------------------------------------------------------------
# hexahedral boxes
# ---------------------------------------------------------------------------
def _hex_nodes(xs, ys, zs):
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    # node (i,j,k) -> i + (nx+1)*(j + (ny+1)*k): i fastest
    return np.stack([X.transpose(2, 1, 0).ravel(),
                     Y.transpose(2, 1, 0).ravel(),
                     Z.transpose(2, 1, 0).ravel()], axis=1).astype(np.float64)


def hex_box(xs, ys, zs, zone_ids=None, jitter=0.0, seed=0, periodic=(), tile_x=None):
    """Hexahedral box on the tensor grid xs x ys x zs.

    zone_ids maps side names ('xmin','xmax','ymin','ymax','zmin','zmax') to
    face zone ids; the interior zone is 3 and the cell zone 2, as in the
    reference's shipped cavity mesh.  Interior faces are emitted per owner
    cell in natural cell order (+x, +y, +z), then one block per boundary side.
    ``jitter`` moves interior nodes by a seeded uniform fraction of the local
    spacing in every direction.  Axes named in ``periodic`` ('x', 'z') get
    the same jitter on their two end planes, so the end faces are exact
    translates of each other.  ``tile_x`` = m repeats the jitter of an m-cell
    wide box along x (the jitter of node i is that of node i mod m), so a box
    of k*m uniform x cells is k exact copies of the m-cell box.
    """
    xs = np.asarray(xs, np.float64)
    ys = np.asarray(ys, np.float64)
    zs = np.asarray(zs, np.float64)
    nx, ny, nz = len(xs) - 1, len(ys) - 1, len(zs) - 1
    zid = dict(xmin=5, xmax=5, ymin=5, ymax=5, zmin=4, zmax=4)
    if zone_ids:
        zid.update(zone_ids)
    x = _hex_nodes(xs, ys, zs)
    if jitter > 0.0:
        rng = np.random.default_rng(seed)
        N1, N2 = nx + 1, ny + 1
        idx = np.arange(x.shape[0])
        i = idx % N1
        j = (idx // N1) % N2
        k = idx // (N1 * N2)
        dx = np.minimum(np.diff(xs)[np.clip(i - 1, 0, nx - 1)], np.diff(xs)[np.clip(i, 0, nx - 1)])
        dy = np.minimum(np.diff(ys)[np.clip(j - 1, 0, ny - 1)], np.diff(ys)[np.clip(j, 0, ny - 1)])
        dz = np.minimum(np.diff(zs)[np.clip(k - 1, 0, nz - 1)], np.diff(zs)[np.clip(k, 0, nz - 1)])
        m1 = (tile_x + 1) if tile_x else N1
        r = rng.uniform(-jitter, jitter, size=(m1 * N2 * (nz + 1), 3))
        src_i = i % tile_x if tile_x else i
        src_i = np.where(i == nx, 0, src_i) if "x" in periodic else src_i
        src_k = np.where(k == nz, 0, k) if "z" in periodic else k
        r = r[src_i + m1 * (j + N2 * src_k)]
        inner_x = (i > 0) & (i < nx) & ((i % tile_x > 0) if tile_x else True)
        inner_y = (j > 0) & (j < ny)
        inner_z = (k > 0) & (k < nz)
        x[:, 0] += np.where(inner_x, r[:, 0] * dx, 0.0)
        x[:, 1] += np.where(inner_y, r[:, 1] * dy, 0.0)
        x[:, 2] += np.where(inner_z, r[:, 2] * dz, 0.0)

    def node(i, j, k):
        return (i + (nx + 1) * (j + (ny + 1) * k) + 1).astype(np.int32)

    def cell(i, j, k):
        return (i + nx * (j + ny * k) + 1).astype(np.int32)

    K, J, I = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    I, J, K = I.ravel(), J.ravel(), K.ravel()            # natural cell order

    recs = []   # (owner, sub, nodes(4), c0, c1)
    m = I < nx - 1                                         # +x interior
    i, j, k = I[m], J[m], K[m]
    recs.append((cell(i, j, k), 0,
                 np.stack([node(i + 1, j, k), node(i + 1, j + 1, k),
                           node(i + 1, j + 1, k + 1), node(i + 1, j, k + 1)], 1),
                 cell(i, j, k), cell(i + 1, j, k)))
    m = J < ny - 1                                         # +y interior
    i, j, k = I[m], J[m], K[m]
    recs.append((cell(i, j, k), 1,
                 np.stack([node(i, j + 1, k), node(i, j + 1, k + 1),
                           node(i + 1, j + 1, k + 1), node(i + 1, j + 1, k)], 1),
                 cell(i, j, k), cell(i, j + 1, k)))
    m = K < nz - 1                                         # +z interior
    i, j, k = I[m], J[m], K[m]
    recs.append((cell(i, j, k), 2,
                 np.stack([node(i, j, k + 1), node(i + 1, j, k + 1),
                           node(i + 1, j + 1, k + 1), node(i, j + 1, k + 1)], 1),
                 cell(i, j, k), cell(i, j, k + 1)))
    own = np.concatenate([r[0] for r in recs])
    sub = np.concatenate([np.full(r[0].shape, r[1]) for r in recs])
    order = np.lexsort((sub, own))
    inodes = np.concatenate([r[2] for r in recs])[order]
    ic0 = np.concatenate([r[3] for r in recs])[order]
    ic1 = np.concatenate([r[4] for r in recs])[order]

    bn, bc0, bz = [], [], []

    def side(mask, nodes, name):
        i, j, k = I[mask], J[mask], K[mask]
        bn.append(nodes(i, j, k))
        bc0.append(cell(i, j, k))
        bz.append(np.full(i.shape, zid[name], np.int32))

    side(I == 0, lambda i, j, k: np.stack([node(i, j, k), node(i, j, k + 1),
                                           node(i, j + 1, k + 1), node(i, j + 1, k)], 1), "xmin")
    side(I == nx - 1, lambda i, j, k: np.stack([node(i + 1, j, k), node(i + 1, j + 1, k),
                                                node(i + 1, j + 1, k + 1), node(i + 1, j, k + 1)], 1), "xmax")
    side(J == 0, lambda i, j, k: np.stack([node(i, j, k), node(i + 1, j, k),
                                           node(i + 1, j, k + 1), node(i, j, k + 1)], 1), "ymin")
    side(J == ny - 1, lambda i, j, k: np.stack([node(i, j + 1, k), node(i, j + 1, k + 1),
                                                node(i + 1, j + 1, k + 1), node(i + 1, j + 1, k)], 1), "ymax")
    side(K == 0, lambda i, j, k: np.stack([node(i, j, k), node(i, j + 1, k),
                                           node(i + 1, j + 1, k), node(i + 1, j, k)], 1), "zmin")
    side(K == nz - 1, lambda i, j, k: np.stack([node(i, j, k + 1), node(i + 1, j, k + 1),
                                                node(i + 1, j + 1, k + 1), node(i, j + 1, k + 1)], 1), "zmax")

    # boundary faces grouped by zone id (ascending), keeping side order inside
    bnodes = np.concatenate(bn)
    bown = np.concatenate(bc0)
    bzone = np.concatenate(bz)
    border = np.argsort(bzone, kind="stable")
    fn = np.concatenate([inodes, bnodes[border]]).astype(np.int32)
    c0 = np.concatenate([ic0, bown[border]]).astype(np.int32)
    c1 = np.concatenate([ic1, np.zeros(bown.shape[0], np.int32)]).astype(np.int32)
    zone = np.concatenate([np.full(ic0.shape[0], 3, np.int32), bzone[border]])
    nf = c0.shape[0]
    fptr = np.arange(0, 4 * nf + 1, 4, dtype=np.int64)
    zones = {3: ("interior", "interior")}
    for name, z in zid.items():
        zones.setdefault(int(z), ("boundary", name))
    return Mesh(x=x, fptr=fptr, fnodes=fn.ravel(), c0=c0, c1=c1, zone=zone,
                ncells=nx * ny * nz, ctype=HEX, zones=zones)


def cavity(n=128, nz=1):
    """Unit-square lid-driven cavity, n x n uniform quads extruded to z in [0,1].

    Zones: 2 fluid cells, 3 interior, 4 symmetry (z planes), 5 wall (x/y walls).
    """
    xs = np.linspace(0.0, 1.0, n + 1)
    zs = np.linspace(0.0, 1.0, nz + 1)
    return hex_box(xs, xs, zs)


def tanh_stretch(n, ratio=1.1, lo=-1.0, hi=1.0):
    """n cells on [lo,hi] clustered at both ends by y = tanh(g*s)/tanh(g),
    s in [-1,1]; g is chosen so the largest adjacent-spacing ratio is `ratio`."""
    s = np.linspace(-1.0, 1.0, n + 1)

    def worst(g):
        y = np.tanh(g * s) / np.tanh(g)
        d = np.diff(y)
        return float(np.max(np.maximum(d[1:] / d[:-1], d[:-1] / d[1:])))

    a, b = 1e-6, 10.0
    for _ in range(200):
        mid = 0.5 * (a + b)
        if worst(mid) > ratio:
            b = mid
        else:
            a = mid
    y = np.tanh(a * s) / np.tanh(a)
    return lo + (y + 1.0) * 0.5 * (hi - lo)


def channel(nx=400, ny=100, nz=50, lx=4 * np.pi, lz=2 * np.pi, ratio=1.1, periodic=False, jitter=0.0, seed=0):
    """Plane channel [0,lx] x [-1,1] x [0,lz] with tanh wall clustering.

    Zones: 3 interior, 6 inlet (x=0), 7 outlet (x=lx), 5 walls (y=+-1),
    4 side planes (z=0, z=lz).  periodic=True gives the z planes their own
    zones, 8 (z=0) and 9 (z=lz), for periodic pairs 7-6 and 9-8.
    """
    xs = np.linspace(0.0, lx, nx + 1)
    ys = tanh_stretch(ny, ratio)
    zs = np.linspace(0.0, lz, nz + 1)
    z = dict(xmin=6, xmax=7, ymin=5, ymax=5, zmin=8, zmax=9) if periodic else \
        dict(xmin=6, xmax=7, ymin=5, ymax=5, zmin=4, zmax=4)
    return hex_box(xs, ys, zs, zone_ids=z, jitter=jitter, seed=seed, periodic=("x", "z") if periodic else ())


# ---------------------------------------------------------------------------
# extruded triangle (wedge) meshes
# ---------------------------------------------------------------------------
def _extrude_triangles(p2, tris, edge_zone_of, nl, lz, zbot=4, ztop=4):
    """Extrude a 2D triangulation (p2: (np,2), tris: (nt,3) 0-based CCW) into
    nl layers of wedges.  edge_zone_of(a, b) -> zone id for a boundary edge.

    Cells: cell(t, l) = t + nt*l + 1.  Interior faces first (per cell, in cell
    order: lateral interior edges owned by the cell, then its top triangle if
    not the last layer), then boundary faces grouped by zone id.
    """
    npt, nt = p2.shape[0], tris.shape[0]
    zs = np.linspace(0.0, lz, nl + 1)
    x = np.concatenate([np.column_stack([p2, np.full(npt, z)]) for z in zs])

    # unique edges with owner / neighbour triangles
    e = np.concatenate([tris[:, [0, 1]], tris[:, [1, 2]], tris[:, [2, 0]]])
    et = np.concatenate([np.arange(nt)] * 3)
    key = np.sort(e, axis=1)
    order = np.lexsort((key[:, 1], key[:, 0]))
    key, e, et = key[order], e[order], et[order]
    same = np.all(key[1:] == key[:-1], axis=1)
    first = np.ones(key.shape[0], bool)
    first[1:] = ~same
    starts = np.flatnonzero(first)
    counts = np.diff(np.append(starts, key.shape[0]))
    if np.any(counts > 2):
        raise ValueError("non-manifold triangulation")
    ea, eb = e[starts, 0], e[starts, 1]
    t0 = et[starts]
    t1 = np.where(counts == 2, et[np.minimum(starts + 1, key.shape[0] - 1)], -1)
    # owner = lower triangle id
    swap = (t1 >= 0) & (t1 < t0)
    t0, t1 = np.where(swap, t1, t0), np.where(swap, t0, t1)
    ea2 = np.where(swap, eb, ea)
    eb2 = np.where(swap, ea, eb)
    ea, eb = ea2, eb2

    def node(p, l):
        return (p + npt * l + 1).astype(np.int32)

    def cell(t, l):
        return (t + nt * l + 1).astype(np.int32)

    owners, subs, nlist, fc0, fc1 = [], [], [], [], []
    inner = t1 >= 0
    for l in range(nl):
        a, b = ea[inner], eb[inner]
        owners.append(cell(t0[inner], l))
        subs.append(np.zeros(a.shape[0], np.int64))
        nlist.append(np.stack([node(a, l), node(b, l), node(b, l + 1), node(a, l + 1)], 1))
        fc0.append(cell(t0[inner], l))
        fc1.append(cell(t1[inner], l))
    for l in range(nl - 1):
        t = np.arange(nt)
        owners.append(cell(t, l))
        subs.append(np.ones(nt, np.int64))
        tri = tris
        nlist.append(np.stack([node(tri[:, 0], l + 1), node(tri[:, 1], l + 1),
                               node(tri[:, 2], l + 1), np.full(nt, -1, np.int32)], 1))
        fc0.append(cell(t, l))
        fc1.append(cell(t, l + 1))
    own = np.concatenate(owners)
    sub = np.concatenate(subs)
    order = np.lexsort((sub, own))
    inodes = np.concatenate(nlist)[order]
    ic0 = np.concatenate(fc0)[order]
    ic1 = np.concatenate(fc1)[order]

    bnl, bc0, bz = [], [], []
    bnd = ~inner
    bzone_edge = np.array([edge_zone_of(int(a), int(b)) for a, b in zip(ea[bnd], eb[bnd])], np.int32) \
        if np.any(bnd) else np.zeros(0, np.int32)
    for l in range(nl):
        a, b = ea[bnd], eb[bnd]
        bnl.append(np.stack([node(a, l), node(b, l), node(b, l + 1), node(a, l + 1)], 1))
        bc0.append(cell(t0[bnd], l))
        bz.append(bzone_edge)
    t = np.arange(nt)
    bnl.append(np.stack([node(tris[:, 0], 0), node(tris[:, 2], 0), node(tris[:, 1], 0),
                         np.full(nt, -1, np.int32)], 1))
    bc0.append(cell(t, 0))
    bz.append(np.full(nt, zbot, np.int32))
    bnl.append(np.stack([node(tris[:, 0], nl), node(tris[:, 1], nl), node(tris[:, 2], nl),
                         np.full(nt, -1, np.int32)], 1))
    bc0.append(cell(t, nl - 1))
    bz.append(np.full(nt, ztop, np.int32))
    bnodes = np.concatenate(bnl)
    bown = np.concatenate(bc0)
    bzone = np.concatenate(bz)
    border = np.argsort(bzone, kind="stable")
    allnodes = np.concatenate([inodes, bnodes[border]])
    c0 = np.concatenate([ic0, bown[border]]).astype(np.int32)
    c1 = np.concatenate([ic1, np.zeros(bown.shape[0], np.int32)]).astype(np.int32)
    zone = np.concatenate([np.full(ic0.shape[0], 3, np.int32), bzone[border]])
    nn = np.where(allnodes[:, 3] < 0, 3, 4)
    fptr = np.zeros(c0.shape[0] + 1, np.int64)
    fptr[1:] = np.cumsum(nn)
    fnodes = allnodes[allnodes >= 0].astype(np.int32)     # row-major keeps order
    return Mesh(x=x, fptr=fptr, fnodes=fnodes, c0=c0, c1=c1, zone=zone,
                ncells=nt * nl, ctype=WEDGE)


def cylinder_prisms(ntheta=1000, nr=300, r_out=20.0, d=1.0, layers=2, lz=0.2,
                    jitter=0.15, seed=200):
    """O-grid around a cylinder of diameter d, each quad split into two
    triangles, extruded `layers` layers over lz.  Radial geometric growth from
    r = d/2 to r_out.  Interior 2D vertices are moved by a seeded uniform
    jitter of +-`jitter` times the local minimum edge length.

    Zones: 3 interior, 5 cylinder wall, 6 far field, 4 z planes.
    """
    r0 = 0.5 * d
    # geometric radial spacing with nr cells from r0 to r_out
    q = 1.05
    for _ in range(200):   # solve r0*(q^nr - 1)/(q - 1)*h0 = r_out - r0 with h0 = 2*pi*r0/ntheta
        h0 = 2 * np.pi * r0 / ntheta
        f = h0 * (q ** nr - 1) / (q - 1) - (r_out - r0)
        df = h0 * (nr * q ** (nr - 1) * (q - 1) - (q ** nr - 1)) / (q - 1) ** 2
        q = max(1.0 + 1e-9, q - f / df)
    h = h0 * q ** np.arange(nr)
    r = np.concatenate([[r0], r0 + np.cumsum(h)])
    r[-1] = r_out
    th = np.linspace(0.0, 2 * np.pi, ntheta, endpoint=False)
    R, T = np.meshgrid(r, th, indexing="ij")             # (nr+1, ntheta)
    p2 = np.column_stack([(R * np.cos(T)).ravel(), (R * np.sin(T)).ravel()])

    def pid(j, i):
        return j * ntheta + (i % ntheta)

    J, I = np.meshgrid(np.arange(nr), np.arange(ntheta), indexing="ij")
    J, I = J.ravel(), I.ravel()
    a, b, c, dd = pid(J, I), pid(J, I + 1), pid(J + 1, I + 1), pid(J + 1, I)
    alt = ((I + J) % 2) == 0
    t1 = np.where(alt[:, None], np.stack([a, b, c], 1), np.stack([a, b, dd], 1))
    t2 = np.where(alt[:, None], np.stack([a, c, dd], 1), np.stack([b, c, dd], 1))
    tris = np.stack([t1, t2], 1).reshape(-1, 3).astype(np.int64)
    if jitter > 0.0:
        rng = np.random.default_rng(seed)
        jj = np.arange(p2.shape[0]) // ntheta
        hloc = np.minimum(np.diff(r)[np.clip(jj - 1, 0, nr - 1)],
                          np.diff(r)[np.clip(jj, 0, nr - 1)])
        hloc = np.minimum(hloc, 2 * np.pi * r[jj] / ntheta)
        mv = rng.uniform(-jitter, jitter, size=p2.shape) * hloc[:, None]
        inner = (jj > 0) & (jj < nr)
        p2 = p2 + np.where(inner[:, None], mv, 0.0)

    wall_ids = set(range(ntheta))

    def ez(a, b):
        return 5 if (a in wall_ids and b in wall_ids) else 6

    m = _extrude_triangles(p2, tris, ez, layers, lz)
    m.zones = {3: ("interior", "interior"), 4: ("boundary", "zplanes"),
               5: ("boundary", "cylinder"), 6: ("boundary", "farfield")}
    return m


def perturbed_box(ncells_target, kind="hex", jitter=0.10, seed=0):
    """Seeded jittered unit box with about ncells_target cells.

    kind 'hex' gives n^3 hexahedra; 'prism' gives a jittered right-triangle
    grid extruded into wedges.  All outer faces are zone 5.
    """
    if kind == "hex":
        n = max(2, int(round(ncells_target ** (1.0 / 3.0))))
        g = np.linspace(0.0, 1.0, n + 1)
        return hex_box(g, g, g, zone_ids=dict(xmin=5, xmax=5, ymin=5, ymax=5, zmin=5, zmax=5),
                       jitter=jitter, seed=seed)
    n = max(2, int(round((ncells_target / 2.0) ** (1.0 / 3.0))))
    nl = n
    g = np.linspace(0.0, 1.0, n + 1)
    X, Y = np.meshgrid(g, g, indexing="ij")
    p2 = np.column_stack([X.T.ravel(), Y.T.ravel()])
    rng = np.random.default_rng(seed)
    ii = np.arange(p2.shape[0]) % (n + 1)
    jj = np.arange(p2.shape[0]) // (n + 1)
    inner = (ii > 0) & (ii < n) & (jj > 0) & (jj < n)
    p2 = p2 + np.where(inner[:, None], rng.uniform(-jitter, jitter, p2.shape) / n, 0.0)
    J, I = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    J, I = J.ravel(), I.ravel()
    a = J * (n + 1) + I
    b, c, d = a + 1, a + n + 2, a + n + 1
    tris = np.stack([np.stack([a, b, c], 1), np.stack([a, c, d], 1)], 1).reshape(-1, 3)
    m = _extrude_triangles(p2, tris, lambda a, b: 5, nl, 1.0, zbot=5, ztop=5)
    m.zones = {3: ("interior", "interior"), 5: ("boundary", "walls")}
    return m
