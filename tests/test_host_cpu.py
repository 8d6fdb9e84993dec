"""Host-side pieces that need no GPU: the C-ABI library loads and exports
every declared entry point, the control parser follows the reference's
grammar, the mesh generators produce valid owner/neighbour meshes and the
.cas writer/reader round-trips them."""
import ctypes as C
import os
import shutil

import numpy as np
import pytest

from mixsolver_amd import _lib, cas, meshgen
from mixsolver_amd.control import BC_OUTFLOW, BC_VINLET_PARAB, BC_WALL, ControlError, parse_control


def test_library_exports_every_header_symbol():
    L = C.CDLL(_lib.LIB_PATH)
    syms = _lib.header_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    # the binding's signature table covers every declared entry point
    _lib.lib()


def test_control_grammar():
    ctrl, groups = parse_control("""
rho = 1.2
mu = 1e-3
ppe_precond = ic0
transient = true
time_scheme = 2
body_force = 1.0 0.0 0.0
bc = 6 velocity-inlet-parabolic 15.63 2.0 2 -1.0
bc = 7 outflow
bc = 5 wall
lid = 5 2 1.0 1.0 0.0 0.0
tbc = 5 1 300.0
some_new_key = 3
""", warn=False)
    assert ctrl.rho == 1.2 and ctrl.mu == 1e-3 and ctrl.ppe_precond == 1 and ctrl.gpu_ppe_precond == 1
    assert ctrl.transient == 1 and ctrl.time_scheme == 2 and list(ctrl.body_force) == [1.0, 0.0, 0.0]
    assert [g.btype for g in groups] == [BC_VINLET_PARAB, BC_OUTFLOW, BC_WALL]
    assert groups[0].prfl_mean == 15.63 and groups[0].prfl_axis == 2 and groups[0].prfl_origin == -1.0
    assert groups[2].has_lid == 1 and groups[2].lid_dir == 2 and groups[2].ttype == 1 and groups[2].tval == 300.0
    # any ppe_precond other than ic0 selects Jacobi (mod_uns_simple_mpi.f90:437-444)
    c2, _ = parse_control("ppe_precond = amg\nbc = 1 wall\n", warn=False)
    assert c2.ppe_precond == 0
    with pytest.raises(ControlError) as e:
        parse_control("rho = 1\n", warn=False)
    assert e.value.code == 5
    with pytest.raises(ControlError) as e:
        parse_control("bc = 3 nonsense\n", warn=False)
    assert e.value.code == 7
    with pytest.raises(ControlError) as e:
        parse_control("bc = 3 wall\nlid = 4 2 1.0 1 0 0\n", warn=False)
    assert e.value.code == 8


@pytest.mark.parametrize("mk", [lambda: meshgen.cavity(8),
                                lambda: meshgen.channel(6, 5, 4),
                                lambda: meshgen.cylinder_prisms(12, 4, 3.0, layers=2),
                                lambda: meshgen.perturbed_box(500, "hex", seed=1),
                                lambda: meshgen.perturbed_box(500, "prism", seed=2)])
def test_meshgen_valid(mk):
    m = mk()
    nc = m.ncells
    assert m.c0.min() >= 1 and m.c0.max() <= nc
    assert m.c1.min() >= 0 and m.c1.max() <= nc
    # interior faces first, each unordered cell pair at most once
    inner = m.c1 > 0
    assert np.all(inner[: m.n_interior])
    pairs = np.sort(np.stack([m.c0[inner], m.c1[inner]], 1), axis=1)
    assert len(np.unique(pairs, axis=0)) == len(pairs)
    # every cell is closed: count faces per cell
    cnt = np.bincount(m.c0 - 1, minlength=nc) + np.bincount(m.c1[inner] - 1, minlength=nc)
    assert cnt.min() >= 5


def test_cas_roundtrip(tmp_path):
    m = meshgen.cylinder_prisms(10, 3, 2.0, layers=2, jitter=0.1, seed=3)
    p = os.path.join(tmp_path, "m.cas")
    cas.write_cas(p, m, {5: "wall", 6: "pressure-far-field", 4: "symmetry"})
    r = cas.read_cas(p)
    assert r.ncells == m.ncells
    assert np.array_equal(r.x, m.x)
    assert np.array_equal(r.fptr, m.fptr) and np.array_equal(r.fnodes, m.fnodes)
    assert np.array_equal(r.c0, m.c0) and np.array_equal(r.c1, m.c1) and np.array_equal(r.zone, m.zone)


def test_tanh_stretch_ratio():
    y = meshgen.tanh_stretch(100, 1.1)
    d = np.diff(y)
    assert abs(y[0] + 1) < 1e-15 and abs(y[-1] - 1) < 1e-15
    r = np.maximum(d[1:] / d[:-1], d[:-1] / d[1:])
    assert abs(r.max() - 1.1) < 1e-6


def _grid_centres(nx, ny, nz, h=(1.0, 1.0, 1.0)):
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return np.stack([(i.ravel() + 0.5) * h[0], (j.ravel() + 0.5) * h[1], (k.ravel() + 0.5) * h[2]], 1)


def test_coarse_aggregates_closed_form():
    """On a uniform grid, bisection along the longest side with count-weighted
    cuts gives boxes: 8x4x2 cells into 4 -> x-quarters (x is split twice, the
    x/y tie goes to x); 8x8x8 into 8 -> octants."""
    from mixsolver_amd.solver import coarse_aggregates
    xc = _grid_centres(8, 4, 2)
    i = np.arange(len(xc)) % 8
    a = coarse_aggregates(xc, 4)
    assert np.array_equal(a, i // 2)
    xc = _grid_centres(8, 8, 8)
    a = coarse_aggregates(xc, 8)
    i, j, k = np.arange(512) % 8, (np.arange(512) // 8) % 8, np.arange(512) // 64
    octant = (i // 4) * 4 + (j // 4) * 2 + (k // 4)   # split order x, then y, then z
    # aggregates are numbered by the split tree: x halves, then y, then z
    assert np.array_equal(a, octant)


def test_coarse_aggregates_balanced_deterministic():
    from mixsolver_amd.solver import coarse_aggregates
    rng = np.random.default_rng(5)
    xc = rng.uniform(size=(5000, 3)) * [4.0, 1.0, 2.0]
    for k in (1, 3, 37, 256):
        a = coarse_aggregates(xc, k)
        cnt = np.bincount(a, minlength=k)
        assert len(cnt) == k and cnt.max() - cnt.min() <= 1, k
        assert np.array_equal(a, coarse_aggregates(xc, k))
    # every cell its own aggregate
    assert np.array_equal(np.sort(coarse_aggregates(xc[:50], 50)), np.arange(50))
    # compact: 64 aggregates of uniform points have small boxes
    a = coarse_aggregates(xc, 64)
    ext = np.array([np.ptp(xc[a == g], axis=0).prod() for g in range(64)])
    assert ext.max() < 4.0 * (8.0 / 64)


def _cell_graph(m):
    """1-based CSR pattern (diagonal included) of a mesh's cell graph and the
    interior face pairs (0-based)."""
    inner = m.c1 > 0
    a, b = m.c0[inner] - 1, m.c1[inner] - 1
    n = m.ncells
    r = np.concatenate([a, b, np.arange(n)])
    c = np.concatenate([b, a, np.arange(n)])
    key = np.unique(r.astype(np.int64) * n + c)
    r, c = key // n, key % n
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, r + 1, 1)
    return (np.cumsum(rp) + 1).astype(np.int32), (c + 1).astype(np.int32), a, b


def _bandwidth(a, b, o2n):
    return int(np.abs(o2n[a] - o2n[b]).max())


@pytest.mark.parametrize("mesh", ["hex", "prism"])
@pytest.mark.parametrize("rcm", [False, True])
@pytest.mark.parametrize("colored,coarse", [(False, 0), (True, 0), (True, 16), (False, 16)])
def test_cell_ordering_invariants(mesh, rcm, colored, coarse):
    """The device numbering: a permutation; with colouring no two neighbours
    share a colour and the classes are contiguous and ascending; with a
    coarse space every aggregate is contiguous inside each class (the same
    aggregates as the bisection alone); deterministic."""
    from mixsolver_amd.solver import cell_ordering, coarse_aggregates
    m = (meshgen.channel(16, 8, 6) if mesh == "hex" else
         meshgen.cylinder_prisms(ntheta=40, nr=12, layers=2, jitter=0.15, seed=3))
    rp, ci, a, b = _cell_graph(m)
    xc = _centroids(m)
    n = m.ncells
    n2o, color, ncol = cell_ordering(rp, ci, xc, rcm=rcm, colored=colored, coarse=coarse)
    assert np.array_equal(np.sort(n2o), np.arange(n))
    o2n = np.empty(n, np.int64)
    o2n[n2o] = np.arange(n)
    if colored:
        assert ncol >= 2 and color.max() == ncol - 1
        assert np.all(color[a] != color[b])
        assert np.all(np.diff(color[n2o]) >= 0)
    else:
        assert ncol == 1 and not color.any()
    if coarse:
        agg = coarse_aggregates(xc, coarse)
        cls = color[n2o]
        for c in range(ncol):
            seq = agg[n2o][cls == c]
            # contiguous: each aggregate appears as one run inside the class
            runs = seq[np.r_[True, seq[1:] != seq[:-1]]]
            assert len(runs) == len(np.unique(seq)), c
    again = cell_ordering(rp, ci, xc, rcm=rcm, colored=colored, coarse=coarse)
    assert np.array_equal(again[0], n2o) and np.array_equal(again[1], color)


def _centroids(m):
    """Cell centroids as the mean of the cells' face centroids (enough for a
    geometric split)."""
    x = m.x.reshape(-1, 3)
    fc = np.array([x[m.fnodes[m.fptr[f]:m.fptr[f + 1]] - 1].mean(axis=0) for f in range(m.nfaces)])
    s = np.zeros((m.ncells, 3))
    cnt = np.zeros(m.ncells)
    for side in (m.c0, m.c1):
        ok = side > 0
        np.add.at(s, side[ok] - 1, fc[ok])
        np.add.at(cnt, side[ok] - 1, 1)
    return s / cnt[:, None]


def test_greedy_colouring_of_a_box_is_the_checkerboard():
    """A structured hex box is a bipartite cell graph: greedy colouring in the
    natural order gives exactly two colours, (i + j + k) mod 2."""
    from mixsolver_amd.solver import cell_ordering
    nx, ny, nz = 7, 5, 4
    m = meshgen.hex_box(np.linspace(0, 1, nx + 1), np.linspace(0, 1, ny + 1), np.linspace(0, 1, nz + 1),
                        zone_ids=dict(xmin=5, xmax=5, ymin=5, ymax=5, zmin=5, zmax=5))
    rp, ci, a, b = _cell_graph(m)
    n2o, color, ncol = cell_ordering(rp, ci, colored=True)
    idx = np.arange(m.ncells)
    i, j, k = idx % nx, (idx // nx) % ny, idx // (nx * ny)
    assert ncol == 2 and np.array_equal(color, (i + j + k) % 2)
    # classes in ascending order, natural order inside each class
    assert np.array_equal(n2o, np.r_[idx[color == 0], idx[color == 1]])


def test_rcm_recovers_a_narrow_band_from_a_shuffled_mesh():
    """RCM on a randomly shuffled channel mesh brings the matrix bandwidth
    from about n back to the order of one cross-section (the structured
    numbering's band is nx*ny = 128 cells here); it starts from a
    minimum-degree cell (the last of the reversed order)."""
    from mixsolver_amd.solver import cell_ordering
    m = meshgen.channel(16, 8, 6)
    n = m.ncells
    rp0, ci0, a0, b0 = _cell_graph(m)
    perm = np.random.default_rng(9).permutation(n)        # old -> shuffled id
    a, b = perm[a0], perm[b0]
    r = np.concatenate([a, b, np.arange(n)])
    c = np.concatenate([b, a, np.arange(n)])
    key = np.unique(r.astype(np.int64) * n + c)
    rr, cc = key // n, key % n
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, rr + 1, 1)
    rp, ci = (np.cumsum(rp) + 1).astype(np.int32), (cc + 1).astype(np.int32)
    ident = np.arange(n)
    before = _bandwidth(a, b, ident)
    n2o, _, _ = cell_ordering(rp, ci, rcm=True)
    o2n = np.empty(n, np.int64)
    o2n[n2o] = np.arange(n)
    after = _bandwidth(a, b, o2n)
    assert before > n // 2 and after <= 2 * 16 * 8, (before, after)
    deg = np.diff(rp) - 1
    assert deg[n2o[-1]] == deg.min()


def test_coarse_aggregates_rejects_bad_k():
    from mixsolver_amd._lib import MixSolverError
    from mixsolver_amd.solver import coarse_aggregates
    xc = np.zeros((10, 3))
    for k in (0, 11):
        with pytest.raises(MixSolverError):
            coarse_aggregates(xc, k)


PER_CTL = "bc = 5 wall\nperiodic = 7 6\nperiodic = 9 8\n"


def test_periodic_control_grammar():
    ctrl, groups = parse_control(PER_CTL + "periodic = 11 10 0.5 0 -2.25\n", warn=False)
    assert ctrl.nperiodic == 3
    assert [list(ctrl.periodic_zone[k]) for k in range(3)] == [[7, 6], [9, 8], [11, 10]]
    assert list(ctrl.periodic_auto)[:3] == [1, 1, 0]
    assert list(ctrl.periodic_shift[2]) == [0.5, 0.0, -2.25]
    for bad in ("periodic = 7\n", "periodic = 7 7\n", "periodic = 7 6 1 2\n", "periodic = 1 2\n" * 5,
                "bc = 7 wall\n"):
        with pytest.raises(ControlError):
            parse_control(PER_CTL + bad, warn=False)


def test_periodic_match_closed_form():
    """hex_box emits the x-min and x-max faces (and z-min / z-max) in the same
    (j, k) order, so face i of one zone pairs with face i of the other; the
    jittered end planes are exact translates."""
    from mixsolver_amd.solver import periodic_match
    nx, ny, nz, lx, lz = 6, 4, 5, 2.0, 1.5
    m = meshgen.channel(nx, ny, nz, lx=lx, lz=lz, periodic=True, jitter=0.15, seed=1)
    auto = periodic_match(m, PER_CTL)
    given = periodic_match(m, f"bc = 5 wall\nperiodic = 7 6 {lx!r} 0 0\nperiodic = 9 8 0 0 {lz!r}\n")
    assert np.array_equal(auto, given)
    z = m.zone
    for za, zb, cnt in ((7, 6, ny * nz), (9, 8, nx * ny)):
        fa, fb = np.nonzero(z == za)[0], np.nonzero(z == zb)[0]
        assert len(fa) == len(fb) == cnt
        assert np.array_equal(auto[fa], fb) and np.array_equal(auto[fb], fa)
    assert np.all(auto[np.isin(z, [3, 5])] == -1)


def test_periodic_match_rejects_bad_pairs():
    from mixsolver_amd._lib import MixSolverError
    from mixsolver_amd.solver import periodic_match
    m = meshgen.channel(6, 4, 5, lx=2.0, lz=1.5, periodic=True, jitter=0.15, seed=1)
    for ctl in ("bc = 5 wall\nperiodic = 7 6 1.0 0 0\nperiodic = 9 8\n",     # wrong shift
                "bc = 5 wall\nperiodic = 7 9\n",                               # face counts differ
                "bc = 5 wall\nperiodic = 7 6\nperiodic = 7 8\n"):             # zone in two pairs
        with pytest.raises(MixSolverError):
            periodic_match(m, ctl)
    # end planes jittered independently are not translates of each other
    zid = dict(xmin=6, xmax=7, ymin=5, ymax=5, zmin=8, zmax=9)
    j = meshgen.hex_box(np.linspace(0, 2, 7), np.linspace(-1, 1, 5), np.linspace(0, 1.5, 6), zone_ids=zid,
                        jitter=0.15, seed=1)
    with pytest.raises(MixSolverError):
        periodic_match(j, "bc = 5 wall\nperiodic = 7 6\n")
    # one cell across: a cell would be its own neighbour
    one = meshgen.channel(1, 4, 3, lx=1.0, lz=1.0, periodic=True)
    with pytest.raises(MixSolverError, match="itself"):
        periodic_match(one, PER_CTL)


FLANG = shutil.which("amdflang") or shutil.which("flang")


@pytest.mark.skipif(FLANG is None, reason="no Fortran compiler in this image")
def test_fortran_binding_layout_and_link(tmp_path):
    """mixsolver_c.f90 (the ISO_C_BINDING interface a Fortran caller uses)
    compiles, links against the library, has the C struct sizes, and calls a
    host-only entry point."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    fdir = os.path.join(root, "mixsolver_amd", "fortran")
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run([FLANG, "-c", os.path.join(fdir, "mixsolver_c.f90"), "-o", "mixsolver_c.o"],
                   cwd=tmp_path, check=True)
    subprocess.run([FLANG, os.path.join(fdir, "check_binding.f90"), "mixsolver_c.o", f"-L{libdir}",
                    "-lmixsolver_hip", f"-Wl,-rpath,{libdir}", "-o", "check_binding"], cwd=tmp_path, check=True)
    csz = (tmp_path / "sz.c")
    csz.write_text('#include <stdio.h>\n#include "%s"\nint main(){printf("%%zu %%zu", sizeof(ms_ctrl_t), '
                   'sizeof(ms_bcgroup_t));return 0;}\n' % os.path.join(root, "include", "mixsolver.h"))
    subprocess.run(["gcc", str(csz), "-o", str(tmp_path / "sz")], check=True)
    sizes = subprocess.run([str(tmp_path / "sz")], capture_output=True, text=True, check=True).stdout.split()
    # the ctypes mirror has the same sizes
    from mixsolver_amd.control import MsBcGroup, MsCtrl
    assert [C.sizeof(MsCtrl), C.sizeof(MsBcGroup)] == [int(v) for v in sizes]
    r = subprocess.run([str(tmp_path / "check_binding")] + sizes, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.parametrize("kind", ["hex", "prism"])
def test_serving_workload_systems(kind):
    """The serving workload's systems: 1-based sorted CSR with the diagonal,
    symmetric, zero row sums on interior cells and positive ones on cells with
    a Dirichlet face (weak diagonal dominance), SPD; RHS zero-mean."""
    from mixsolver_amd.serve import LaplacianPattern, request_mesh
    P = LaplacianPattern(request_mesh(700, kind, 4))
    rng = np.random.default_rng(2)
    val, b = P.values(rng), P.rhs(rng)
    A = P.dense(val)
    assert np.array_equal(A, A.T)
    rs = A.sum(axis=1)
    bnd = np.zeros(P.n, bool)
    bnd[P.mesh.c0[P.mesh.c1 == 0] - 1] = True
    assert np.all(np.abs(rs[~bnd]) <= 1e-12 * np.abs(np.diag(A)[~bnd]))
    assert np.all(rs[bnd] > 0)
    assert np.linalg.eigvalsh(A).min() > 0
    assert abs(b.sum()) < 1e-10
    for i in range(P.n):
        cols = P.col_idx[P.row_ptr[i] - 1:P.row_ptr[i + 1] - 1]
        assert np.all(np.diff(cols) > 0) and (i + 1) in cols


def test_packed_bed_closures():
    """Kozeny-Carman written with the particle diameter (constant 180) equals
    the Kozeny form with the specific surface S0 = 6/dp (Kozeny constant 5):
    K = eps^3 / (5 S0^2 (1-eps)^2) = eps^3 / (5 a_sf^2) with a_sf = S0 (1-eps);
    the Ergun sink at a superficial speed U reproduces the Ergun pressure
    gradient 150 mu (1-eps)^2 U / (eps^3 dp^2) + 1.75 rho (1-eps) U^2 / (eps^3 dp)
    once the viscous constant 150 is used (ratio 180/150 on the Darcy part)."""
    from mixsolver_amd import zones
    eps = np.array([0.35, 0.4, 0.6, 0.9])
    dp = 1.0e-3
    K = zones.kozeny_carman_permeability(eps, dp)
    a = zones.specific_surface(eps, dp)
    np.testing.assert_allclose(K, eps ** 3 / (5.0 * a ** 2), rtol=1e-14)
    S0 = 6.0 / dp
    np.testing.assert_allclose(K, eps ** 3 / (5.0 * S0 ** 2 * (1 - eps) ** 2), rtol=1e-14)
    mu, rho, U = 1.8e-5, 1.2, 0.05
    C = zones.ergun_inertial(eps, dp)
    sink = mu / K * U * (150.0 / 180.0) + rho * C * U * U
    ergun = 150 * mu * (1 - eps) ** 2 * U / (eps ** 3 * dp ** 2) + 1.75 * rho * (1 - eps) * U ** 2 / (eps ** 3 * dp)
    np.testing.assert_allclose(sink, ergun, rtol=1e-13)
    # K grows with dp^2 and falls as the bed closes
    assert np.all(np.diff(K) > 0)
    np.testing.assert_allclose(zones.kozeny_carman_permeability(eps, 2 * dp), 4 * K, rtol=1e-14)
    f = zones.packed_bed_fields(eps, dp, k_s=1.0)
    assert f["perm_dir"].shape == (4, 3) and np.all(f["perm_dir"][:, 0] == K)


REF_PRECISION = "/root/reference/src/common/mod_precision.f90"


@pytest.mark.skipif(FLANG is None or not os.path.exists(REF_PRECISION),
                    reason="needs a Fortran compiler and the reference's mod_precision")
def test_dropin_module_host_factorizations(tmp_path):
    """The drop-in mod_uns_linsolver keeps the reference's host helpers:
    its ilu0_factor / ic0_factor reproduce the reference's factors bit for bit
    (solvers_spd fixture), and csr_matvec / ilu0_apply / ic0_apply invert
    them (host only, no device call)."""
    import subprocess
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "solvers_spd.npz"))
    rp, ci, val = d["row_ptr"].astype(np.int32), d["col_idx"].astype(np.int32), d["val"]
    n, nnz = len(rp) - 1, len(ci)
    with open(tmp_path / "in.bin", "wb") as f:
        np.array([n, nnz], np.int32).tofile(f)
        rp.tofile(f)
        ci.tofile(f)
        val.tofile(f)
    prog = tmp_path / "t.f90"
    prog.write_text("""program t
   use mod_precision, only: dp
   use mod_uns_linsolver
   implicit none
   type(csr_t) :: A, L, U
   integer :: n, nnz, ierr, u1, u2
   real(dp), allocatable :: x(:), y(:), z(:)
   open(newunit=u1, file='in.bin', access='stream', form='unformatted', status='old')
   read(u1) n, nnz
   A%nrows = n
   allocate(A%row_ptr(n+1), A%col_idx(nnz), A%val(nnz), x(n), y(n), z(n))
   read(u1) A%row_ptr, A%col_idx, A%val
   close(u1)
   L = A; call ic0_factor(L, ierr)
   open(newunit=u2, file='out.bin', access='stream', form='unformatted', status='replace')
   write(u2) ierr, L%val
   U = A; call ilu0_factor(U, ierr)
   write(u2) ierr, U%val
   x = 1.0_dp
   call csr_matvec(A, x, y)
   call ilu0_apply(U, y, z)
   write(u2) z
   call ic0_apply(L, y, z)
   write(u2) z
   close(u2)
end program t
""")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    fdir = os.path.join(root, "mixsolver_amd", "fortran")
    libdir = os.path.join(root, "mixsolver_amd")
    for src in (REF_PRECISION, os.path.join(fdir, "mixsolver_c.f90"), os.path.join(fdir, "mixsolver_dropin.f90"),
                os.path.join(fdir, "mod_uns_linsolver.f90")):
        subprocess.run([FLANG, "-O2", "-c", src], cwd=tmp_path, check=True)
    subprocess.run([FLANG, "-O2", "t.f90", "mod_precision.o", "mixsolver_c.o", "mixsolver_dropin.o", "mod_uns_linsolver.o",
                    f"-L{libdir}", "-lmixsolver_hip", f"-Wl,-rpath,{libdir}", "-o", "t"], cwd=tmp_path, check=True)
    subprocess.run(["./t"], cwd=tmp_path, check=True)
    b = np.fromfile(tmp_path / "out.bin", dtype=np.uint8)
    o = 0

    def take(dt, k):
        nonlocal o
        a = np.frombuffer(b, dtype=dt, count=k, offset=o)
        o += a.nbytes
        return a
    assert take(np.int32, 1)[0] == d["ic0_ierr"] == 0
    assert np.array_equal(take(np.float64, nnz), d["ic0_L"])
    assert take(np.int32, 1)[0] == d["ilu0_ierr"] == 0
    assert np.array_equal(take(np.float64, nnz), d["ilu0_LU"])
    # for this symmetric matrix (LU)^-1 A 1 is 1 only if the factors are exact; ILU(0) and
    # IC(0) of a matrix with fill differ from A, so check the solves against dense algebra
    z_ilu, z_ic = take(np.float64, n), take(np.float64, n)
    A = np.zeros((n, n))
    for i in range(n):
        A[i, ci[rp[i] - 1:rp[i + 1] - 1] - 1] = val[rp[i] - 1:rp[i + 1] - 1]
    Lu = np.zeros((n, n))
    for i in range(n):
        Lu[i, ci[rp[i] - 1:rp[i + 1] - 1] - 1] = d["ilu0_LU"][rp[i] - 1:rp[i + 1] - 1]
    Lo = np.tril(Lu, -1) + np.eye(n)
    Up = np.triu(Lu)
    np.testing.assert_allclose(Lo @ Up @ z_ilu, A @ np.ones(n), rtol=1e-12, atol=1e-12)
    Lc = np.zeros((n, n))
    for i in range(n):
        Lc[i, ci[rp[i] - 1:rp[i + 1] - 1] - 1] = d["ic0_L"][rp[i] - 1:rp[i + 1] - 1]
    Lc = np.tril(Lc)
    np.testing.assert_allclose(Lc @ Lc.T @ z_ic, A @ np.ones(n), rtol=1e-12, atol=1e-12)


# ------------------------------------------------------------ VTU output
@pytest.mark.parametrize("tag", ["cav", "prism", "solid"])
@pytest.mark.parametrize("binary", [False, True])
def test_vtu_byte_identical_to_reference(tag, binary, tmp_path):
    """ms_vtu_write (host only) writes the reference's vtk_write file byte for
    byte: tetrahedral fan of every face from its first node, reversed for the
    neighbour cell, centroid points, ES24.16E3 ASCII numbers (negative zero,
    1e-300, -2.5e200 among them) or appended Float32 binary with UInt32 block
    headers; the solid bar adds temperature_solid and heat_source_solid."""
    import hashlib
    from conftest import load_golden, mesh_from
    from mixsolver_amd.control import parse_control
    from mixsolver_amd.vtu import optional_arrays, write_vtu
    from mixsolver_amd.zones import resolve_cell_zones, setup_porous_fields
    from oracle import pyoracle
    d = load_golden("vtu_files")
    sub = {k[len(tag) + 1:]: v for k, v in d.items() if k.startswith(tag + "_")}
    m = mesh_from(sub)
    ctrl, groups = parse_control(str(sub["control"]), warn=False)
    xc = pyoracle.OracleCase(m, ctrl, groups).geometry()["xc"]
    cz = resolve_cell_zones(m, ctrl.cell_zones) if m.czone is not None else None
    f = setup_porous_fields(m, ctrl.cell_zones, cz) if cz is not None else None
    with_ts, with_q = optional_arrays(cz, f["h_sf"] if f else [], f["a_sf"] if f else [])
    assert (with_ts, with_q) == ((True, True) if tag == "solid" else (False, False))
    T = sub["T"]
    out = tmp_path / "x.vtu"
    write_vtu(out, m, xc, sub["p"], sub["u"], T, T_s=0.5 * T + 1.0 if with_ts else None,
              q_src=f["q_src"] if with_q else None, binary=binary)
    data = out.read_bytes()
    kind = "b" if binary else "a"
    assert len(data) == int(sub[f"{kind}_size"])
    assert hashlib.sha256(data).hexdigest() == str(sub[f"{kind}_sha256"])


# ------------------------------------------------------------ field dump v1
def test_fortran_number_formats(tmp_path):
    """The dump's ES24.16 numbers and the VTU's ES24.16E3 numbers equal
    flang's own output for the same values (signed zero, denormals, the
    extremes, rounding into the next power of ten, two- and three-digit
    exponents: ES24.16 drops the letter for three-digit exponents)."""
    from conftest import load_golden
    from mixsolver_amd.restart import FIELDS, write_dump
    from mixsolver_amd.vtu import write_vtu
    d = load_golden("fortran_formats")
    v = d["values"]
    n = len(v)
    f = {k: np.zeros((n, 3)) if k.startswith("u") else np.zeros(n) for k in FIELDS}
    f["p"] = v.copy()
    write_dump(tmp_path / "d.txt", n, f, 0, "m.cas", "2026-01-01 00:00:00")
    lines = (tmp_path / "d.txt").read_text().split("\n")
    i = lines.index("# end u") + 1
    assert lines[i:i + n] == list(d["es24_16"])
    assert lines[i + n] == "# end p"
    m = meshgen.hex_box(np.linspace(0.0, 1.0, n + 1), np.array([0.0, 1.0]), np.array([0.0, 1.0]))
    write_vtu(tmp_path / "a.vtu", m, np.zeros((n, 3)), v, np.zeros((n, 3)), np.zeros(n))
    txt = (tmp_path / "a.vtu").read_text()
    blk = txt.split('Name="pressure" NumberOfComponents="1" format="ascii">\n')[1].split("\n        </DataArray>")[0]
    got = blk.split("\n")[::12]        # 6 quad faces -> 12 tetrahedra per cell
    assert got == list(d["es24_16e3"])


def test_dump_round_trip_and_errors(tmp_path):
    """write_field_dump / read_field_dump restated: the seven arrays and
    ts_order survive a round trip bit for bit (17 significant digits), the
    reader takes list-directed input (values split or joined across lines,
    D exponents, commas), and returns the reference's codes: 1 cannot open,
    2 header ended or malformed, 3 version, 4 ncells, 5 data."""
    from mixsolver_amd.restart import FIELDS, read_dump, write_dump
    n = 7
    rng = np.random.default_rng(3)
    f = {k: rng.normal(size=(n, 3)) if k.startswith("u") else rng.normal(size=n) * 1e5 for k in FIELDS}
    f["T"][2] = -0.0
    path = tmp_path / "d.txt"
    write_dump(path, n, f, 2, "case.cas", "2026-01-01 00:00:00")
    ier, g, ts, src = read_dump(path, n)
    assert (ier, ts, src) == (0, 2, "case.cas")
    for k in FIELDS:
        assert np.array_equal(g[k], f[k]), k
    assert np.signbit(g["T"][2])
    text = path.read_text()
    lines = text.split("\n")
    # list-directed: u's 21 values regrouped 7 per line with commas and D exponents
    iu = lines.index("# fields: u p T u_old u_old_old T_old T_old_old") + 1
    vals = [f"{x:.17e}".replace("e", "D") for x in f["u"].ravel()]
    alt = lines[:iu] + [", ".join(vals[i:i + 7]) for i in range(0, 21, 7)] + lines[iu + n:]
    (tmp_path / "alt.txt").write_text("\n".join(alt))
    ier, g2, _, _ = read_dump(tmp_path / "alt.txt", n)
    assert ier == 0 and np.array_equal(g2["u"], f["u"]) and np.array_equal(g2["p"], f["p"])
    assert read_dump(tmp_path / "missing.txt", n)[0] == 1
    (tmp_path / "h.txt").write_text("\n".join(lines[:3]) + "\n")
    assert read_dump(tmp_path / "h.txt", n)[0] == 2
    (tmp_path / "m.txt").write_text("\n".join(lines[:2] + ["1.0"] + lines[2:]))
    assert read_dump(tmp_path / "m.txt", n)[0] == 2
    (tmp_path / "v.txt").write_text(text.replace("field dump v1", "field dump v2"))
    assert read_dump(tmp_path / "v.txt", n)[0] == 3
    assert read_dump(path, n + 1)[0] == 4
    (tmp_path / "t.txt").write_text("\n".join(lines[: iu + n + 3]))
    assert read_dump(tmp_path / "t.txt", n)[0] == 5


def _chan_cells(ys, nx=300):
    xs = np.arange(nx + 1.0)
    yc = 0.5 * (ys[1:] + ys[:-1])
    X, Y = np.meshgrid(0.5 * (xs[1:] + xs[:-1]), yc, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), np.full(X.size, 5.0)], 1)


def test_channel_check_closed_forms():
    """The channel indicators on fields with known answers, on wall-clustered
    rows (spacing grows away from both walls, so the end rows' widths stop
    at the walls): uniform u gives U_bulk = u exactly in every layer; the
    Poiseuille profile sampled at the centres has zero profile error and
    u_max / U = max 6 eta (1 - eta); a pressure linear in x gives its slope
    and the rise -slope * 299; a centreline ramp crosses the 1 % band at the
    layer it was built to; a dropped outer row or unclipped end width
    changes the uniform-u bulk."""
    from casecheck import channel_check
    s = 0.5 * (1.0 - np.cos(np.pi * np.linspace(0.0, 1.0, 41)))
    ys = 10.0 * s
    xc = _chan_cells(ys)
    n = len(xc)
    ub = 100.0 * 1.846e-5 / (1.177 * 10.0)
    u = np.zeros((n, 3))
    u[:, 0] = 0.8 * ub
    c = channel_check(xc, u, np.zeros(n))
    assert c["u_bulk_in"] == pytest.approx(0.8 * ub, rel=1e-14)
    assert c["u_bulk_out"] == pytest.approx(0.8 * ub, rel=1e-14)
    assert c["mass_drift"] == pytest.approx(0.0, abs=1e-13)
    assert c["n_layers"] == 300
    eta = xc[:, 1] / 10.0
    u[:, 0] = 6.0 * ub * eta * (1.0 - eta)
    p = 3.0e-9 - 2.5e-10 * xc[:, 0]
    c = channel_check(xc, u, p)
    assert c["profile_err_max"] < 1e-14 and c["profile_err_l1"] < 1e-14
    assert c["umax_ratio_out"] == pytest.approx((6 * eta * (1 - eta)).max(), rel=1e-14)
    assert c["dpdx_fit"] == pytest.approx(-2.5e-10, rel=1e-10)
    assert c["p_rise"] == pytest.approx(2.5e-10 * 299.0, rel=1e-12)
    assert c["dpdx_exact"] == pytest.approx(-12 * 1.846e-5 * ub / 100.0, rel=1e-15)
    assert 0.99 < c["u_bulk_in"] / ub < 1.01
    # centreline ramp: reaches the 1 % band at x = 40.5
    mid = np.abs(xc[:, 1] - 5.0) == np.abs(xc[:, 1] - 5.0).min()
    u[:, 0] = 0.0
    u[mid, 0] = 1.5 * ub * np.minimum(xc[mid, 0] / 40.5, 1.0) * 0.995 + 0.0075 * ub * (xc[mid, 0] >= 40.5)
    assert channel_check(xc, u, p)["entrance_length"] == 40.5


@pytest.mark.skipif(FLANG is None, reason="needs a Fortran compiler")
def test_dropin_mpi_module_host_helpers(tmp_path):
    """The one-partition drop-in mod_uns_linsolver_mpi: owned_block keeps the
    owned rows' columns <= nowned in stored order and matvec_owned is A x over
    the owned rows (owned_block / matvec_owned, mod_uns_linsolver_mpi.f90:
    56-149, host only), and a solver handed a partitioned matrix (halo rows)
    stops with a message before any device call."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    fdir = os.path.join(root, "mixsolver_amd", "fortran")
    libdir = os.path.join(root, "mixsolver_amd")
    prog = tmp_path / "t.f90"
    prog.write_text("""program t
   use mod_precision, only: dp
   use mod_uns_linsolver, only: csr_t
   use mod_uns_linsolver_mpi
   implicit none
   type(csr_t) :: A, B
   real(dp) :: x(4), y(3), res
   integer :: it, ierr, halo, k
   character(len=8) :: arg
   ! 3 owned rows, one halo column (4) in rows 1 and 3
   A%nrows = 4
   A%row_ptr = [1, 4, 6, 9, 9]
   A%col_idx = [1, 2, 4, 1, 2, 2, 3, 4]
   A%val = [4.0_dp, -1.0_dp, -0.5_dp, -1.0_dp, 3.0_dp, -1.0_dp, 5.0_dp, -2.0_dp]
   call owned_block(A, 3, B)
   print '(4i3)', B%row_ptr
   print '(6i3)', B%col_idx
   print '(6f6.2)', B%val
   x = [1.0_dp, 2.0_dp, 3.0_dp, 4.0_dp]
   halo = 0
   call matvec_owned(A, x, y, halo, 3)
   print '(3f8.3)', y
   call get_command_argument(1, arg)
   if (trim(arg) == 'solve') then
      x = 0.0_dp
      call cg_jacobi_mpi(A, y, x, 3, halo, 1.0e-10_dp, 10, it, res, ierr)
   end if
end program t
""")
    srcs = [os.path.join(root, "tests", "dropin", "mod_precision.f90"), os.path.join(fdir, "mixsolver_c.f90"),
            os.path.join(fdir, "mixsolver_dropin.f90"), os.path.join(fdir, "mod_uns_linsolver.f90"),
            os.path.join(fdir, "mod_uns_linsolver_mpi.f90")]
    for src in srcs:
        subprocess.run([FLANG, "-O2", "-c", src], cwd=tmp_path, check=True)
    objs = ["mod_precision.o", "mixsolver_c.o", "mixsolver_dropin.o", "mod_uns_linsolver.o", "mod_uns_linsolver_mpi.o"]
    subprocess.run([FLANG, "-O2", "t.f90"] + objs + [f"-L{libdir}", "-lmixsolver_hip", f"-Wl,-rpath,{libdir}",
                    "-o", "t"], cwd=tmp_path, check=True)
    out = subprocess.run(["./t"], cwd=tmp_path, check=True, capture_output=True, text=True).stdout.split("\n")
    assert out[0].split() == ["1", "3", "5", "7"]
    assert out[1].split() == ["1", "2", "1", "2", "2", "3"]
    assert [float(v) for v in out[2].split()] == [4.0, -1.0, -1.0, 3.0, -1.0, 5.0]
    assert [float(v) for v in out[3].split()] == [4.0 - 2.0 - 2.0, -1.0 + 6.0, -2.0 + 15.0 - 8.0]
    r = subprocess.run(["./t", "solve"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode != 0 and "one partition only" in r.stdout
