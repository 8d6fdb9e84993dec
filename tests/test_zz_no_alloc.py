"""ms_cg_jacobi_noalloc (the entry point that allocates nothing) against
ms_cg_jacobi on the same systems, two calls in a row with the second system
smaller than the first, plus its size limits and the reference's special
cases (zero right-hand side, a row without a usable diagonal)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_ROWS = 131072
MAX_NNZ = 1048576


def laplacian(nx, seed):
    """1-based CSR of a 7-point finite-volume Laplacian on an nx^3 grid with
    log-normal face coefficients and Dirichlet walls (SPD, diagonally
    dominant), and a seeded right-hand side."""
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
        for face in (0, nx - 1):   # Dirichlet walls
            w = np.take(idx, [face], axis=ax).ravel()
            np.add.at(diag, w, np.exp(rng.standard_normal(w.size)))
    rows.append(np.arange(n))
    cols.append(np.arange(n))
    vals.append(diag)
    r, c, v = np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)
    order = np.lexsort((c, r))
    r, c, v = r[order], c[order], v[order]
    row_ptr = np.zeros(n + 1, np.int32)
    np.add.at(row_ptr, r + 1, 1)
    row_ptr = np.cumsum(row_ptr).astype(np.int32) + 1
    b = rng.standard_normal(n)
    return row_ptr, (c + 1).astype(np.int32), v, b


def spmv(row_ptr, col_idx, val, x):
    n = len(row_ptr) - 1
    r = np.repeat(np.arange(n), np.diff(row_ptr))
    y = np.zeros(n)
    np.add.at(y, r, val * x[col_idx - 1])
    return y


def call(name, row_ptr, col_idx, val, b, x0, tol, maxit):
    from mixsolver_amd import _lib
    L = _lib.lib()
    x = np.array(x0, np.float64, copy=True)
    it, res, ierr = C.c_int(0), C.c_double(0.0), C.c_int(0)
    rc = getattr(L, name)(len(b), _lib.ip(row_ptr), _lib.ip(col_idx), _lib.dp(val), _lib.dp(b), _lib.dp(x),
                          float(tol), int(maxit), C.byref(it), C.byref(res), C.byref(ierr))
    msg = L.ms_last_error().decode() if rc else ""
    return rc, msg, x, it.value, res.value, ierr.value


@pytest.fixture(scope="module")
def lib():
    from mixsolver_amd import _lib
    L = _lib.lib()
    _lib.check(L.ms_init(0))
    return L


def test_noalloc_matches_cg_jacobi_two_calls(lib):
    tol = 1e-10
    for nx, seed in ((48, 11), (20, 12)):      # 110592 rows, then 8000
        rp, ci, v, b = laplacian(nx, seed)
        assert len(b) <= MAX_ROWS and rp[-1] - 1 <= MAX_NNZ
        x0 = np.zeros(len(b))
        print(f"noalloc vs cg_jacobi: {nx}^3 rows, nnz {rp[-1] - 1}", flush=True)
        rc, msg, xr, itr, resr, ierr_r = call("ms_cg_jacobi", rp, ci, v, b, x0, tol, 5000)
        assert rc == 0, msg
        rc, msg, xn, itn, resn, ierr_n = call("ms_cg_jacobi_noalloc", rp, ci, v, b, x0, tol, 5000)
        assert rc == 0, msg
        assert ierr_n == ierr_r == 0
        assert abs(itn - itr) <= 1, (itn, itr)
        assert resn <= tol and resr <= tol
        assert np.linalg.norm(xn - xr) <= 1e-10 * np.linalg.norm(xr)
        true_res = np.linalg.norm(b - spmv(rp, ci, v, xn)) / np.linalg.norm(b)
        assert true_res <= 2 * tol, true_res


def test_noalloc_exhaustion_and_special_cases(lib):
    rp, ci, v, b = laplacian(12, 13)
    n = len(b)
    x0 = np.random.default_rng(3).standard_normal(n)
    # exhaustion: it = maxit + 1, ierr = 1, the same iterate as cg_jacobi
    rc, msg, xr, itr, resr, er = call("ms_cg_jacobi", rp, ci, v, b, x0, 1e-14, 7)
    assert rc == 0, msg
    rc, msg, xn, itn, resn, en = call("ms_cg_jacobi_noalloc", rp, ci, v, b, x0, 1e-14, 7)
    assert rc == 0, msg
    assert (itn, en) == (itr, er) == (8, 1)
    assert np.linalg.norm(xn - xr) <= 1e-12 * np.linalg.norm(xr)
    assert abs(resn - resr) <= 1e-10 * resr
    # zero right-hand side: x = 0, it = 0
    rc, msg, xn, itn, resn, en = call("ms_cg_jacobi_noalloc", rp, ci, v, np.zeros(n), x0, 1e-10, 100)
    assert rc == 0 and (itn, en, resn) == (0, 0, 0.0) and not xn.any()
    # a zero diagonal value: cg_jacobi divides by it and runs out of
    # iterations (mod_uns_linsolver.f90:252-261 checks presence only)
    v2 = v.copy()
    row = 5
    k = [k for k in range(rp[row] - 1, rp[row + 1] - 1) if ci[k] == row + 1][0]
    v2[k] = 0.0
    rc, msg, xr, itr, resr, er = call("ms_cg_jacobi", rp, ci, v2, b, x0, 1e-10, 100)
    rc2, msg2, xn, itn, resn, en = call("ms_cg_jacobi_noalloc", rp, ci, v2, b, x0, 1e-10, 100)
    assert rc == rc2 == 0
    assert (itn, en) == (itr, er) == (101, 1)
    # a row without its diagonal entry is an argument error for both
    keep = np.ones(len(ci), bool)
    keep[k] = False
    rp3 = rp.copy()
    rp3[row + 1:] -= 1
    rc, msg, *_ = call("ms_cg_jacobi", rp3, ci[keep], v[keep], b, x0, 1e-10, 100)
    rc2, msg2, *_ = call("ms_cg_jacobi_noalloc", rp3, ci[keep], v[keep], b, x0, 1e-10, 100)
    assert rc == rc2 == -1 and "diagonal" in msg and "diagonal" in msg2


def test_noalloc_rejects_sizes_beyond_scratch(lib):
    n = MAX_ROWS + 1
    rp = np.arange(1, n + 2, dtype=np.int32)
    ci = np.arange(1, n + 1, dtype=np.int32)
    rc, msg, *_ = call("ms_cg_jacobi_noalloc", rp, ci, np.ones(n), np.ones(n), np.zeros(n), 1e-10, 10)
    assert rc == -1 and "131072 rows" in msg
    n = 4
    per_row = MAX_NNZ // n + 1        # nnz = MAX_NNZ + 4
    rp = (np.arange(n + 1) * per_row + 1).astype(np.int32)
    ci = np.tile(np.arange(1, n + 1, dtype=np.int32), per_row)
    rc, msg, *_ = call("ms_cg_jacobi_noalloc", rp, ci, np.ones(rp[-1] - 1), np.ones(n), np.zeros(n), 1e-10, 10)
    assert rc == -1 and "1048576 nonzeros" in msg
    # malformed CSR is rejected on the host before anything reaches the device
    rp, ci, v, b = laplacian(4, 1)
    bad = ci.copy()
    bad[3] = len(b) + 1
    rc, msg, *_ = call("ms_cg_jacobi_noalloc", rp, bad, v, b, np.zeros(len(b)), 1e-10, 10)
    assert rc == -1 and "out of range" in msg
