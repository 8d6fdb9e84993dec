// ms_schur.hip -- red-black reduction of the pressure-correction system.
//
// When the cell graph has two colours, the colour-blocked numbering puts
// colour 0 in rows [0, n0) and colour 1 in rows [n0, n), and the p' matrix
// is A = [D0 B; B^T D1] with D0 and D1 diagonal (no two cells of one colour
// are neighbours).  Eliminating colour 0 leaves the Schur complement on
// colour 1, S = D1 - B^T D0^-1 B, and A x = b becomes
//     S x1 = b1 - B^T D0^-1 b0,   x0 = D0^-1 (b0 - B x1).
// PCG on S with the Jacobi preconditioner diag(S) is the colour-blocked
// DILU-preconditioned PCG on A seen on colour 1: diag(S) is DILU's colour-1
// diagonal, and the two produce the same iterates up to rounding.  An
// iteration on S costs one product with S (19 entries per row on
// hexahedra, on half the cells) instead of a product with A and the two
// triangular sweeps of DILU.  After the back substitution the colour-0 rows
// of b - A x vanish and the colour-1 rows are the residual of the reduced
// system, so the FP64 tolerance on ||b - A x|| / ||b|| carries over to the
// reduced solve.  S is formed once per matrix (FP64, a fixed summation
// order per entry); the reduced solve is the PCG of ms_linalg.hip (FP64, or
// FP32 inner iterations with FP64 reliable updates) with the aggregate
// coarse space of the colour-1 cells.
#include <algorithm>
#include <cmath>

#include "ms_kernels.hpp"
#include "ms_system.hpp"

namespace ms {
void build_coarse(SellPattern &P, const std::vector<int> &agg_dev, int k);
void csr_to_sell(SellPattern &P, int n, const std::vector<int> &rp, const std::vector<int> &ci);

namespace {

dim3 gc(int n) { return dim3(nblocks(n)); }

__device__ __forceinline__ long long sl(const Mat &M, int r, int k) {
  return M.off[r >> 6] + (long long)k * kSlice + (r & 63);
}

// S = D1 - B^T D0^-1 B on colour-1 row i (device row n0 + i): the diagonal
// value of A, then for every colour-0 neighbour k of the row (in the row's
// column order) and every colour-1 neighbour j of k (in k's column order)
// S(i, j) -= (a_rk / a_kk) a_kj.
__global__ void k_schur_values(Mat A, const double *aval, int n0, int n1, Mat S, double *sval) {
  const int i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;   // XCD-local rows
  if (i >= n1) return;
  const int r = n0 + i;
  const int ls = S.rowlen[i];
  for (int m = 0; m < ls; m++) sval[sl(S, i, m)] = 0.0;
  sval[sl(S, i, S.diagk[i])] = aval[sl(A, r, A.diagk[r])];
  const int la = A.rowlen[r];
  for (int kk = 0; kk < la; kk++) {
    const long long q = sl(A, r, kk);
    const int k = A.col[q];
    if (k == r) continue;
    const double w = aval[q] / aval[sl(A, k, A.diagk[k])];
    const int lk = A.rowlen[k];
    for (int jj = 0; jj < lk; jj++) {
      const long long q2 = sl(A, k, jj);
      const int j = A.col[q2];
      if (j == k) continue;
      const int jl = j - n0;
      for (int m = 0; m < ls; m++) {
        const long long qs = sl(S, i, m);
        if (S.col[qs] == jl) {
          sval[qs] = sval[qs] - w * aval[q2];
          break;
        }
      }
    }
  }
}

// r = b - A x (FP64) and the partial sums of r.r per workgroup
__global__ void k_schur_resid(int n, Mat A, const double *aval, const double *b, const double *x, double *r,
                              double *part) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double v[1] = {0.0};
  if (i < n) {
    double s = 0.0;
    const int len = A.rowlen[i];
    for (int k = 0; k < len; k++) {
      const long long q = sl(A, i, k);
      s = s + aval[q] * x[A.col[q]];
    }
    const double ri = b[i] - s;
    r[i] = ri;
    v[0] = ri * ri;
  }
  block_partials<OpSum, 1>(v, part);
}

// partial sums of v.v
__global__ void k_schur_sumsq(int n, const double *v, double *part) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double a[1] = {i < n ? v[i] * v[i] : 0.0};
  block_partials<OpSum, 1>(a, part);
}

// reduced right-hand side rs_i = r_(n0+i) - sum_k (a_rk / a_kk) r_k, its
// partial sums of squares, and d = 0
__global__ void k_schur_rhs(Mat A, const double *aval, int n0, int n1, const double *r, double *rs, double *d,
                            double *part) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double v[1] = {0.0};
  if (i < n1) {
    const int row = n0 + i;
    double s = r[row];
    const int la = A.rowlen[row];
    for (int kk = 0; kk < la; kk++) {
      const long long q = sl(A, row, kk);
      const int k = A.col[q];
      if (k == row) continue;
      s = s - aval[q] / aval[sl(A, k, A.diagk[k])] * r[k];
    }
    rs[i] = s;
    d[i] = 0.0;
    v[0] = s * s;
  }
  block_partials<OpSum, 1>(v, part);
}

// x1 += d; x0 += (r0 - B d) / D0
__global__ void k_schur_back(Mat A, const double *aval, int n0, int n, const double *r, const double *d, double *x) {
  const int row = blockIdx.x * kBlock + threadIdx.x;
  if (row >= n) return;
  if (row >= n0) {
    x[row] = x[row] + d[row - n0];
    return;
  }
  double s = r[row];
  const int la = A.rowlen[row];
  for (int kk = 0; kk < la; kk++) {
    const long long q = sl(A, row, kk);
    const int j = A.col[q];
    if (j == row) continue;
    s = s - aval[q] * d[j - n0];
  }
  x[row] = x[row] + s / aval[sl(A, row, A.diagk[row])];
}

__global__ void k_schur_zero(int n, double *x) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] = 0.0;
}

// weights of the matrix-free S (MatFree) from the FP64 matrix A: colour-0
// rows -a_kj / a_kk off the diagonal and (row sum) / a_kk on it, colour-1
// rows -a_ij and the row sum; padding 0.  Quotients and sums in FP64.
__global__ void k_mf_values(Mat A, const double *val, int n0, int n, float *w) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= n) return;
  const int s = r >> 6;
  const long long o0 = A.off[s];
  const int wd = (int)((A.off[s + 1] - o0) >> 6);
  const long long base = o0 + (r & 63);
  const int dk = A.diagk[r];
  double d = 0.0, sum = 0.0;
  for (int k = 0; k < wd; k++) {
    const double a = val[base + (long long)k * kSlice];
    sum = sum + a;
    if (k == dk) d = a;
  }
  const bool c0 = r < n0;
  for (int k = 0; k < wd; k++) {
    const long long q = base + (long long)k * kSlice;
    float v;
    if (k == dk) v = c0 ? (d != 0.0 ? (float)(sum / d) : 0.f) : (float)sum;
    else if (A.col[q] == r) v = 0.f;   // padding
    else v = c0 ? (d != 0.0 ? (float)(-val[q] / d) : 0.f) : (float)(-val[q]);
    w[q] = v;
  }
}

// sum of the workgroup partials on the host, in index order
double host_sum(SchurSystem &S, int nparts) {
  S.hpart.resize(nparts);
  MS_HIP(hipMemcpyAsync(S.hpart.data(), S.part.p, nparts * sizeof(double), hipMemcpyDeviceToHost, stream()));
  MS_HIP(hipStreamSynchronize(stream()));
  double s = 0.0;
  for (int i = 0; i < nparts; i++) s += S.hpart[i];
  return s;
}

}  // namespace

// The reduced system of a colour-blocked system with two colours: the
// pattern of S (colour-1 rows, the distance-2 colour-1 neighbours of each),
// its aggregate coarse space (the colour-1 cells' aggregates, contiguous as
// in the full numbering) and its solver.
void build_schur(System &s, const std::vector<int> &agg_dev, int nagg) {
  const SellPattern &P = s.pat;
  if (P.color_ptr.size() != 3)
    fail(MS_EINVAL, "gpu_ppe_solver = schur needs a two-colour cell graph (this mesh has " +
                        std::to_string(P.color_ptr.size() - 1) + " colours)");
  auto S = std::make_unique<SchurSystem>();
  const int n = P.n, n0 = P.color_ptr[1], n1 = n - n0;
  S->n0 = n0;
  S->n1 = n1;
  std::vector<int> rp(n1 + 1, 0), ci, row;
  for (int i = 0; i < n1; i++) {
    const int r = n0 + i;
    row.clear();
    row.push_back(i);
    for (int kk = 0; kk < P.rowlen[r]; kk++) {
      const int k = P.col[P.slot(r, kk)];
      if (k == r) continue;
      for (int jj = 0; jj < P.rowlen[k]; jj++) {
        const int j = P.col[P.slot(k, jj)];
        if (j != k) row.push_back(j - n0);
      }
    }
    std::sort(row.begin(), row.end());
    row.erase(std::unique(row.begin(), row.end()), row.end());
    ci.insert(ci.end(), row.begin(), row.end());
    rp[i + 1] = (int)ci.size();
  }
  csr_to_sell(S->pat, n1, rp, ci);
  S->pat.upload();
  if (nagg > 0) {
    // aggregates with colour-1 cells, renumbered in order (an aggregate
    // without any would leave a zero row in Z^T S Z)
    std::vector<int> agg1(agg_dev.begin() + n0, agg_dev.end()), newid(nagg, -1);
    int k1 = 0;
    for (int a : agg1)
      if (newid[a] < 0) newid[a] = 0;
    for (int a = 0; a < nagg; a++)
      if (newid[a] == 0) newid[a] = k1++;
    for (int &a : agg1) a = newid[a];
    build_coarse(S->pat, agg1, k1);
  }
  S->kry = std::make_unique<Krylov>(&S->pat);
  S->val.alloc(S->pat.nstore);
  S->rhs.alloc(n1);
  S->d.alloc(n1);
  S->r64.alloc(n);
  S->part.alloc(nblocks(n) + 64);
  attach_matfree(*S, s.P);
  s.schur = std::move(S);
}

void attach_matfree(SchurSystem &S, const SellPattern *A) {
  S.mf_w.alloc(A->nstore);
  S.mf_hx.alloc(std::max(S.n0, 1));
  S.mf = MatFree{S.n0, S.n1, A, S.mf_w.p, S.mf_hx.p, nullptr};
  const Coarse &C = S.P->coarse;
  if (C.k > 0) {   // aggregate of every colour-1 row, from the aggregate chunks
    std::vector<int> agg(S.n1, 0);
    for (int c = 0; c < C.nchunks; c++)
      for (int i = C.chunk_ptr[c]; i < C.chunk_ptr[c + 1]; i++) agg[i] = C.chunk_agg[c];
    S.mf_agg.upload(agg);
    S.mf.agg1 = S.mf_agg.p;
  }
  S.kry->set_matfree(&S.mf);
}

SolveResult schur_solve(System &s, const double *val, const double *b, double *x, double tol, int maxit, bool mixed,
                        double inner_tol, bool reuse) {
  SchurSystem &S = *s.schur;
  const int n = s.P->n, n0 = S.n0, n1 = S.n1;
  const Mat A = matof(s.P), Sm = matof(S.P);
  hipStream_t st = stream();
  SolveResult R;
  hipLaunchKernelGGL(k_schur_sumsq, gc(n), dim3(kBlock), 0, st, n, b, S.part.p);
  MS_CHECK_LAUNCH();
  const double bnrm = std::sqrt(host_sum(S, nblocks(n)));
  if (bnrm == 0.0) {
    hipLaunchKernelGGL(k_schur_zero, gc(n), dim3(kBlock), 0, st, n, x);
    MS_CHECK_LAUNCH();
    return R;
  }
  if (!(reuse && S.have_values)) {
    hipLaunchKernelGGL(k_schur_values, gc(n1), dim3(kBlock), 0, st, A, val, n0, n1, Sm, S.val.p);
    if (mixed && S.kry->matfree_on()) hipLaunchKernelGGL(k_mf_values, gc(n), dim3(kBlock), 0, st, A, val, n0, n, S.mf_w.p);
    MS_CHECK_LAUNCH();
    S.have_values = true;
    reuse = false;
  }
  const bool coarse = S.P->coarse.k > 0;
  int total = 0;
  // a pass: r = b - A x, the reduced solve to ||b|| tol, the back
  // substitution; a second pass only when rounding left the full residual
  // above tol
  for (int pass = 0;; pass++) {
    hipLaunchKernelGGL(k_schur_resid, gc(n), dim3(kBlock), 0, st, n, A, val, b, x, S.r64.p, S.part.p);
    MS_CHECK_LAUNCH();
    R.res = std::sqrt(host_sum(S, nblocks(n))) / bnrm;
    if (R.res <= tol || pass == 3) break;
    if (total >= maxit) {
      R.ierr = 1;
      break;
    }
    hipLaunchKernelGGL(k_schur_rhs, gc(n1), dim3(kBlock), 0, st, A, val, n0, n1, S.r64.p, S.rhs.p, S.d.p, S.part.p);
    MS_CHECK_LAUNCH();
    const double rsn = std::sqrt(host_sum(S, nblocks(n1)));
    double tol_s = rsn > 0.0 ? tol * bnrm / rsn : tol;
    // A correction pass follows a pass whose reduced residual met its target
    // while the full residual, evaluated with A rather than with the formed
    // S, stayed just above tol by rounding; the correction then asks for a
    // tenfold reduction at least, so that it does not end at the same margin.
    if (pass > 0) tol_s = std::min(tol_s, 0.1);
    SolveResult I = mixed ? S.kry->mixed(S.val.p, S.rhs.p, S.d.p, tol_s, maxit - total, PC_JACOBI, 0, inner_tol,
                                         coarse, reuse || pass > 0)
                          : S.kry->pcg(S.val.p, S.rhs.p, S.d.p, tol_s, maxit - total, PC_JACOBI, coarse,
                                       reuse || pass > 0);
    if (!mixed) I.it64 = I.it;
    total += I.it;
    R.it32 += I.it32;
    R.it64 += I.it64;
    R.stalled |= I.stalled;
    R.outer += I.outer;
    if (I.ierr == 2) {
      R.ierr = 2;
      break;
    }
    hipLaunchKernelGGL(k_schur_back, gc(n), dim3(kBlock), 0, st, A, val, n0, n, S.r64.p, S.d.p, x);
    MS_CHECK_LAUNCH();
    if (I.ierr == 1) {
      R.ierr = 1;
    }
  }
  R.it = total;
  if (R.ierr == 1 && R.res <= tol) R.ierr = 0;
  return R;
}

}  // namespace ms
