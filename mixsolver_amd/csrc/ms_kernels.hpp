// ms_kernels.hpp -- device code shared by the single-system Krylov solvers
// (ms_linalg.hip) and the multi-system request server (ms_serve.hip): SELL-64
// row kernels, the finisher functors of the solver recurrences, the FP32
// matrix copy, preconditioner set-up and the coarse-space kernels.
#pragma once
#include "ms_linalg.hpp"
#include "ms_reduce.hpp"

namespace ms {
namespace {

constexpr double kTiny = 2.2250738585072014e-308;   // tiny(1.0_dp)
constexpr double kRefineTol = 1.0e-5;               // FP32 inner BiCGSTAB reduction per refinement cycle

struct Mat {
  const long long *off;
  const int *rowlen, *col, *diagk;
  const unsigned short *col16;   // 16-bit offsets for the FP32 kernels (nullptr: use col)
  const int *cbase;
};

// column of stored slot q of row r (slice s) in the FP32 kernels
__device__ __forceinline__ int col_f32(const Mat &A, long long q, int r, int s) {
  if (A.col16) {
    const unsigned c = A.col16[q];
    return c < 0xFFFEu ? A.cbase[s] + (int)c : (c == 0xFFFFu ? r : A.col[q]);
  }
  return A.col[q];
}

// column of slot q with the column format fixed at compile time: C16 = 1
// 16-bit offsets, 0 32-bit columns, -1 decided per call (col_f32).  The
// fixed forms keep the slice loop free of the format branch (measured 7%
// on the FP32 SpMV).
template <int C16>
__device__ __forceinline__ int col_at(const Mat &A, long long q, int r, int s) {
  if constexpr (C16 == 1) {
    const unsigned c = A.col16[q];
    return c < 0xFFFEu ? A.cbase[s] + (int)c : (c == 0xFFFFu ? r : A.col[q]);
  } else if constexpr (C16 == 0) {
    return A.col[q];
  } else {
    return col_f32(A, q, r, s);
  }
}

// Row set of a cell-wise kernel: plain blocks of kBlock rows, or (coarse
// space active) one chunk of rows of a single aggregate per workgroup.
struct Rows {
  const int *cptr;   // chunk row boundaries, nullptr = plain blocks
  const int *cagg;   // aggregate of each chunk
  int n;
  __device__ __forceinline__ bool get(int &i, int &a) const {
    if (cptr) {
      const int r0 = cptr[blockIdx.x], r1 = cptr[blockIdx.x + 1];
      i = r0 + (int)threadIdx.x;
      a = cagg[blockIdx.x];
      return i < r1;
    }
    i = blockIdx.x * kBlock + threadIdx.x;
    a = 0;
    return i < n;
  }
};

// Coarse solve inside a finisher: c = Z^T r from the chunk partials (last
// of the NV values), nu = E^-1 c, st->cnu = c.nu.
struct CoDev {
  int k;             // 0 = no coarse space
  const int *agg_cptr, *agg_cidx;
  const float *Einv;
  double *nu;
  KS *st;
};

template <class T>
__device__ __forceinline__ T row_dot(int r, const Mat &A, const T *val, const T *x) {
  const long long base = A.off[r >> 6] + (r & 63);
  const int len = A.rowlen[r];
  T s = T(0);
  for (int k = 0; k < len; k++) {
    const long long q = base + (long long)k * kSlice;
    s = s + val[q] * x[A.col[q]];
  }
  return s;
}

// FP32 rows are stored with the FP64 row sum in the diagonal slot and
// applied in difference form, q_i = s_i p_i + sum_{j!=i} a_ij (p_j - p_i):
// algebraically equal to A p, but constants are annihilated exactly, so the
// near-null space of the pressure-correction Laplacian survives the rounding
// of its coefficients to FP32.
// Slice of a fixed width W: every column is loaded first, then every value,
// so the gathers of x start as soon as the first column arrives while the
// value loads are in flight; accumulation runs k = 0..W-1 as in the loop.
template <class T, int C16, int W>
__device__ __forceinline__ T row_apply_w(int r, int s, long long base, const Mat &A, const T *val, const T *x,
                                         T xi) {
  int c[W];
  T a[W];
#pragma unroll
  for (int k = 0; k < W; k++) c[k] = col_at<C16>(A, base + (long long)k * kSlice, r, s);
#pragma unroll
  for (int k = 0; k < W; k++) a[k] = val[base + (long long)k * kSlice];
  T acc = T(0);
#pragma unroll
  for (int k = 0; k < W; k++) acc = acc + a[k] * (c[k] == r ? xi : x[c[k]] - xi);
  return acc;
}

template <class T, int C16 = -1>
__device__ __forceinline__ T row_apply(int r, const Mat &A, const T *val, const T *x) {
  if constexpr (sizeof(T) == 8) {
    return row_dot(r, A, val, x);
  } else {
    // every slot of the slice width contributes: a_k (x_col - x_i) off the
    // diagonal, s_i x_i in the diagonal slot (col = r, holding the row sum)
    // and 0 in padding (col = r, a = 0); the loop runs the slice width,
    // uniform across the wavefront, and needs no diagonal position.  The
    // widths of tetrahedral, prismatic and hexahedral rows (5, 6, 7) take
    // fully unrolled bodies.
    const int s = r >> 6;
    const long long o0 = A.off[s];
    const int w = (int)((A.off[s + 1] - o0) >> 6);
    const long long base = o0 + (r & 63);
    const T xi = x[r];
    if (w == 7) return row_apply_w<T, C16, 7>(r, s, base, A, val, x, xi);
    if (w == 6) return row_apply_w<T, C16, 6>(r, s, base, A, val, x, xi);
    if (w == 5) return row_apply_w<T, C16, 5>(r, s, base, A, val, x, xi);
    T acc = T(0);
#pragma unroll 4
    for (int k = 0; k < w; k++) {
      const long long q = base + (long long)k * kSlice;
      const int c = col_at<C16>(A, q, r, s);
      acc = acc + val[q] * (c == r ? xi : x[c] - xi);
    }
    return acc;
  }
}

// FP32 copy of an FP64 matrix for row_apply: off-diagonals rounded, diagonal
// slot = FP64 row sum rounded.
__global__ void k_mat32_diff(int n, Mat A, const double *v64, float *v32) {
  int r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= n) return;
  const long long base = A.off[r >> 6] + (r & 63);
  const int len = A.rowlen[r], dk = A.diagk[r];
  double s = 0.0;
  for (int k = 0; k < len; k++) {
    const long long q = base + (long long)k * kSlice;
    s = s + v64[q];
    if (k != dk) v32[q] = (float)v64[q];
  }
  v32[base + (long long)dk * kSlice] = (float)s;
  // padding slots are read by row_apply: zero them
  const int w = (int)((A.off[(r >> 6) + 1] - A.off[r >> 6]) >> 6);
  for (int k = len; k < w; k++) v32[base + (long long)k * kSlice] = 0.0f;
}

// ------------------------------------------------------------- finishers
struct FinBnorm {
  KS *st;
  __device__ bool skip() const { return false; }
  template <class T> __device__ void operator()(const T *a) const { st->bnrm = sqrt((double)a[0]); }
};
struct FinResidual {
  KS *st;
  __device__ bool skip() const { return false; }
  template <class T> __device__ void operator()(const T *a) const {
    st->rr = a[0];
    st->res = sqrt((double)a[0]) / st->bnrm;
    st->rmax = sqrt((double)a[0]);
    if (st->res <= st->tol) st->done = 1;
  }
};
// r.z of the two-level preconditioner: the coarse part c.nu is added when
// the coarse space is active (kept separate so the one-level path is
// unchanged bit for bit).
__device__ __forceinline__ double rz_total(const KS *st, double a) { return st->coarse ? a + st->cnu : a; }
// r.z formed by k_pcg_p_split waits in st->rz_next (st->rzp = 1) until the
// next finisher that reads st->rz takes it over.
__device__ __forceinline__ void take_rz(KS *st) {
  if (st->rzp) {
    st->rz = st->rz_next;
    st->rzp = 0;
  }
}

struct FinNop {
  KS *st;
  __device__ bool skip() const { return st->done == 1; }
  template <class T> __device__ void operator()(const T *) const {}
};
struct FinRz0 {
  KS *st;
  __device__ bool skip() const { return st->done != 0; }
  template <class T> __device__ void operator()(const T *a) const { st->rz = rz_total(st, (double)a[0]); }
};
struct FinPap {
  KS *st;
  __device__ bool skip() const { return st->done != 0; }
  template <class T> __device__ void operator()(const T *a) const {
    take_rz(st);
    st->it = st->it + 1;
    st->pap = a[0];
    if (fabs((double)a[0]) < kTiny) {
      st->ierr = 2;
      st->done = 1;
    } else {
      st->alpha = st->rz / (double)a[0];
    }
  }
};
struct FinUpdate {
  KS *st;
  int jac;
  __device__ bool skip() const { return st->done != 0; }
  template <class T> __device__ void operator()(const T *a) const {
    st->rr = a[0];
    st->res = sqrt((double)a[0]) / st->bnrm;
    if (st->res <= st->tol) {
      st->done = 1;
      return;
    }
    if (jac) {
      if (fabs(st->rz) < kTiny) {
        st->ierr = 2;
        st->done = 1;
      } else {
        const double rzn = rz_total(st, (double)a[1]);
        st->beta = rzn / st->rz;
        st->rz = rzn;
      }
    }
  }
};
struct FinRz {
  KS *st;
  __device__ bool skip() const { return st->done != 0; }
  template <class T> __device__ void operator()(const T *a) const {
    if (st->ru == 2) st->ru = 0;           // reliable update fully applied
    if (fabs(st->rz) < kTiny) {
      st->ierr = 2;
      st->done = 1;
    } else {
      const double rzn = rz_total(st, (double)a[0]);
      st->beta = rzn / st->rz;
      st->rz = rzn;
    }
  }
};
// Mixed PCG runs the FP32 iteration on r/sigma (sigma = |r| at the last
// reliable update) so its vectors stay O(1): a p' right-hand side can be
// ~1e-14, which would otherwise push FP32 products into the denormal range.
// Sloppy residual check of the FP32 iteration.  When the recursive residual
// meets the tolerance the iteration pauses (done = 2) and the host runs a
// reliable update at once; when it has dropped below delta times its largest
// value since the last reliable update, an update is flagged (ru = 1) and
// the host applies it at the end of the chunk.  `trigger` is 0 on the pass
// that follows a paused update.  jac: Jacobi r.z from the second partial.
struct FinUpdateRU {
  KS *st;
  int jac, trigger;
  __device__ bool skip() const { return st->done != 0; }
  template <class T> __device__ void operator()(const T *a) const {
    const double rn = sqrt((double)a[0]);              // scaled sloppy |r|
    st->res = rn * st->sigma / st->bnrm;
    if (trigger) {
      if (st->res <= st->tol) {            // pause: the host verifies convergence in FP64
        st->ru = 1;
        st->done = 2;
        return;
      }
      if (rn < st->delta * st->rmax) st->ru = 1;   // applied at the end of the chunk
    }
    if (rn > st->rmax) st->rmax = rn;
    if (!trigger) st->ru = 0;              // rescale pass done
    if (jac) {
      if (fabs(st->rz) < kTiny) {
        st->ierr = 2;
        st->done = 1;
      } else {
        const double rzn = rz_total(st, (double)a[1]);
        st->beta = rzn / st->rz;
        st->rz = rzn;
      }
    }
  }
};
// FinUpdateRU for k_finish_split (Jacobi): the same residual checks; the
// Jacobi part of r.z goes to st->rzd, and beta is formed by k_pcg_p_split
// once the coarse parts c.nu are in.
struct FinUpdateSplit {
  KS *st;
  int trigger;
  __device__ bool skip() const { return st->done != 0; }
  template <class T> __device__ void operator()(const T *a) const {
    take_rz(st);   // the r.z this iteration's alpha used (k_pcg_update_fa)
    const double rn = sqrt((double)a[0]);
    st->res = rn * st->sigma / st->bnrm;
    if (trigger) {
      if (st->res <= st->tol) {
        st->ru = 1;
        st->done = 2;
        return;
      }
      if (rn < st->delta * st->rmax) st->ru = 1;
    }
    if (rn > st->rmax) st->rmax = rn;
    if (!trigger) st->ru = 0;
    if (fabs(st->rz) < kTiny) {
      st->ierr = 2;
      st->done = 1;
    } else {
      st->rzd = (double)a[1];
      st->pup = 1;   // the p update of this iteration is pending (k_pcg_p_split or the fused product)
    }
  }
};
struct FinRU {         // after x += sigma*y and r = b - A x in FP64
  KS *st;
  __device__ bool skip() const { return st->done == 1 || st->ru != 1; }
  template <class T> __device__ void operator()(const T *a) const {
    take_rz(st);
    const double rt = sqrt((double)a[0]);
    st->rr = a[0];
    st->res = rt / st->bnrm;
    st->nru = st->nru + 1;
    if (st->res <= st->tol) {
      st->done = 1;
      return;
    }
    st->scale = st->sigma / rt;            // p and r.z move to the new scale
    st->rz = st->rz * st->scale * st->scale;
    st->sigma = rt;
    st->rmax = 1.0;
    st->ru = 2;                            // k_ru_rescale pending
    st->done = 0;                          // resume the iteration
  }
};
// DILU iterations finish the update's r.r (first of the three chunk partials)
// and the sweeps' r.z in one kernel after the sweeps: the convergence and
// reliable-update checks of r.r move behind the preconditioner, which leaves
// the iterates unchanged (one unneeded sweep on the last iteration).
struct FinUpdRz {        // FP64 PCG: FinUpdate then FinRz
  KS *st;
  __device__ bool skip() const { return st->done != 0; }
  template <class T> __device__ void operator()(const T *a1, const T *a2) const {
    st->rr = a1[0];
    st->res = sqrt((double)a1[0]) / st->bnrm;
    if (st->res <= st->tol) {
      st->done = 1;
      return;
    }
    if (fabs(st->rz) < kTiny) {
      st->ierr = 2;
      st->done = 1;
    } else {
      const double rzn = rz_total(st, (double)a2[0]);
      st->beta = rzn / st->rz;
      st->rz = rzn;
    }
  }
};
struct FinUpdRzRU {      // mixed PCG: FinUpdateRU then FinRz
  KS *st;
  int trigger;
  __device__ bool skip() const { return st->done != 0; }
  template <class T> __device__ void operator()(const T *a1, const T *a2) const {
    const double rn = sqrt((double)a1[0]);
    st->res = rn * st->sigma / st->bnrm;
    if (trigger) {
      if (st->res <= st->tol) {
        st->ru = 1;
        st->done = 2;
        return;
      }
      if (rn < st->delta * st->rmax) st->ru = 1;
    }
    if (rn > st->rmax) st->rmax = rn;
    if (st->ru == 2) st->ru = 0;
    if (fabs(st->rz) < kTiny) {
      st->ierr = 2;
      st->done = 1;
    } else {
      const double rzn = rz_total(st, (double)a2[0]);
      st->beta = rzn / st->rz;
      st->rz = rzn;
    }
  }
};

template <class T, class Fin>
__global__ __launch_bounds__(kFinThreads) void k_finish2(int n1, const T *p1, int n2, const T *p2, Fin fin) {
  if (fin.skip()) return;
  T a1[3], a2[1];
  fin_gather<OpSum, 3, T>(n1, p1, a1);
  fin_gather<OpSum, 1, T>(n2, p2, a2);
  __shared__ T l1[3 * (kFinThreads / 64)], l2[kFinThreads / 64];
  block_reduce<OpSum, 3, T, kFinThreads>(a1, l1);
  block_reduce<OpSum, 1, T, kFinThreads>(a2, l2);
  if (threadIdx.x == 0) fin(a1, a2);
}

struct FinBiRho {
  KS *st;
  __device__ bool skip() const { return st->done != 0; }
  template <class T> __device__ void operator()(const T *a) const {
    st->it = st->it + 1;
    st->rho = a[0];
    if (fabs((double)a[0]) < kTiny) {
      st->ierr = 2;
      st->done = 1;
    } else {
      st->beta = ((double)a[0] / st->rho_old) * (st->alpha / st->omega);
    }
  }
};
struct FinBiAlpha {
  KS *st;
  __device__ bool skip() const { return st->done != 0; }
  template <class T> __device__ void operator()(const T *a) const { st->alpha = st->rho / (double)a[0]; }
};
struct FinBiS {
  KS *st;
  __device__ bool skip() const { return st->done != 0; }
  template <class T> __device__ void operator()(const T *a) const {
    st->snrm = sqrt((double)a[0]);
    if (st->snrm / st->bnrm <= st->tol) {
      st->res = st->snrm / st->bnrm;
      st->done = 2;
    }
  }
};
struct FinBiOmega {
  KS *st;
  __device__ bool skip() const { return st->done != 0; }
  template <class T> __device__ void operator()(const T *a) const {
    st->tt = a[0];
    if ((double)a[0] < kTiny) {
      st->ierr = 2;
      st->done = 1;
      return;
    }
    st->omega = (double)a[1] / (double)a[0];
    if (fabs(st->omega) < kTiny) {
      st->ierr = 2;
      st->done = 1;
    }
  }
};
struct FinBiXR {
  KS *st;
  __device__ bool skip() const { return st->done != 0; }
  template <class T> __device__ void operator()(const T *a) const {
    st->rr = a[0];
    st->res = sqrt((double)a[0]) / st->bnrm;
    if (st->res <= st->tol) {
      st->done = 1;
      return;
    }
    st->rho_old = st->rho;
  }
};

// ------------------------------------------------------------ basic kernels
template <class T>
__global__ void k_spmv(int n, Mat A, const T *val, const T *x, T *y) {
  int r = blockIdx.x * kBlock + threadIdx.x;
  if (r < n) y[r] = row_dot(r, A, val, x);
}

template <class T>
__global__ void k_extract_diag(int n, Mat A, const T *val, T *d) {
  int r = blockIdx.x * kBlock + threadIdx.x;
  if (r < n) d[r] = val[A.off[r >> 6] + (r & 63) + (long long)A.diagk[r] * kSlice];
}

__global__ void k_to_float(long long n, const double *a, float *b) {
  long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

template <class T>
__global__ void k_sumsq(int n, const T *b, T *part) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  T v[1] = {i < n ? b[i] * b[i] : T(0)};
  block_partials<OpSum, 1>(v, part);
}

template <class T>
__global__ void k_residual(int n, Mat A, const T *val, const T *b, const T *x, T *r, T *part) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  T v[1] = {T(0)};
  if (i < n) {
    T ri = b[i] - row_dot(i, A, val, x);
    r[i] = ri;
    v[0] = ri * ri;
  }
  block_partials<OpSum, 1>(v, part);
}

// ------------------------------------------------ preconditioner set-up
template <class T>
__global__ void k_dilu_setup(int r0, int r1, Mat A, const long long *tpos, const T *val, T *d) {
  int i = r0 + blockIdx.x * kBlock + threadIdx.x;
  if (i >= r1) return;
  const long long bi = A.off[i >> 6] + (i & 63);
  const int dpi = A.diagk[i];
  T di = val[bi + (long long)dpi * kSlice];
  for (int k = 0; k < dpi; k++) {
    const long long q = bi + (long long)k * kSlice;
    di = di - val[q] * val[tpos[q]] / d[A.col[q]];
  }
  d[i] = di;
}

// ------------------------------------------------------------ coarse space
// Finisher of the update reduction with the coarse solve (one workgroup):
// the scalar sums as k_finish, then c_a = sum of the last partial over the
// chunks of aggregate a (fixed order), nu = E^-1 c and c.nu, before the
// functor runs on thread 0.

// Coarse solve riding along a DILU forward-sweep launch: nwg extra
// workgroups of kBlock threads, each recomputing c = Z^T r from the chunk
// partials (third of three values per chunk) and four rows of nu = E^-1 c;
// c.nu partials go to cnu_part[g], summed with the sweep's r.z partials.
template <class T>
struct CoWork {
  int k, nwg;
  const int *agg_cptr, *agg_cidx;
  const float *Einv;
  double *nu;
  const T *cpart;
  T *cnu_part;
};

template <class T>
__device__ void coarse_ride(const CoWork<T> &cw, int g) {
  __shared__ double c[kMaxAgg];
  const int t = threadIdx.x;
  for (int b = t; b < cw.k; b += kBlock) {
    const int j0 = cw.agg_cptr[b], j1 = cw.agg_cptr[b + 1];
    double s = 0.0;
    int j = j0;
    for (; j + 7 < j1; j += 8) {       // eight independent far loads in flight, summed in index order
      T x[8];
#pragma unroll
      for (int u = 0; u < 8; u++) x[u] = cw.cpart[cw.agg_cidx[j + u] * 3 + 2];
#pragma unroll
      for (int u = 0; u < 8; u++) s = s + (double)x[u];
    }
    for (; j < j1; j++) s = s + (double)cw.cpart[cw.agg_cidx[j] * 3 + 2];
    c[b] = s;
  }
  __syncthreads();
  const int w = t >> 6, lane = t & 63, a = g * (kBlock / 64) + w;
  double acc = 0.0;
  if (a < cw.k)
    for (int b = lane; b < cw.k; b += 64) acc = acc + (double)cw.Einv[(size_t)a * cw.k + b] * c[b];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc = acc + __shfl_down(acc, o, 64);
  __shared__ double pr[kBlock / 64];
  if (lane == 0) {
    double cn = 0.0;
    if (a < cw.k) {
      cw.nu[a] = acc;
      cn = acc * c[a];
    }
    pr[w] = cn;
  }
  __syncthreads();
  if (t == 0) cw.cnu_part[g] = (T)(((pr[0] + pr[1]) + pr[2]) + pr[3]);
}

// Coarse solve by the whole finisher workgroup (kFinThreads threads): Q
// threads per aggregate (Q = 4, 2, 1 for k <= 256, 512, 1024; part q =
// t / (kFinThreads / Q)) each sum every Q-th chunk partial (the last of the
// NV values) and one Q-th of the E^-1 columns; the parts are combined in q
// order.  Writes nu and co.st->cnu.
// c = Z^T r into c[0, k) (LDS) by the whole workgroup, Q threads per
// aggregate as described above; ends with the workgroup synchronised.
template <int NV, class T>
__device__ void coarse_c(const T *partial, const CoDev &co, double *c, double *partq) {
  const int Q = co.k <= kFinThreads / 4 ? 4 : (co.k <= kFinThreads / 2 ? 2 : 1);
  const int span = kFinThreads / Q;
  const int a = threadIdx.x % span, q = threadIdx.x / span;
  if (a < co.k) {
    // four chunk indices, then their four partials, in flight at once; the
    // sum runs in chunk order
    double s = 0.0;
    int j = co.agg_cptr[a] + q;
    const int je = co.agg_cptr[a + 1];
    for (; j + 3 * Q < je; j += 4 * Q) {
      const int i0 = co.agg_cidx[j], i1 = co.agg_cidx[j + Q], i2 = co.agg_cidx[j + 2 * Q], i3 = co.agg_cidx[j + 3 * Q];
      const T v0 = partial[i0 * NV + NV - 1], v1 = partial[i1 * NV + NV - 1];
      const T v2 = partial[i2 * NV + NV - 1], v3 = partial[i3 * NV + NV - 1];
      s = s + (double)v0;
      s = s + (double)v1;
      s = s + (double)v2;
      s = s + (double)v3;
    }
    for (; j < je; j += Q) s = s + (double)partial[co.agg_cidx[j] * NV + NV - 1];
    partq[q * span + a] = s;
  }
  __syncthreads();
  if (q == 0 && a < co.k) {
    double s = partq[a];
    for (int qq = 1; qq < Q; qq++) s = s + partq[qq * span + a];
    c[a] = s;
  }
  __syncthreads();
}

template <int NV, class T>
__device__ void coarse_solve(const T *partial, const CoDev &co) {
  __shared__ double c[kMaxAgg], partq[kFinThreads];
  coarse_c<NV, T>(partial, co, c, partq);
  const int Q = co.k <= kFinThreads / 4 ? 4 : (co.k <= kFinThreads / 2 ? 2 : 1);
  const int span = kFinThreads / Q;
  const int a = threadIdx.x % span, q = threadIdx.x / span;
  const int kq = (co.k + Q - 1) / Q, b0 = q * kq, b1 = min(co.k, b0 + kq);
  if (a < co.k) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;    // E^-1 is symmetric: column a
    int b = b0;
    // 16 loads in flight per thread (the loop is latency-bound); accumulator
    // u % 4 takes the entries b = u mod 4 in ascending order, as below
    for (; b + 15 < b1; b += 16) {
      float e[16];
#pragma unroll
      for (int u = 0; u < 16; u++) e[u] = co.Einv[(size_t)(b + u) * co.k + a];
#pragma unroll
      for (int u = 0; u < 16; u += 4) {
        s0 = s0 + (double)e[u] * c[b + u];
        s1 = s1 + (double)e[u + 1] * c[b + u + 1];
        s2 = s2 + (double)e[u + 2] * c[b + u + 2];
        s3 = s3 + (double)e[u + 3] * c[b + u + 3];
      }
    }
    for (; b + 3 < b1; b += 4) {
      const float e0 = co.Einv[(size_t)b * co.k + a], e1 = co.Einv[(size_t)(b + 1) * co.k + a];
      const float e2 = co.Einv[(size_t)(b + 2) * co.k + a], e3 = co.Einv[(size_t)(b + 3) * co.k + a];
      s0 = s0 + (double)e0 * c[b];
      s1 = s1 + (double)e1 * c[b + 1];
      s2 = s2 + (double)e2 * c[b + 2];
      s3 = s3 + (double)e3 * c[b + 3];
    }
    for (; b < b1; b++) s0 = s0 + (double)co.Einv[(size_t)b * co.k + a] * c[b];
    partq[q * span + a] = ((s0 + s1) + s2) + s3;
  }
  __syncthreads();
  double prod[1] = {0.0};
  if (q == 0 && a < co.k) {
    double nu = partq[a];
    for (int qq = 1; qq < Q; qq++) nu = nu + partq[qq * span + a];
    co.nu[a] = nu;
    prod[0] = nu * c[a];
  }
  __shared__ double lds2[kFinThreads / 64];
  block_reduce<OpSum, 1, double, kFinThreads>(prod, lds2);
  if (threadIdx.x == 0) co.st->cnu = prod[0];
}

template <int NV, class T, class Fin>
__global__ __launch_bounds__(kFinThreads) void k_finish_co(int nparts, const T *partial, Fin fin, CoDev co) {
  if (fin.skip()) return;
  T acc[NV];
  fin_gather<OpSum, NV, T>(nparts, partial, acc);
  __shared__ T lds[NV * (kFinThreads / 64)];
  block_reduce<OpSum, NV, T, kFinThreads>(acc, lds);
  if (co.k > 0) coarse_solve<NV, T>(partial, co);
  if (threadIdx.x == 0) fin(acc);
}

// Finisher of the mixed Jacobi PCG with a coarse space, spread over 1 + G
// workgroups (G = ceil(k / kSplitRows)) so the coarse solve no longer runs
// in one workgroup: workgroup 0 reduces the update's scalar partials and
// runs the functor; workgroup g >= 1 recomputes c = Z^T r from the chunk
// partials, forms rows [kSplitRows (g-1), kSplitRows g) of nu = E^-1 c (one
// wavefront per row, E^-1 read along its rows) and writes the c.nu sum of
// its rows to cnu_part[g-1].  k_pcg_p_split adds the parts (fixed order).
constexpr int kSplitRows = kFinThreads / 64;
constexpr int kSplitMaxParts = (kMaxAgg + kSplitRows - 1) / kSplitRows;   // coarse workgroups at most (64)
template <int NV, class T, class Fin>
__global__ __launch_bounds__(kFinThreads) void k_finish_split(int nparts, const T *partial, Fin fin, CoDev co,
                                                              double *cnu_part) {
  if (fin.skip()) return;
  if (blockIdx.x == 0) {
    T acc[NV];
    fin_gather<OpSum, NV, T>(nparts, partial, acc);
    __shared__ T lds[NV * (kFinThreads / 64)];
    block_reduce<OpSum, NV, T, kFinThreads>(acc, lds);
    if (threadIdx.x == 0) fin(acc);
    return;
  }
  __shared__ double c[kMaxAgg], partq[kFinThreads], pr[kSplitRows];
  coarse_c<NV, T>(partial, co, c, partq);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int a = (blockIdx.x - 1) * kSplitRows + w;
  double prod = 0.0;
  if (a < co.k) {
    double s = 0.0;
    for (int b = lane; b < co.k; b += 64) s = s + (double)co.Einv[(size_t)a * co.k + b] * c[b];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s = s + __shfl_down(s, o, 64);
    if (lane == 0) {
      co.nu[a] = s;
      prod = s * c[a];
    }
  }
  if (lane == 0) pr[w] = prod;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int j = 0; j < kSplitRows; j++) t = t + pr[j];
    cnu_part[blockIdx.x - 1] = t;
  }
}

__global__ void k_fill64(long long n, double *v, double x) {
  long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) v[i] = x;
}

// E = Z^T A Z: one wavefront per nonzero (a,b) sums its matrix slots (listed
// in ascending order) and writes E[a*k+b]; absent entries stay zero.
__global__ void k_coarse_E(int nE, const int *e_ptr, const int *e_key, const int *e_slot, const double *val,
                           double *E) {
  const int e = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (e >= nE) return;
  double s = 0.0;
  for (int j = e_ptr[e] + lane; j < e_ptr[e + 1]; j += 64) s = s + val[e_slot[j]];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s = s + __shfl_down(s, o, 64);
  if (lane == 0) E[e_key[e]] = s;
}

// E from the slot map in two passes, for entries with many slots: kEParts
// wavefronts per entry each sum an interleaved share of its slots (lane
// order, then shuffles), and one thread per entry adds the parts in part
// order; the fixed order keeps E deterministic, and the entry's slots are
// spread over kEParts times more wavefronts than in k_coarse_E.
constexpr int kEParts = 16;
__global__ void k_coarse_E_parts(int nE, const int *e_ptr, const int *e_slot, const double *val, double *parts) {
  const long long w = ((long long)blockIdx.x * kBlock + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  const long long e = w / kEParts;
  const int part = (int)(w % kEParts);
  if (e >= nE) return;
  double s = 0.0;
  for (int j = e_ptr[e] + part * 64 + lane; j < e_ptr[e + 1]; j += 64 * kEParts) s = s + val[e_slot[j]];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s = s + __shfl_down(s, o, 64);
  if (lane == 0) parts[e * kEParts + part] = s;
}
__global__ void k_coarse_E_sum(int nE, const int *e_key, const double *parts, double *E) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= nE) return;
  double s = parts[(long long)e * kEParts];
  for (int q = 1; q < kEParts; q++) s = s + parts[(long long)e * kEParts + q];
  E[e_key[e]] = s;
}

// Gauss-Jordan inversion of the SPD matrix E (no pivoting) in FP64 over the
// whole grid, one launch per pivot j: W holds the current matrix; row j and
// column j of it come from snapshots written by the previous launch (rowj,
// colj), so every entry is read and written by its own thread only, and the
// threads of row / column j + 1 write the next snapshots.
//   a_jj <- 1/p,  a_jc <- a_jc/p,  a_rj <- -a_rj/p,  a_rc <- a_rc - a_rj a_jc/p.
__global__ void k_gj_init(int k, const double *E, double *W, double *row0, double *col0) {
  const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (t >= (long long)k * k) return;
  const int r = (int)(t / k), c = (int)(t % k);
  const double v = E[t];
  W[t] = v;
  if (r == 0) row0[c] = v;
  if (c == 0) col0[r] = v;
}

__global__ void k_gj_step(int k, int j, double *W, const double *rowj, const double *colj, double *rown,
                          double *coln) {
  const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (t >= (long long)k * k) return;
  const int r = (int)(t / k), c = (int)(t % k);
  const double inv = 1.0 / rowj[j];
  const double rowc = rowj[c] * inv;
  double v;
  if (r == j) v = (c == j) ? inv : rowc;
  else if (c == j) v = -colj[r] * inv;
  else v = W[t] - colj[r] * rowc;
  W[t] = v;
  if (r == j + 1) rown[c] = v;
  if (c == j + 1) coln[r] = v;
}

__global__ void k_gj_out(long long kk, const double *W, float *Einv) {
  const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (t < kk) Einv[t] = (float)W[t];
}

// In-place Gauss-Jordan inversion of the SPD matrix E (k <= 256, no
// pivoting needed) in FP32, one workgroup of 1024 threads holding the
// padded 256 x 256 matrix in registers: thread t owns column t % 256 of rows
// t/256 + 4m, m = 0..63.  Pivot step j with pivot p = a_jj:
//   a_jj <- 1/p,  a_jc <- a_jc/p,  a_rj <- -a_rj/p,  a_rc <- a_rc - a_rj a_jc/p.
// While updating, the owners of row j+1 and column j+1 publish them to a
// second LDS buffer (static register indices only), so each step needs one
// barrier and no dynamically indexed register array.
constexpr int kInvThreads = 1024;
__global__ __launch_bounds__(kInvThreads) void k_coarse_inv(int k, const double *E, float *Einv) {
  const int t = threadIdx.x, c = t & 255, rb = t >> 8;
  float a[64];
  __shared__ float prow[2][256], pcol[2][256];
#pragma unroll
  for (int m = 0; m < 64; m++) {
    const int r = rb + 4 * m;
    a[m] = (r < k && c < k) ? (float)E[(size_t)r * k + c] : (r == c ? 1.0f : 0.0f);
    if (r == 0) prow[0][c] = a[m];
    if (c == 0) pcol[0][r] = a[m];
  }
  __syncthreads();
  for (int j = 0; j < k; j++) {
    const int cur = j & 1, nxt = cur ^ 1, jn = j + 1;
    const float inv = 1.0f / prow[cur][j];
    const float rowc = prow[cur][c] * inv;
#pragma unroll
    for (int m = 0; m < 64; m++) {
      const int r = rb + 4 * m;
      float v;
      if (r == j) v = (c == j) ? inv : rowc;
      else if (c == j) v = -pcol[cur][r] * inv;
      else v = a[m] - pcol[cur][r] * rowc;
      a[m] = v;
      if (r == jn) prow[nxt][c] = v;
      if (c == jn) pcol[nxt][r] = v;
    }
    __syncthreads();
  }
#pragma unroll
  for (int m = 0; m < 64; m++) {
    const int r = rb + 4 * m;
    if (r < k && c < k) Einv[(size_t)r * k + c] = a[m];
  }
}

}  // namespace

static inline Mat matof(const SellPattern *p) {
  return Mat{p->d_off.p, p->d_rowlen.p, p->d_col.p, p->d_diagk.p, p->has16 ? p->d_col16.p : nullptr,
             p->has16 ? p->d_cbase.p : nullptr};
}

}  // namespace ms
