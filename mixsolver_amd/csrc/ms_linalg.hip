// ms_linalg.hip -- SELL-64 sparse kernels and device-resident Krylov solvers
// for gfx950.
//
// Recurrences follow the reference solvers at one partition
// (mod_uns_linsolver_mpi.f90: bicgstab_ilu0_mpi :155-229, cg_jacobi_mpi
// :235-300, cg_ic0_mpi :306-363; factorizations mod_uns_linsolver.f90:79-142
// and :328-399).  Scalars (alpha, beta, rho, omega, residuals, iteration
// count, breakdown flags) live in device memory; they are updated by the
// one-workgroup finisher of each two-stage reduction, so an iteration needs
// no host round trip.  The host enqueues iterations in chunks and reads the
// state once per chunk; every iteration kernel returns at once when the
// state says done.
//
// Kernel fusion per PCG iteration (FP64 path):
//   Jacobi: [q = A p, p.q] [x,r update, r.r, r.(r/d)] [p = r/d + beta p]
//   DILU  : [q = A p, p.q] [x,r update, r.r, z = r/d on colour 0]
//           [forward sweeps colours 1..c-1, backward sweeps c-2..0, r.z]
//           [p = z + beta p]
// Vectors are templated on the value type: double for the FP64 path, float
// for the inner iterations of mixed precision (FP32 products and FP32
// accumulation of dot-product partials; scalar recurrences in double).
#include <atomic>
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include <hip/hip_ext.h>

#include "ms_kernels.hpp"
#include "ms_linalg.hpp"

namespace ms {
namespace {

// ------------------------------------------------------------------- PCG
// p = z (+ nu of the row's aggregate); partial r.z of the one-level part
template <class T>
__global__ void k_pcg_start(Rows R, const T *r, const T *z, const T *dg, int jac, T *p, const double *nu,
                            const KS *st, T *part) {
  if (st->done) return;
  int i, a;
  T v[1] = {T(0)};
  if (R.get(i, a)) {
    T zi = jac ? r[i] / dg[i] : z[i];
    p[i] = nu ? zi + (T)nu[a] : zi;
    v[0] = r[i] * zi;
  }
  block_partials<OpSum, 1>(v, part);
}

// partials (0, 0, sum r) per row set: Z^T r of the initial residual
template <class T>
__global__ void k_rsum(Rows R, const T *r, const KS *st, T *part) {
  if (st->done == 1) return;
  int i, a;
  T v[3] = {T(0), T(0), T(0)};
  if (R.get(i, a)) v[2] = r[i];
  block_partials<OpSum, 3>(v, part);
}

template <class T, int C16 = -1>
__global__ void k_pcg_spmv(int n, Mat A, const T *val, const T *p, T *q, const KS *st, T *part) {
  if (st->done) return;
  int i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  T v[1] = {T(0)};
  if (i < n) {
    const T pi = p[i];   // read before the store to q, which may alias p for the compiler
    const T qi = row_apply<T, C16>(i, A, val, p);
    q[i] = qi;
    v[0] = pi * qi;
  }
  block_partials<OpSum, 1>(v, part);
}

// Matrix-free S p (MatFree): with g_k = sum_j (-a_kj / a_kk) x_j over the
// colour-1 neighbours of colour-0 row k, (S x)_i = sum_k (-a_ik) (x_i - g_k)
// + rowsum_i x_i.  Stage 1 forms, per colour-0 row, h_k = x_ref - g_k =
// sum_j w_kj (x_ref - x_j) + rho_k x_ref with x_ref the row's first
// colour-1 neighbour and rho_k = rowsum_k / a_kk; stage 2 forms
// sum_k (-a_ik) ((x_i - x_ref_k) + h_k) + rowsum_i x_i.  Every term is a
// difference or a row-sum term, as in row_apply's difference form, so a
// constant vector gives the row sums' product exactly.  The slot loop runs
// the slice width of A (diagonal and padding slots have col = row).
// The pending p update of the split PCG (p = r/d + nu[agg] + beta p, as
// k_pcg_p_split) can ride along the product (PU): both stages compute beta
// from the coarse finisher's parts in every workgroup, stage 1 forms the
// updated p of each neighbour it reads, stage 2 that of its own row and
// stores it (stage 1 has finished reading the old p by then).  The values
// are those k_pcg_p_split would store, bit for bit.
__device__ __forceinline__ float pu_beta(KS *st, const double *cnu, int G, bool writer) {
  __shared__ double parts[kSplitMaxParts + 2];
  __shared__ float beta_f;
  const int t = threadIdx.x;
  if (t < G) parts[t] = cnu[t];
  else if (t == kSplitMaxParts) parts[t] = st->rzd;
  else if (t == kSplitMaxParts + 1) parts[t] = st->rz;
  __syncthreads();
  if (t == 0) {
    double rzn = parts[kSplitMaxParts];
    for (int g = 0; g < G; g++) rzn = rzn + parts[g];
    const double beta = rzn / parts[kSplitMaxParts + 1];
    beta_f = (float)beta;
    if (writer) {
      st->rz_next = rzn;
      st->rzp = 1;
      st->beta = beta;
    }
  }
  __syncthreads();
  return beta_f;
}
template <bool PU>
__device__ __forceinline__ float pu_val(int j, const float *p, const PUp &u, float beta, bool up) {
  if constexpr (PU) {
    if (up) {
      const float zj = u.r[j] / u.dg[j] + (float)u.nu[u.agg1[j]];
      return zj + beta * p[j];
    }
  }
  return p[j];
}

// x_ref (slot 1's column, the row's first colour-1 neighbour) is taken from
// the columns already loaded, so its gather is issued with the others
template <int C16, int W, bool PU>
__device__ __forceinline__ float mf1_row(int k, int s, long long base, int n0, const Mat &A, const float *w,
                                         const float *x, float &xr, const PUp &u, float beta, bool up) {
  int c[W];
  float a[W], xv[W];
#pragma unroll
  for (int j = 0; j < W; j++) c[j] = col_at<C16>(A, base + (long long)j * kSlice, k, s);
#pragma unroll
  for (int j = 0; j < W; j++) a[j] = w[base + (long long)j * kSlice];
#pragma unroll
  for (int j = 0; j < W; j++) xv[j] = c[j] < n0 ? 0.f : pu_val<PU>(c[j] - n0, x, u, beta, up);
  xr = c[1] >= n0 ? xv[1] : 0.f;
  float acc = 0.f;
#pragma unroll
  for (int j = 0; j < W; j++) acc = acc + a[j] * (c[j] < n0 ? xr : xr - xv[j]);
  return acc;
}

template <int C16, bool PU>
__global__ void k_mf_stage1(int n0, Mat A, const float *w, const float *x, float2 *hx, KS *st, PUp u) {
  if (st->done) return;
  float beta = 0.f;
  bool up = false;
  if constexpr (PU) {
    up = st->pup != 0;   // the same for every workgroup: pup changes only in k_pcg_update_fa
    if (up) beta = pu_beta(st, u.cnu, u.G, false);
  }
  const int k = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (k >= n0) return;
  const int s = k >> 6;
  const long long o0 = A.off[s];
  const int wd = (int)((A.off[s + 1] - o0) >> 6);
  const long long base = o0 + (k & 63);
  // slot 0 is the diagonal (colour-1 columns are above every colour-0 row)
  float xr = 0.f, acc;
  if (wd == 7) acc = mf1_row<C16, 7, PU>(k, s, base, n0, A, w, x, xr, u, beta, up);
  else if (wd == 6) acc = mf1_row<C16, 6, PU>(k, s, base, n0, A, w, x, xr, u, beta, up);
  else if (wd == 5) acc = mf1_row<C16, 5, PU>(k, s, base, n0, A, w, x, xr, u, beta, up);
  else {
    const int c1 = wd > 1 ? col_at<C16>(A, base + kSlice, k, s) : k;
    xr = c1 >= n0 ? pu_val<PU>(c1 - n0, x, u, beta, up) : 0.f;
    acc = 0.f;
    for (int j = 0; j < wd; j++) {
      const long long q = base + (long long)j * kSlice;
      const int c = col_at<C16>(A, q, k, s);
      acc = acc + w[q] * (c < n0 ? xr : xr - pu_val<PU>(c - n0, x, u, beta, up));
    }
  }
  hx[k] = make_float2(acc, xr);
}

template <int C16, int W>
__device__ __forceinline__ float mf2_row(int r, int s, long long base, int n0, const Mat &A, const float *w,
                                         const float2 *hx, float xi) {
  int c[W];
  float a[W];
#pragma unroll
  for (int j = 0; j < W; j++) c[j] = col_at<C16>(A, base + (long long)j * kSlice, r, s);
#pragma unroll
  for (int j = 0; j < W; j++) a[j] = w[base + (long long)j * kSlice];
  float2 h[W];
#pragma unroll
  for (int j = 0; j < W; j++) h[j] = c[j] < n0 ? hx[c[j]] : make_float2(0.f, 0.f);
  float acc = 0.f;
#pragma unroll
  for (int j = 0; j < W; j++) acc = acc + a[j] * (c[j] < n0 ? (xi - h[j].y) + h[j].x : xi);
  return acc;
}

template <int C16, bool PU>
__global__ void k_mf_stage2(int n0, int n1, Mat A, const float *w, float *x, const float2 *hx, float *q, KS *st,
                            float *part, PUp u) {
  if (st->done) return;
  float beta = 0.f;
  bool up = false;
  if constexpr (PU) {
    up = st->pup != 0;
    if (up) beta = pu_beta(st, u.cnu, u.G, blockIdx.x == 0);
  }
  const int i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  float v[1] = {0.f};
  if (i < n1) {
    const int r = n0 + i, s = r >> 6;
    const long long o0 = A.off[s];
    const int wd = (int)((A.off[s + 1] - o0) >> 6);
    const long long base = o0 + (r & 63);
    const float xi = pu_val<PU>(i, x, u, beta, up);
    if (up) x[i] = xi;
    float acc;
    if (wd == 7) acc = mf2_row<C16, 7>(r, s, base, n0, A, w, hx, xi);
    else if (wd == 6) acc = mf2_row<C16, 6>(r, s, base, n0, A, w, hx, xi);
    else if (wd == 5) acc = mf2_row<C16, 5>(r, s, base, n0, A, w, hx, xi);
    else {
      acc = 0.f;
      for (int j = 0; j < wd; j++) {
        const long long qq = base + (long long)j * kSlice;
        const int c = col_at<C16>(A, qq, r, s);
        if (c < n0) {
          const float2 h = hx[c];
          acc = acc + w[qq] * ((xi - h.y) + h.x);
        } else {
          acc = acc + w[qq] * xi;
        }
      }
    }
    q[i] = acc;
    v[0] = xi * acc;
  }
  block_partials<OpSum, 1>(v, part);
}

// x += alpha p; r -= alpha q; partials r.r, (Jacobi) r.(r/d) and sum r
// (Z^T r when the rows are aggregate chunks).
// mode 0 plain, 1 Jacobi, 2 DILU: z = r/d on the first colour block [0,c0end)
template <class T, class X>
__global__ void k_pcg_update(Rows R, const T *p, const T *q, X *x, T *r, T *z, const T *dg, int mode, int c0end,
                             const KS *st, T *part) {
  if (st->done) return;
  int i, a;
  T v[3] = {T(0), T(0), T(0)};
  if (R.get(i, a)) {
    // every load before the first store: the compiler cannot tell the
    // pointers apart, and loads after a store would wait for it
    const T alpha = (T)st->alpha;
    const T pi = p[i], qi = q[i], r0 = r[i];
    const X xi = x[i];
    const T di = (mode == 1 || (mode == 2 && i < c0end)) ? dg[i] : T(1);
    x[i] = xi + (X)(alpha * pi);
    const T ri = r0 - alpha * qi;
    r[i] = ri;
    v[0] = ri * ri;
    v[2] = ri;
    if (mode == 1) v[1] = ri * (ri / di);
    else if (mode == 2 && i < c0end) z[i] = ri / di;
  }
  block_partials<OpSum, 3>(v, part);
}

// p = z (+ nu of the row's aggregate) + beta p
template <class T>
__global__ void k_pcg_p(Rows R, const T *r, const T *z, const T *dg, int jac, T *p, const double *nu,
                        const KS *st) {
  if (st->done) return;
  int i, a;
  if (R.get(i, a)) {
    T zi = jac ? r[i] / dg[i] : z[i];
    if (nu) zi = zi + (T)nu[a];
    p[i] = zi + (T)st->beta * p[i];
  }
}

// k_pcg_update (Jacobi) with the alpha finisher folded in, for the mixed
// Jacobi PCG with a coarse space: every workgroup reduces the SpMV's p.q
// partials itself, in the same fixed order, so all hold the same pap and
// alpha (r.z is the value the last p update left pending, when it left one);
// workgroup 0 keeps the solver state as FinPap does (iteration count, pap,
// alpha, the breakdown test).  One launch and one kernel boundary less per
// iteration.
template <class T>
__global__ void k_pcg_update_fa(Rows R, int np, const T *ppart, const T *p, const T *q, T *y, T *r, const T *dg,
                                KS *st, T *part) {
  if (st->done) return;
  __shared__ T lds[kBlock / 64];
  __shared__ double alpha_s;
  __shared__ int bad_s;
  T acc[1] = {T(0)};
  int b = threadIdx.x;
  for (; b + 3 * kBlock < np; b += 4 * kBlock) {   // four loads in flight, summed in index order
    const T x0 = ppart[b], x1 = ppart[b + kBlock], x2 = ppart[b + 2 * kBlock], x3 = ppart[b + 3 * kBlock];
    acc[0] = acc[0] + x0;
    acc[0] = acc[0] + x1;
    acc[0] = acc[0] + x2;
    acc[0] = acc[0] + x3;
  }
  for (; b < np; b += kBlock) acc[0] = acc[0] + ppart[b];
  block_reduce<OpSum, 1, T, kBlock>(acc, lds);
  if (threadIdx.x == 0) {
    const double pap = (double)acc[0];
    const double rz = st->rzp ? st->rz_next : st->rz;
    const int bad = fabs(pap) < kTiny;
    bad_s = bad;
    alpha_s = bad ? 0.0 : rz / pap;
    if (blockIdx.x == 0) {
      st->it = st->it + 1;
      st->pap = pap;
      st->pup = 0;   // the product before this kernel applied any pending p update
      if (bad) {
        st->ierr = 2;
        st->done = 1;
      } else {
        st->alpha = alpha_s;
      }
    }
  }
  __syncthreads();
  if (bad_s) return;
  int i, a;
  T v[3] = {T(0), T(0), T(0)};
  if (R.get(i, a)) {
    const T alpha = (T)alpha_s;
    const T pi = p[i], qi = q[i], r0 = r[i], yi = y[i], di = dg[i];
    y[i] = yi + alpha * pi;
    const T ri = r0 - alpha * qi;
    r[i] = ri;
    v[0] = ri * ri;
    v[1] = ri * (ri / di);
    v[2] = ri;
  }
  block_partials<OpSum, 3>(v, part);
}

// p = r/d + nu_agg + beta p after k_finish_split: every workgroup forms
// r.z = rzd + the coarse parts (fixed order) and beta = r.z / rz; workgroup 0
// leaves r.z in st->rz_next for the next finisher.
template <class T>
__global__ void k_pcg_p_split(Rows R, const T *r, const T *dg, T *p, const double *nu, KS *st,
                              const double *cnu_part, int G, int flush = 0) {
  if (st->done) return;
  if (flush && !st->pup) return;   // flush: apply a pending update only
  // the parts, rzd and rz are loaded by different threads at once (one load
  // latency), then added by thread 0 in part order
  __shared__ double parts[kSplitMaxParts + 2], beta_s;
  const int t = threadIdx.x;
  if (t < G) parts[t] = cnu_part[t];
  else if (t == kSplitMaxParts) parts[t] = st->rzd;
  else if (t == kSplitMaxParts + 1) parts[t] = st->rz;
  __syncthreads();
  if (t == 0) {
    double rzn = parts[kSplitMaxParts];
    for (int g = 0; g < G; g++) rzn = rzn + parts[g];
    const double beta = rzn / parts[kSplitMaxParts + 1];
    beta_s = beta;
    if (blockIdx.x == 0) {
      st->rz_next = rzn;
      st->rzp = 1;
      st->beta = beta;
    }
  }
  __syncthreads();
  int i, a;
  if (R.get(i, a)) {
    const T zi = r[i] / dg[i] + (T)nu[a];
    p[i] = zi + (T)beta_s * p[i];
  }
}

__global__ void k_clear_pup(KS *st) {
  if (threadIdx.x == 0) st->pup = 0;
}

template <class T>
__global__ void k_dot(int n, const T *a, const T *b, const KS *st, T *part) {
  if (st->done) return;
  int i = blockIdx.x * kBlock + threadIdx.x;
  T v[1] = {i < n ? a[i] * b[i] : T(0)};
  block_partials<OpSum, 1>(v, part);
}

template <class T>
__global__ void k_jac_rz(int n, const T *r, const T *dg, T *z, const KS *st, T *part) {
  if (st->done) return;
  int i = blockIdx.x * kBlock + threadIdx.x;
  T v[1] = {T(0)};
  if (i < n) {
    T zi = r[i] / dg[i];
    z[i] = zi;
    v[0] = r[i] * zi;
  }
  block_partials<OpSum, 1>(v, part);
}

// ------------------------------------------------------- reliable updates
// x64 += sigma * y; y = 0   (only when a reliable update is pending)
__global__ void k_fold(int n, float *y, double *x, const KS *st) {
  if (st->done == 1 || st->ru != 1) return;
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) {
    x[i] = x[i] + st->sigma * (double)y[i];
    y[i] = 0.0f;
  }
}
// r64 = b - A64 x and r64.r64 (FP64)
__global__ void k_true_res(int n, Mat A, const double *val, const double *b, const double *x, double *r64,
                           const KS *st, double *part) {
  if (st->done == 1 || st->ru != 1) return;
  int i = blockIdx.x * kBlock + threadIdx.x;
  double v[1] = {0.0};
  if (i < n) {
    double ri = b[i] - row_dot(i, A, val, x);
    r64[i] = ri;
    v[0] = ri * ri;
  }
  block_partials<OpSum, 1>(v, part);
}
// r32 = r64 / sigma; p *= scale (after FinRU moved to the new scale), with
// the partials and colour-0 DILU values of k_pcg_update for the new r
__global__ void k_ru_rescale(Rows R, const double *r64, float *r32, float *p, float *z, const float *dg, int mode,
                             int c0end, const KS *st, float *part) {
  if (st->done || st->ru != 2) return;
  int i, a;
  float v[3] = {0.0f, 0.0f, 0.0f};
  if (R.get(i, a)) {
    const float ri = (float)(r64[i] / st->sigma);
    r32[i] = ri;
    p[i] = p[i] * (float)st->scale;
    v[0] = ri * ri;
    v[2] = ri;
    if (mode == 1) v[1] = ri * (ri / dg[i]);
    else if (mode == 2 && i < c0end) z[i] = ri / dg[i];
  }
  block_partials<OpSum, 3>(v, part);
}
// initial scaled residual r32 = r64 / sigma
__global__ void k_scale_down(int n, const double *r64, float *r32, const KS *st) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) r32[i] = (float)(r64[i] / st->sigma);
}
__global__ void k_set_sigma(KS *st) {
  if (threadIdx.x == 0) {
    st->sigma = sqrt(st->rr);
    st->rmax = 1.0;
  }
}

// -------------------------------------------------------------- BiCGSTAB
// Additive coarse correction of a BiCGSTAB preconditioner application,
// out += Z E^-1 Z^T in (the p' matrix and so E are symmetric): chunk sums of
// in (FP64, one workgroup per aggregate chunk), nu = E^-1 c by one finisher
// workgroup, then nu of each row's aggregate added to out.  All three skip a
// finished solve, as the other BiCGSTAB kernels do.
template <class T>
__global__ void k_co_restrict(Rows RW, const T *in, double *cpart, const int *done) {
  if (*done) return;
  int i, a;
  double v[1] = {0.0};
  if (RW.get(i, a)) v[0] = (double)in[i];
  block_partials<OpSum, 1, double>(v, cpart);
}
__global__ __launch_bounds__(kFinThreads) void k_co_solve(const double *cpart, CoDev co, const int *done) {
  if (*done) return;
  coarse_solve<1, double>(cpart, co);
}
template <class T>
__global__ void k_co_prolong(Rows RW, const double *nu, T *out, const int *done) {
  if (*done) return;
  int i, a;
  if (RW.get(i, a)) out[i] = out[i] + (T)nu[a];
}

template <class T>
__global__ void k_bi_init(int n, const T *r, T *r0, T *p, T *v, KS *st) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) {
    r0[i] = r[i];
    p[i] = T(0);
    v[i] = T(0);
  }
  if (i == 0) {
    st->rho_old = 1.0;
    st->alpha = 1.0;
    st->omega = 1.0;
  }
}

template <class T>
__global__ void k_bi_p(int n, const T *r, const T *v, T *p, const KS *st) {
  if (st->done) return;
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) p[i] = r[i] + (T)st->beta * (p[i] - (T)st->omega * v[i]);
}

template <class T, int C16 = -1>
__global__ void k_bi_spmv_dot(int n, Mat A, const T *val, const T *ph, const T *r0, T *v, const KS *st, T *part) {
  if (st->done) return;
  int i = blockIdx.x * kBlock + threadIdx.x;
  T w[1] = {T(0)};
  if (i < n) {
    const T r0i = r0[i];
    const T vi = row_apply<T, C16>(i, A, val, ph);
    v[i] = vi;
    w[0] = r0i * vi;
  }
  block_partials<OpSum, 1>(w, part);
}

template <class T>
__global__ void k_bi_s(int n, const T *r, const T *v, T *s, const KS *st, T *part) {
  if (st->done) return;
  int i = blockIdx.x * kBlock + threadIdx.x;
  T w[1] = {T(0)};
  if (i < n) {
    T si = r[i] - (T)st->alpha * v[i];
    s[i] = si;
    w[0] = si * si;
  }
  block_partials<OpSum, 1>(w, part);
}

template <class T>
__global__ void k_bi_early(int n, const T *ph, T *x, const KS *st) {
  if (st->done != 2) return;
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] = x[i] + (T)st->alpha * ph[i];
}

__global__ void k_done2_to_1(KS *st) {
  if (threadIdx.x == 0 && st->done == 2) st->done = 1;
}

template <class T, int C16 = -1>
__global__ void k_bi_spmv_omega(int n, Mat A, const T *val, const T *sh, const T *s, T *t, const KS *st, T *part) {
  if (st->done) return;
  int i = blockIdx.x * kBlock + threadIdx.x;
  T w[2] = {T(0), T(0)};
  if (i < n) {
    const T si = s[i];
    const T ti = row_apply<T, C16>(i, A, val, sh);
    t[i] = ti;
    w[0] = ti * ti;
    w[1] = ti * si;
  }
  block_partials<OpSum, 2>(w, part);
}

template <class T>
__global__ void k_bi_xr(int n, const T *ph, const T *sh, const T *s, const T *t, T *x, T *r, const KS *st, T *part) {
  if (st->done) return;
  int i = blockIdx.x * kBlock + threadIdx.x;
  T w[1] = {T(0)};
  if (i < n) {
    x[i] = x[i] + (T)st->alpha * ph[i] + (T)st->omega * sh[i];
    T ri = s[i] - (T)st->omega * t[i];
    r[i] = ri;
    w[0] = ri * ri;
  }
  block_partials<OpSum, 1>(w, part);
}

// ------------------------------------------------ preconditioner kernels
template <class T>
__global__ void k_jacobi(int n, const T *in, const T *dg, T *out, const int *done) {
  if (*done) return;
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) out[i] = in[i] / dg[i];
}

// natural IC(0) of row i, its lower neighbours final (mod_uns_linsolver.f90:328-365)
template <class T>
__device__ void ic0_factor_row(int i, const Mat &A, T *F, int *flag) {
  const long long bi = A.off[i >> 6] + (i & 63);
  const int dpi = A.diagk[i];
  for (int kp = 0; kp < dpi; kp++) {
    const long long qkp = bi + (long long)kp * kSlice;
    const int k = A.col[qkp];
    const long long bk = A.off[k >> 6] + (k & 63);
    const int dpk = A.diagk[k];
    int p = 0, q = 0;
    T dsum = T(0);
    while (p <= kp - 1 && q <= dpk - 1) {
      const int cp = A.col[bi + (long long)p * kSlice], cq = A.col[bk + (long long)q * kSlice];
      if (cp < cq) p++;
      else if (cp > cq) q++;
      else {
        dsum = dsum + F[bi + (long long)p * kSlice] * F[bk + (long long)q * kSlice];
        p++;
        q++;
      }
    }
    const T s = F[qkp] - dsum;
    F[qkp] = s / F[bk + (long long)dpk * kSlice];
  }
  const long long qd = bi + (long long)dpi * kSlice;
  T s = F[qd];
  for (int kp = 0; kp < dpi; kp++) {
    const T v = F[bi + (long long)kp * kSlice];
    s = s - v * v;
  }
  if (s <= T(0)) {
    *flag = 2;
    return;
  }
  F[qd] = tsqrt(s);
}

// natural ILU(0) of row i, IKJ variant (mod_uns_linsolver.f90:79-115)
template <class T>
__device__ void ilu0_factor_row(int i, const Mat &A, T *F, int *flag) {
  const long long bi = A.off[i >> 6] + (i & 63);
  const int dpi = A.diagk[i], leni = A.rowlen[i];
  for (int kk = 0; kk < dpi; kk++) {
    const long long q = bi + (long long)kk * kSlice;
    const int k = A.col[q];
    const long long bk = A.off[k >> 6] + (k & 63);
    const int dpk = A.diagk[k], lenk = A.rowlen[k];
    const T piv = F[bk + (long long)dpk * kSlice];
    if (tabs(piv) < (T)kTiny) {
      *flag = 2;
      return;
    }
    const T aik = F[q] / piv;
    F[q] = aik;
    for (int jj = dpk; jj < lenk; jj++) {
      const int j = A.col[bk + (long long)jj * kSlice];
      if (j <= k) continue;
      for (int m = 0; m < leni; m++) {
        const long long qm = bi + (long long)m * kSlice;
        const int cm = A.col[qm];
        if (cm == j) {
          F[qm] = F[qm] - aik * F[bk + (long long)jj * kSlice];
          break;
        }
        if (cm > j) break;
      }
    }
  }
}

// forward substitution of row i (unit lower for ILU(0), L of IC(0))
template <class T>
__device__ void lower_row(int i, const Mat &A, const T *F, const T *r, T *z, int ic0) {
  const long long bi = A.off[i >> 6] + (i & 63);
  const int dpi = A.diagk[i];
  T s = r[i];
  for (int k = 0; k < dpi; k++) {
    const long long q = bi + (long long)k * kSlice;
    s = s - F[q] * z[A.col[q]];
  }
  z[i] = ic0 ? s / F[bi + (long long)dpi * kSlice] : s;
}

// backward substitution of row i (U of ILU(0), L^T of IC(0) via tpos)
template <class T>
__device__ void upper_row(int i, const Mat &A, const long long *tpos, const T *F, T *z, int ic0) {
  const long long bi = A.off[i >> 6] + (i & 63);
  const int dpi = A.diagk[i], len = A.rowlen[i];
  T s = z[i];
  if (ic0) {
    for (int k = len - 1; k > dpi; k--) {
      const long long q = bi + (long long)k * kSlice;
      s = s - F[tpos[q]] * z[A.col[q]];
    }
  } else {
    for (int k = dpi + 1; k < len; k++) {
      const long long q = bi + (long long)k * kSlice;
      s = s - F[q] * z[A.col[q]];
    }
  }
  z[i] = s / F[bi + (long long)dpi * kSlice];
}

// one wide level: a thread per row
template <class T>
__global__ void k_factor_level(const int *rows, int nr, Mat A, T *F, int ic0, int *flag) {
  int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= nr) return;
  if (ic0) ic0_factor_row(rows[t], A, F, flag);
  else ilu0_factor_row(rows[t], A, F, flag);
}

template <class T>
__global__ void k_lower_level(const int *rows, int nr, Mat A, const T *F, const T *r, T *z, int ic0, const int *done) {
  if (*done) return;
  int t = blockIdx.x * kBlock + threadIdx.x;
  if (t < nr) lower_row(rows[t], A, F, r, z, ic0);
}

template <class T>
__global__ void k_upper_level(const int *rows, int nr, Mat A, const long long *tpos, const T *F, T *z, int ic0,
                              const int *done) {
  if (*done) return;
  int t = blockIdx.x * kBlock + threadIdx.x;
  if (t < nr) upper_row(rows[t], A, tpos, F, z, ic0);
}

// a run of narrow levels [l0, l1) in one workgroup: rows of a level are
// spread over the threads, a barrier orders the levels (the workgroup shares
// one L1, so rows of the next level see the values just stored).  Every
// thread reaches every barrier: the level loop is uniform, only the row loop
// inside a level depends on the thread.
template <class T>
__global__ __launch_bounds__(kSegThreads) void k_factor_seg(const int *lev_ptr, int l0, int l1, const int *rows,
                                                             Mat A, T *F, int ic0, int *flag) {
  for (int l = l0; l < l1; l++) {
    const int a = lev_ptr[l], b = lev_ptr[l + 1];
    for (int t = a + (int)threadIdx.x; t < b; t += kSegThreads) {
      if (ic0) ic0_factor_row(rows[t], A, F, flag);
      else ilu0_factor_row(rows[t], A, F, flag);
    }
    __syncthreads();
  }
}

template <class T>
__global__ __launch_bounds__(kSegThreads) void k_lower_seg(const int *lev_ptr, int l0, int l1, const int *rows, Mat A,
                                                            const T *F, const T *r, T *z, int ic0, const int *done) {
  if (*done) return;
  for (int l = l0; l < l1; l++) {
    const int a = lev_ptr[l], b = lev_ptr[l + 1];
    for (int t = a + (int)threadIdx.x; t < b; t += kSegThreads) lower_row(rows[t], A, F, r, z, ic0);
    __syncthreads();
  }
}

template <class T>
__global__ __launch_bounds__(kSegThreads) void k_upper_seg(const int *lev_ptr, int l0, int l1, const int *rows, Mat A,
                                                            const long long *tpos, const T *F, T *z, int ic0,
                                                            const int *done) {
  if (*done) return;
  for (int l = l0; l < l1; l++) {
    const int a = lev_ptr[l], b = lev_ptr[l + 1];
    for (int t = a + (int)threadIdx.x; t < b; t += kSegThreads) upper_row(rows[t], A, tpos, F, z, ic0);
    __syncthreads();
  }
}

template <class T>
__global__ void k_dilu_first(int r0, int r1, const T *dg, const T *r, T *z, const int *done) {
  if (*done) return;
  int i = r0 + blockIdx.x * kBlock + threadIdx.x;
  if (i < r1) z[i] = r[i] / dg[i];
}

// forward sweep of one colour [r0, r1) over nbr row blocks; the first
// cw.nwg blocks carry the coarse solve when a coarse space rides along (they
// start first and overlap the sweep)
template <class T, int C16 = -1>
__global__ void k_dilu_fwd(int r0, int r1, int nbr, Mat A, const T *val, const T *dg, const T *r, T *z, int last,
                           const int *done, T *part, CoWork<T> cw) {
  if (*done) return;
  if ((int)blockIdx.x < cw.nwg) {
    coarse_ride<T>(cw, blockIdx.x);
    return;
  }
  const int rb = blockIdx.x - cw.nwg;
  int i = r0 + xcd_block(rb, nbr) * kBlock + threadIdx.x;
  T v[1] = {T(0)};
  if (i < r1) {
    const long long bi = A.off[i >> 6] + (i & 63);
    const int dpi = A.diagk[i];
    const T ri = r[i];   // kept: a load after the store to z would wait for it
    T s = ri;
    for (int k = 0; k < dpi; k++) {
      const long long q = bi + (long long)k * kSlice;
      const int c = sizeof(T) == 4 ? col_at<C16>(A, q, i, i >> 6) : A.col[q];
      s = s - val[q] * z[c];
    }
    const T zi = s / dg[i];
    z[i] = zi;
    if (last) v[0] = ri * zi;
  }
  if (part) block_partials<OpSum, 1>(v, part, rb);
}

template <class T, int C16 = -1>
__global__ void k_dilu_bwd(int r0, int r1, Mat A, const T *val, const T *dg, const T *r, T *z, const int *done,
                           T *part) {
  if (*done) return;
  int i = r0 + xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  T v[1] = {T(0)};
  if (i < r1) {
    const long long bi = A.off[i >> 6] + (i & 63);
    const int dpi = A.diagk[i], len = A.rowlen[i];
    const T ri = r[i];   // loaded before the store to z
    T s = T(0);
    for (int k = dpi + 1; k < len; k++) {
      const long long q = bi + (long long)k * kSlice;
      const int c = sizeof(T) == 4 ? col_at<C16>(A, q, i, i >> 6) : A.col[q];
      s = s + val[q] * z[c];
    }
    const T zi = z[i] - s / dg[i];
    z[i] = zi;
    v[0] = ri * zi;
  }
  if (part) block_partials<OpSum, 1>(v, part);
}

// --------------------------------------------------- mixed precision glue
__global__ void k_mp_scale_down(int n, const double *r, const KS *st, float *r32, float *x32, KS *s32,
                                double inner_tol, int inner_max) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  const double inv = 1.0 / sqrt(st->rr);
  if (i < n) {
    r32[i] = (float)(r[i] * inv);
    x32[i] = 0.0f;
  }
  if (i == 0) {
    s32->it = 0;
    s32->ierr = 0;
    s32->done = 0;
    s32->bnrm = 1.0;
    s32->tol = inner_tol;
    s32->maxit = inner_max;
  }
}

__global__ void k_mp_correct(int n, const float *x32, const KS *st, double *x) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] = x[i] + sqrt(st->rr) * (double)x32[i];
}

// streaming copy (calibration of traffic counters and achievable bandwidth)
__global__ void k_copy4(long long n4, const float4 *a, float4 *b) {
  const long long base = (long long)blockIdx.x * kBlock * 4 + threadIdx.x;
  float4 v[4];
#pragma unroll
  for (int u = 0; u < 4; u++)
    if (base + u * kBlock < n4) v[u] = a[base + u * kBlock];
#pragma unroll
  for (int u = 0; u < 4; u++)
    if (base + u * kBlock < n4) b[base + u * kBlock] = v[u];
}

// streaming read of a zero-filled buffer: evicts the last-level cache before
// a cold-cache timing (reads only, so no dirty lines drain during the timed
// kernel); the store never happens for zeros but keeps the loads alive
__global__ void k_touch(long long n4, const float4 *a, float *sink) {
  float acc = 0.0f;
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n4; i += (long long)gridDim.x * kBlock) {
    const float4 v = a[i];
    acc += v.x + v.y + v.z + v.w;
  }
  if (acc == 1.0f) sink[0] = acc;
}

// copy with non-temporal stores (the destination bypasses the caches)
typedef float f4v __attribute__((ext_vector_type(4)));
__global__ void k_copy4_nt(long long n4, const float4 *a, float4 *b) {
  const long long base = (long long)blockIdx.x * kBlock * 4 + threadIdx.x;
  f4v v[4];
#pragma unroll
  for (int u = 0; u < 4; u++)
    if (base + u * kBlock < n4) v[u] = *(const f4v *)&a[base + u * kBlock];
#pragma unroll
  for (int u = 0; u < 4; u++)
    if (base + u * kBlock < n4) __builtin_nontemporal_store(v[u], (f4v *)&b[base + u * kBlock]);
}

}  // namespace

// Average duration of a streaming kernel over `bytes`: kind 0 = k_copy4,
// 1 = copy with non-temporal stores, 2 = read only (k_touch, 8192 blocks).
double time_stream(long long bytes, int kind, int reps) {
  const long long n4 = bytes / 16;
  DBuf<float4> a(n4), b(kind == 2 ? 1 : n4);
  DBuf<float> sink(1);
  MS_HIP(hipMemsetAsync(a.p, 0, n4 * 16, stream()));
  const int grid = (int)((n4 + 4 * kBlock - 1) / (4 * kBlock));
  auto go = [&]() {
    if (kind == 0) hipLaunchKernelGGL(k_copy4, dim3(grid), dim3(kBlock), 0, stream(), n4, a.p, b.p);
    else if (kind == 1) hipLaunchKernelGGL(k_copy4_nt, dim3(grid), dim3(kBlock), 0, stream(), n4, a.p, b.p);
    else hipLaunchKernelGGL(k_touch, dim3(8192), dim3(kBlock), 0, stream(), n4, a.p, sink.p);
  };
  for (int i = 0; i < 2; i++) go();
  hipEvent_t e0, e1;
  MS_HIP(hipEventCreate(&e0));
  MS_HIP(hipEventCreate(&e1));
  MS_HIP(hipEventRecord(e0, stream()));
  for (int i = 0; i < reps; i++) go();
  MS_HIP(hipEventRecord(e1, stream()));
  MS_HIP(hipEventSynchronize(e1));
  float ms = 0.0f;
  MS_HIP(hipEventElapsedTime(&ms, e0, e1));
  MS_HIP(hipEventDestroy(e0));
  MS_HIP(hipEventDestroy(e1));
  MS_CHECK_LAUNCH();
  return (double)ms / reps;
}

double time_copy(long long bytes, int reps) {
  const long long n4 = bytes / 16;
  DBuf<float4> a(n4), b(n4);
  MS_HIP(hipMemsetAsync(a.p, 0, n4 * 16, stream()));
  const int grid = (int)((n4 + 4 * kBlock - 1) / (4 * kBlock));
  for (int i = 0; i < 2; i++) hipLaunchKernelGGL(k_copy4, dim3(grid), dim3(kBlock), 0, stream(), n4, a.p, b.p);
  hipEvent_t e0, e1;
  MS_HIP(hipEventCreate(&e0));
  MS_HIP(hipEventCreate(&e1));
  MS_HIP(hipEventRecord(e0, stream()));
  for (int i = 0; i < reps; i++) hipLaunchKernelGGL(k_copy4, dim3(grid), dim3(kBlock), 0, stream(), n4, a.p, b.p);
  MS_HIP(hipEventRecord(e1, stream()));
  MS_HIP(hipEventSynchronize(e1));
  float ms = 0.0f;
  MS_HIP(hipEventElapsedTime(&ms, e0, e1));
  MS_HIP(hipEventDestroy(e0));
  MS_HIP(hipEventDestroy(e1));
  MS_CHECK_LAUNCH();
  return (double)ms / reps;
}

// ============================================================ host side

template <class Op, int NV, class T, class Fin>
static void finish(int nparts, const T *part, Fin fin) {
  hipLaunchKernelGGL((k_finish<Op, NV, T, Fin>), dim3(1), dim3(kFinThreads), 0, stream(), nparts, part, fin);
}

Krylov::Krylov(const SellPattern *pat_) : pat(pat_), n(pat_->n) {
  nblk_ = nblocks(n);
  for (DBuf<double> *b : {&r_, &r0_, &p_, &q_, &z_, &s_, &t_, &ph_, &sh_, &dg_}) b->alloc(n);
  part_.alloc(4 * ((size_t)std::max(nblk_, pat_->coarse.nchunks) + 64) * 2);
  prz_.alloc(2 * ((size_t)nblk_ + 64) + kMaxAgg);
  st_.alloc(1);
  flag_.alloc(1);
  std::lock_guard<std::recursive_mutex> lk(device_alloc_mutex());
  MS_HIP(hipHostMalloc(&hpin_, 4096, hipHostMallocDefault));
  for (auto &e : state_ev_) MS_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
}

Krylov::~Krylov() {
  std::lock_guard<std::recursive_mutex> lk(device_alloc_mutex());
  if (chunk_exec_) (void)hipGraphExecDestroy(chunk_exec_);
  for (auto &e : state_ev_)
    if (e) (void)hipEventDestroy(e);
  if (hpin_) (void)hipHostFree(hpin_);
}

void Krylov::ensure32() {
  if (have32_) return;
  for (DBuf<float> *v : {&r32_, &r032_, &p32_, &q32_, &z32_, &s32_, &t32_, &ph32_, &sh32_, &x32_, &dg32_}) v->alloc(n);
  part32_.alloc(4 * ((size_t)std::max(nblk_, pat->coarse.nchunks) + 64) * 2);
  prz32_.alloc(2 * ((size_t)nblk_ + 64) + kMaxAgg);
  st32_.alloc(1);
  have32_ = true;
}

template <class T>
KS Krylov::read_state(KS *d) {
  MS_HIP(hipMemcpyAsync(hpin_, d, sizeof(KS), hipMemcpyDeviceToHost, stream()));
  MS_HIP(hipStreamSynchronize(stream()));
  KS h;
  std::memcpy(&h, hpin_, sizeof(KS));
  return h;
}

void Krylov::spmv(const double *val, const double *x, double *y) {
  hipLaunchKernelGGL(k_spmv<double>, dim3(nblk_), dim3(kBlock), 0, stream(), n, matof(pat), val, x, y);
  MS_CHECK_LAUNCH();
}

void Krylov::spmv32(const double *val64, const float *v32, float *p, float *q, KS *st, float *part,
                    hipEvent_t e0, hipEvent_t e1, const PUp *pu) {
  (void)val64;
  const dim3 G(nblk_), B(kBlock);
  hipStream_t S = stream();
  if (matfree_on()) {
    const MatFree &M = *mf_;
    const Mat AM = matof(M.A);
    const dim3 G0(nblocks(M.n0));
    const PUp u = pu ? *pu : PUp{};
    // stage 1 then stage 2; the start event on the first launch, the stop
    // event on the second
    auto launch = [&](auto k1, auto k2) {
      if (e0) hipExtLaunchKernelGGL(k1, G0, B, 0, S, e0, nullptr, 0, M.n0, AM, M.w, p, M.hx, st, u);
      else hipLaunchKernelGGL(k1, G0, B, 0, S, M.n0, AM, M.w, p, M.hx, st, u);
      if (e1) hipExtLaunchKernelGGL(k2, G, B, 0, S, nullptr, e1, 0, M.n0, M.n1, AM, M.w, p, M.hx, q, st, part, u);
      else hipLaunchKernelGGL(k2, G, B, 0, S, M.n0, M.n1, AM, M.w, p, M.hx, q, st, part, u);
    };
    const bool c16 = M.A->has16;
    if (pu) {
      if (c16) launch(k_mf_stage1<1, true>, k_mf_stage2<1, true>);
      else launch(k_mf_stage1<0, true>, k_mf_stage2<0, true>);
    } else {
      if (c16) launch(k_mf_stage1<1, false>, k_mf_stage2<1, false>);
      else launch(k_mf_stage1<0, false>, k_mf_stage2<0, false>);
    }
    return;
  }
  const Mat A = matof(pat);
  if (e0 || e1) {
    if (pat->has16)
      hipExtLaunchKernelGGL((k_pcg_spmv<float, 1>), G, B, 0, S, e0, e1, 0, n, A, v32, p, q, st, part);
    else
      hipExtLaunchKernelGGL((k_pcg_spmv<float, 0>), G, B, 0, S, e0, e1, 0, n, A, v32, p, q, st, part);
  } else if (pat->has16) {
    hipLaunchKernelGGL((k_pcg_spmv<float, 1>), G, B, 0, S, n, A, v32, p, q, st, part);
  } else {
    hipLaunchKernelGGL((k_pcg_spmv<float, 0>), G, B, 0, S, n, A, v32, p, q, st, part);
  }
}

double Krylov::time_spmv(const double *val, bool fp32, int reps, bool cold) {
  Mat A = matof(pat);
  const dim3 G(nblk_), B(kBlock);
  hipStream_t S = stream();
  KS h{};
  MS_HIP(hipMemcpyAsync(st_.p, &h, sizeof(h), hipMemcpyHostToDevice, S));
  if (fp32) {
    last_ = SetupKey{};
    ensure32();
    if (val32_.n < (size_t)pat->nstore) val32_.alloc(pat->nstore);
    hipLaunchKernelGGL(k_mat32_diff, G, B, 0, S, n, A, val, val32_.p);
    MS_HIP(hipMemsetAsync(p32_.p, 0, n * sizeof(float), S));
  } else {
    MS_HIP(hipMemsetAsync(p_.p, 0, n * sizeof(double), S));
  }
  auto launch = [&]() {
    if (fp32)
      spmv32(val, val32_.p, p32_.p, q32_.p, st_.p, part32_.p, nullptr, nullptr);
    else
      hipLaunchKernelGGL(k_pcg_spmv<double>, G, B, 0, S, n, A, val, p_.p, q_.p, st_.p, part_.p);
  };
  for (int i = 0; i < 3; i++) launch();
  if (cold) {
    // each launch timed alone after a 512 MiB read (twice the 256 MB
    // last-level cache), so its matrix and vectors come from HBM
    const long long n4 = (512LL << 20) / 16;
    DBuf<float4> flush(n4);
    DBuf<float> sink(1);
    MS_HIP(hipMemsetAsync(flush.p, 0, n4 * 16, S));
    std::vector<hipEvent_t> ev(2 * (size_t)reps);
    for (auto &e : ev) MS_HIP(hipEventCreate(&e));
    for (int i = 0; i < reps; i++) {
      hipLaunchKernelGGL(k_touch, dim3(8192), B, 0, S, n4, flush.p, sink.p);
      MS_HIP(hipEventRecord(ev[2 * i], S));
      launch();
      MS_HIP(hipEventRecord(ev[2 * i + 1], S));
    }
    MS_HIP(hipEventSynchronize(ev.back()));
    double tot = 0.0;
    for (int i = 0; i < reps; i++) {
      float ms = 0.0f;
      MS_HIP(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
      tot += ms;
    }
    for (auto &e : ev) MS_HIP(hipEventDestroy(e));
    MS_CHECK_LAUNCH();
    return tot / reps;
  }
  hipEvent_t e0, e1;
  MS_HIP(hipEventCreate(&e0));
  MS_HIP(hipEventCreate(&e1));
  MS_HIP(hipEventRecord(e0, S));
  for (int i = 0; i < reps; i++) launch();
  MS_HIP(hipEventRecord(e1, S));
  MS_HIP(hipEventSynchronize(e1));
  float ms = 0.0f;
  MS_HIP(hipEventElapsedTime(&ms, e0, e1));
  MS_HIP(hipEventDestroy(e0));
  MS_HIP(hipEventDestroy(e1));
  MS_CHECK_LAUNCH();
  return (double)ms / reps;
}

int Krylov::setup_pc64(PcKind pc, const double *val) {
  Mat A = matof(pat);
  if ((pc == PC_DILU || pc == PC_NATURAL_IC0) && !pat->symmetric)
    fail(MS_EINVAL, "IC0 / DILU need a structurally symmetric pattern");
  if (pc == PC_JACOBI) {
    hipLaunchKernelGGL(k_extract_diag<double>, dim3(nblk_), dim3(kBlock), 0, stream(), n, A, val, dg_.p);
    MS_CHECK_LAUNCH();
    return 0;
  }
  if (pc == PC_DILU) {
    for (size_t c = 0; c + 1 < pat->color_ptr.size(); c++) {
      int r0 = pat->color_ptr[c], r1 = pat->color_ptr[c + 1];
      if (r1 > r0)
        hipLaunchKernelGGL(k_dilu_setup<double>, dim3(nblocks(r1 - r0)), dim3(kBlock), 0, stream(), r0, r1, A,
                           pat->d_tpos.p, val, dg_.p);
    }
    MS_CHECK_LAUNCH();
    return 0;
  }
  if (!pat->has_levels) fail(MS_EINVAL, "natural-order factorization needs level schedules");
  if (fac_.n < (size_t)pat->nstore) fac_.alloc(pat->nstore);
  MS_HIP(hipMemcpyAsync(fac_.p, val, pat->nstore * sizeof(double), hipMemcpyDeviceToDevice, stream()));
  MS_HIP(hipMemsetAsync(flag_.p, 0, sizeof(int), stream()));
  const int ic0 = pc == PC_NATURAL_IC0;
  for (const auto &g : pat->seg_f) {
    if (g.narrow) {
      hipLaunchKernelGGL(k_factor_seg<double>, dim3(1), dim3(kSegThreads), 0, stream(), pat->d_lev_ptr_f.p, g.l0,
                         g.l1, pat->d_lev_rows_f.p, A, fac_.p, ic0, flag_.p);
      continue;
    }
    const int a = pat->lev_ptr_f[g.l0], b = pat->lev_ptr_f[g.l1];
    hipLaunchKernelGGL(k_factor_level<double>, dim3(nblocks(b - a)), dim3(kBlock), 0, stream(),
                       pat->d_lev_rows_f.p + a, b - a, A, fac_.p, ic0, flag_.p);
  }
  MS_CHECK_LAUNCH();
  int h = 0;
  MS_HIP(hipMemcpyAsync(hpin_, flag_.p, sizeof(int), hipMemcpyDeviceToHost, stream()));
  MS_HIP(hipStreamSynchronize(stream()));
  std::memcpy(&h, hpin_, sizeof(int));
  return h;
}

void Krylov::setup_pc32(PcKind pc, const double *val) {
  Mat A = matof(pat);
  ensure32();
  if (pc == PC_JACOBI || pc == PC_DILU) {
    // the diagonal slot of val32 holds row sums: take d from the FP64 matrix
    setup_pc64(pc, val);
    hipLaunchKernelGGL(k_to_float, dim3(nblk_), dim3(kBlock), 0, stream(), (long long)n, dg_.p, dg32_.p);
  } else {
    int e = setup_pc64(pc, val);
    if (e) fail(MS_EINVAL, "factorization failed in mixed-precision setup");
    if (fac32_.n < (size_t)pat->nstore) fac32_.alloc(pat->nstore);
    hipLaunchKernelGGL(k_to_float, dim3(nblocks(pat->nstore)), dim3(kBlock), 0, stream(), pat->nstore, fac_.p,
                       fac32_.p);
  }
  MS_CHECK_LAUNCH();
}

// the column format of a SELL pattern as a template argument of the FP32
// kernels (the FP64 kernels always read 32-bit columns)
template <class T>
static int c16_of(const SellPattern *pat) {
  return sizeof(T) == 4 && pat->has16 ? 1 : 0;
}

template <class T>
void Krylov::apply_pc(PcKind pc, const T *val, const T *fac, const T *dg, const T *in, T *out, const int *done) {
  Mat A = matof(pat);
  if (pc == PC_JACOBI) {
    hipLaunchKernelGGL(k_jacobi<T>, dim3(nblk_), dim3(kBlock), 0, stream(), n, in, dg, out, done);
  } else if (pc == PC_DILU) {
    const auto &cp = pat->color_ptr;
    const int nc = (int)cp.size() - 1;
    if (cp[1] > cp[0])
      hipLaunchKernelGGL(k_dilu_first<T>, dim3(nblocks(cp[1] - cp[0])), dim3(kBlock), 0, stream(), cp[0], cp[1], dg,
                         in, out, done);
    for (int c = 1; c < nc; c++)
      if (cp[c + 1] > cp[c]) {
        const int gb = nblocks(cp[c + 1] - cp[c]);
        if (c16_of<T>(pat))
          hipLaunchKernelGGL((k_dilu_fwd<T, 1>), dim3(gb), dim3(kBlock), 0, stream(), cp[c], cp[c + 1], gb, A, val, dg,
                             in, out, (int)(c == nc - 1), done, (T *)nullptr, CoWork<T>{});
        else
          hipLaunchKernelGGL((k_dilu_fwd<T, 0>), dim3(gb), dim3(kBlock), 0, stream(), cp[c], cp[c + 1], gb, A, val, dg,
                             in, out, (int)(c == nc - 1), done, (T *)nullptr, CoWork<T>{});
      }
    for (int c = nc - 2; c >= 0; c--)
      if (cp[c + 1] > cp[c])
        if (c16_of<T>(pat))
          hipLaunchKernelGGL((k_dilu_bwd<T, 1>), dim3(nblocks(cp[c + 1] - cp[c])), dim3(kBlock), 0, stream(), cp[c],
                             cp[c + 1], A, val, dg, in, out, done, (T *)nullptr);
        else
          hipLaunchKernelGGL((k_dilu_bwd<T, 0>), dim3(nblocks(cp[c + 1] - cp[c])), dim3(kBlock), 0, stream(), cp[c],
                             cp[c + 1], A, val, dg, in, out, done, (T *)nullptr);
  } else {
    const int ic0 = pc == PC_NATURAL_IC0;
    for (const auto &g : pat->seg_f) {
      if (g.narrow) {
        hipLaunchKernelGGL(k_lower_seg<T>, dim3(1), dim3(kSegThreads), 0, stream(), pat->d_lev_ptr_f.p, g.l0, g.l1,
                           pat->d_lev_rows_f.p, A, fac, in, out, ic0, done);
        continue;
      }
      const int a = pat->lev_ptr_f[g.l0], b = pat->lev_ptr_f[g.l1];
      hipLaunchKernelGGL(k_lower_level<T>, dim3(nblocks(b - a)), dim3(kBlock), 0, stream(), pat->d_lev_rows_f.p + a,
                         b - a, A, fac, in, out, ic0, done);
    }
    for (const auto &g : pat->seg_b) {
      if (g.narrow) {
        hipLaunchKernelGGL(k_upper_seg<T>, dim3(1), dim3(kSegThreads), 0, stream(), pat->d_lev_ptr_b.p, g.l0, g.l1,
                           pat->d_lev_rows_b.p, A, pat->d_tpos.p, fac, out, ic0, done);
        continue;
      }
      const int a = pat->lev_ptr_b[g.l0], b = pat->lev_ptr_b[g.l1];
      hipLaunchKernelGGL(k_upper_level<T>, dim3(nblocks(b - a)), dim3(kBlock), 0, stream(), pat->d_lev_rows_b.p + a,
                         b - a, A, pat->d_tpos.p, fac, out, ic0, done);
    }
  }
  MS_CHECK_LAUNCH();
}

// DILU sweeps of colours 1.. with fused r.z partials (one section per launch);
// returns the number of partials written.  With cw.nwg > 0 the coarse solve
// rides along the first forward launch and its c.nu partials follow the
// sweep sections, so the sum of all partials is r.z of the two-level
// preconditioner.
template <class T>
static int dilu_sweeps(const SellPattern *pat, Mat A, const T *val, const T *dg, const T *r, T *z, const int *done,
                       T *part, CoWork<T> cw = CoWork<T>{}) {
  const auto &cp = pat->color_ptr;
  const int ncol = (int)cp.size() - 1;
  int total = 0;
  for (int c = 1; c < ncol; c++) total += nblocks(cp[c + 1] - cp[c]);
  for (int c = ncol - 2; c >= 0; c--) total += nblocks(cp[c + 1] - cp[c]);
  cw.cnu_part = part + total;
  T *sec = part;
  bool first = true;
  for (int c = 1; c < ncol; c++)
    if (cp[c + 1] > cp[c]) {
      const int gb = nblocks(cp[c + 1] - cp[c]);
      CoWork<T> cwl = cw;
      if (!first) cwl.nwg = 0;
      if (c16_of<T>(pat))
        hipLaunchKernelGGL((k_dilu_fwd<T, 1>), dim3(gb + cwl.nwg), dim3(kBlock), 0, stream(), cp[c], cp[c + 1], gb, A,
                           val, dg, r, z, (int)(c == ncol - 1), done, sec, cwl);
      else
        hipLaunchKernelGGL((k_dilu_fwd<T, 0>), dim3(gb + cwl.nwg), dim3(kBlock), 0, stream(), cp[c], cp[c + 1], gb, A,
                           val, dg, r, z, (int)(c == ncol - 1), done, sec, cwl);
      first = false;
      sec += gb;
    }
  for (int c = ncol - 2; c >= 0; c--)
    if (cp[c + 1] > cp[c]) {
      const int gb = nblocks(cp[c + 1] - cp[c]);
      if (c16_of<T>(pat))
        hipLaunchKernelGGL((k_dilu_bwd<T, 1>), dim3(gb), dim3(kBlock), 0, stream(), cp[c], cp[c + 1], A, val, dg, r, z,
                           done, sec);
      else
        hipLaunchKernelGGL((k_dilu_bwd<T, 0>), dim3(gb), dim3(kBlock), 0, stream(), cp[c], cp[c + 1], A, val, dg, r, z,
                           done, sec);
      sec += gb;
    }
  return total + (first ? 0 : cw.nwg);
}

// ---- PCG (FP64, reference recurrences) -------------------------------------
static Rows rows_of(const SellPattern *pat, bool coarse) {
  if (!coarse) return Rows{nullptr, nullptr, pat->n};
  return Rows{pat->coarse.d_chunk_ptr.p, pat->coarse.d_chunk_agg.p, pat->n};
}
static CoDev codev_of(const SellPattern *pat, const CoarseMat &cm, bool coarse, KS *st) {
  if (!coarse) return CoDev{0, nullptr, nullptr, nullptr, nullptr, st};
  const Coarse &C = pat->coarse;
  return CoDev{C.k, C.d_agg_cptr.p, C.d_agg_cidx.p, cm.Einv.p, cm.nu.p, st};
}
template <class T>
static CoWork<T> cowork_of(const SellPattern *pat, const CoarseMat &cm, bool fused, const T *cpart) {
  CoWork<T> cw{};
  if (!fused) return cw;
  const Coarse &C = pat->coarse;
  cw.k = C.k;
  cw.nwg = (C.k + kBlock / 64 - 1) / (kBlock / 64);
  cw.agg_cptr = C.d_agg_cptr.p;
  cw.agg_cidx = C.d_agg_cidx.p;
  cw.Einv = cm.Einv.p;
  cw.nu = cm.nu.p;
  cw.cpart = cpart;
  return cw;
}

template <int NV, class T, class Fin>
static void finish_co(int nparts, const T *part, Fin fin, const CoDev &co) {
  hipLaunchKernelGGL((k_finish_co<NV, T, Fin>), dim3(1), dim3(kFinThreads), 0, stream(), nparts, part, fin, co);
}

static bool coarse_inv32_enabled();
void Krylov::coarse_setup(const double *val) {
  const Coarse &C = pat->coarse;
  CoarseMat &M = cm_;
  M.ensure(C.k);
  const long long kk = (long long)C.k * C.k;
  const int nE = (int)C.e_key.size();
  hipLaunchKernelGGL(k_fill64, dim3(nblocks(kk)), dim3(kBlock), 0, stream(), kk, M.E.p, 0.0);
  if (M.Eparts.n < (size_t)nE * kEParts) M.Eparts.alloc((size_t)nE * kEParts);
  hipLaunchKernelGGL(k_coarse_E_parts, dim3(nblocks((long long)nE * kEParts, kBlock / 64)), dim3(kBlock), 0,
                     stream(), nE, C.d_e_ptr.p, C.d_e_slot.p, val, M.Eparts.p);
  hipLaunchKernelGGL(k_coarse_E_sum, dim3(nblocks(nE)), dim3(kBlock), 0, stream(), nE, C.d_e_key.p, M.Eparts.p,
                     M.E.p);
  if (C.k <= kInvRegMax && coarse_inv32_enabled()) {
    // one workgroup, E held in registers, Gauss-Jordan in FP32
    hipLaunchKernelGGL(k_coarse_inv, dim3(1), dim3(kInvThreads), 0, stream(), C.k, M.E.p, M.Einv.p);
    MS_CHECK_LAUNCH();
    return;
  }
  // FP64 Gauss-Jordan over the grid, one launch per pivot (k <= 1024)
  double *rc[2][2] = {{M.rc.p, M.rc.p + C.k}, {M.rc.p + 2 * C.k, M.rc.p + 3 * C.k}};
  const dim3 g(nblocks(kk));
  hipLaunchKernelGGL(k_gj_init, g, dim3(kBlock), 0, stream(), C.k, M.E.p, M.W.p, rc[0][0], rc[0][1]);
  for (int j = 0; j < C.k; j++) {
    const int cur = j & 1, nxt = cur ^ 1;
    hipLaunchKernelGGL(k_gj_step, g, dim3(kBlock), 0, stream(), C.k, j, M.W.p, rc[cur][0], rc[cur][1], rc[nxt][0],
                       rc[nxt][1]);
  }
  hipLaunchKernelGGL(k_gj_out, g, dim3(kBlock), 0, stream(), kk, M.W.p, M.Einv.p);
  MS_CHECK_LAUNCH();
}

template <class T>
void Krylov::coarse_apply(const T *in, T *out, const int *done, KS *st) {
  const Coarse &C = pat->coarse;
  if (co_part_.n < (size_t)C.nchunks) co_part_.alloc(C.nchunks);
  const Rows RW = rows_of(pat, true);
  hipLaunchKernelGGL((k_co_restrict<T>), dim3(C.nchunks), dim3(kBlock), 0, stream(), RW, in, co_part_.p, done);
  hipLaunchKernelGGL(k_co_solve, dim3(1), dim3(kFinThreads), 0, stream(), co_part_.p, codev_of(pat, cm_, true, st),
                     done);
  hipLaunchKernelGGL((k_co_prolong<T>), dim3(C.nchunks), dim3(kBlock), 0, stream(), RW, cm_.nu.p, out, done);
}

SolveResult Krylov::pcg(const double *val, const double *b, double *x, double tol, int maxit, PcKind pc,
                        bool coarse, bool reuse) {
  SolveResult R;
  Mat A = matof(pat);
  const dim3 G(nblk_), B(kBlock);
  hipStream_t S = stream();
  KS *st = st_.p;
  double *part = part_.p;
  coarse = coarse && pat->coarse.k > 0;
  // DILU with a coarse space: the coarse solve rides along the forward sweep
  const bool fused = coarse && pc == PC_DILU && pat->color_ptr.size() > 2;
  KS h{};
  h.tol = tol;
  h.maxit = maxit;
  h.coarse = coarse && !fused;
  MS_HIP(hipMemcpyAsync(st, &h, sizeof(h), hipMemcpyHostToDevice, S));
  hipLaunchKernelGGL(k_sumsq<double>, G, B, 0, S, n, b, part);
  finish<OpSum, 1>(nblk_, part, FinBnorm{st});
  MS_CHECK_LAUNCH();
  KS hs = read_state<double>(st);
  if (hs.bnrm == 0.0) {
    MS_HIP(hipMemsetAsync(x, 0, n * sizeof(double), S));
    return R;
  }
  const SetupKey key{val, (int)pc, 0, (int)coarse};
  if (!(reuse && last_ == key)) {
    last_ = SetupKey{};
    int fe = setup_pc64(pc, val);
    if (fe) {
      R.ierr = fe + 10;
      return R;
    }
    if (coarse) coarse_setup(val);
    last_ = key;
  }
  hipLaunchKernelGGL(k_residual<double>, G, B, 0, S, n, A, val, b, x, r_.p, part);
  finish<OpSum, 1>(nblk_, part, FinResidual{st});
  MS_CHECK_LAUNCH();
  hs = read_state<double>(st);
  R.res = hs.res;
  if (hs.done) return R;
  const int jac = pc == PC_JACOBI;
  const int *done = &st->done;
  const Rows RW = rows_of(pat, coarse);
  const dim3 GR(coarse ? pat->coarse.nchunks : nblk_);
  const CoDev co = codev_of(pat, cm_, coarse, st);
  const double *nu = coarse ? cm_.nu.p : nullptr;
  const CoWork<double> cw = cowork_of<double>(pat, cm_, fused, part);
  double *prz = prz_.p;
  if (fused) {
    hipLaunchKernelGGL(k_rsum<double>, GR, B, 0, S, RW, r_.p, st, part);
    const auto &cp = pat->color_ptr;
    hipLaunchKernelGGL(k_dilu_first<double>, dim3(nblocks(cp[1] - cp[0])), B, 0, S, cp[0], cp[1], dg_.p, r_.p, z_.p,
                       done);
    const int np = dilu_sweeps<double>(pat, A, val, dg_.p, r_.p, z_.p, done, prz, cw);
    finish<OpSum, 1>(np, prz, FinRz0{st});
    MS_HIP(hipMemsetAsync(p_.p, 0, n * sizeof(double), S));
    hipLaunchKernelGGL(k_pcg_p<double>, GR, B, 0, S, RW, r_.p, z_.p, dg_.p, 0, p_.p, nu, st);   // beta = 0
  } else {
    if (coarse) {
      hipLaunchKernelGGL(k_rsum<double>, GR, B, 0, S, RW, r_.p, st, part);
      finish_co<3>((int)GR.x, part, FinNop{st}, co);
    }
    if (!jac) apply_pc<double>(pc, val, fac_.p, dg_.p, r_.p, z_.p, done);
    hipLaunchKernelGGL(k_pcg_start<double>, GR, B, 0, S, RW, r_.p, z_.p, dg_.p, jac, p_.p, nu, st, part);
    finish<OpSum, 1>((int)GR.x, part, FinRz0{st});
  }
  const int mode = jac ? 1 : (pc == PC_DILU ? 2 : 0);
  const int c0end = pat->color_ptr[1];
  const CoDev co_upd = fused ? CoDev{0, nullptr, nullptr, nullptr, nullptr, st} : co;
  const bool merged = mode == 2 && (!coarse || fused);
  int enq = 0, chunk = 8;
  while (true) {
    int m = std::min(chunk, maxit - enq);
    for (int it = 0; it < m; it++) {
      hipLaunchKernelGGL(k_pcg_spmv<double>, G, B, 0, S, n, A, val, p_.p, q_.p, st, part);
      finish<OpSum, 1>(nblk_, part, FinPap{st});
      hipLaunchKernelGGL((k_pcg_update<double, double>), GR, B, 0, S, RW, p_.p, q_.p, x, r_.p, z_.p, dg_.p, mode,
                         c0end, st, part);
      if (merged) {      // convergence check and r.z in one finisher after the sweeps
        const int np = dilu_sweeps<double>(pat, A, val, dg_.p, r_.p, z_.p, done, prz, cw);
        hipLaunchKernelGGL((k_finish2<double, FinUpdRz>), dim3(1), dim3(kFinThreads), 0, S, (int)GR.x, part, np, prz,
                           FinUpdRz{st});
      } else {
        finish_co<3>((int)GR.x, part, FinUpdate{st, jac}, co_upd);
      }
      if (mode == 2 && !merged) {
        int np = dilu_sweeps<double>(pat, A, val, dg_.p, r_.p, z_.p, done, prz, cw);
        finish<OpSum, 1>(np, prz, FinRz{st});
      } else if (mode == 0) {
        apply_pc<double>(pc, val, fac_.p, dg_.p, r_.p, z_.p, done);
        hipLaunchKernelGGL(k_dot<double>, G, B, 0, S, n, r_.p, z_.p, st, part);
        finish<OpSum, 1>(nblk_, part, FinRz{st});
      }
      hipLaunchKernelGGL(k_pcg_p<double>, GR, B, 0, S, RW, r_.p, z_.p, dg_.p, jac, p_.p, nu, st);
    }
    MS_CHECK_LAUNCH();
    enq += m;
    hs = read_state<double>(st);
    if (hs.done || enq >= maxit) break;
    chunk = std::min(chunk * 2, 128);
  }
  R.it = hs.it;
  R.res = hs.res;
  R.ierr = hs.ierr;
  if (!hs.done) {        // DO-loop exhaustion: it = maxit + 1, ierr = 1
    R.it = maxit + 1;
    R.ierr = 1;
  }
  return R;
}

// ---- BiCGSTAB -----------------------------------------------------------------
SolveResult Krylov::bicgstab(const double *val, const double *b, double *x, double tol, int maxit, PcKind pc,
                             bool coarse) {
  SolveResult R;
  Mat A = matof(pat);
  const dim3 G(nblk_), B(kBlock);
  hipStream_t S = stream();
  KS *st = st_.p;
  double *part = part_.p;
  KS h{};
  h.tol = tol;
  h.maxit = maxit;
  MS_HIP(hipMemcpyAsync(st, &h, sizeof(h), hipMemcpyHostToDevice, S));
  hipLaunchKernelGGL(k_sumsq<double>, G, B, 0, S, n, b, part);
  finish<OpSum, 1>(nblk_, part, FinBnorm{st});
  MS_CHECK_LAUNCH();
  KS hs = read_state<double>(st);
  if (hs.bnrm == 0.0) {
    MS_HIP(hipMemsetAsync(x, 0, n * sizeof(double), S));
    return R;
  }
  last_ = SetupKey{};
  int fe = setup_pc64(pc, val);
  if (fe) {
    R.ierr = fe + 10;
    return R;
  }
  coarse = coarse && pat->coarse.k > 0;
  if (coarse) coarse_setup(val);
  hipLaunchKernelGGL(k_residual<double>, G, B, 0, S, n, A, val, b, x, r_.p, part);
  finish<OpSum, 1>(nblk_, part, FinResidual{st});
  hipLaunchKernelGGL(k_bi_init<double>, G, B, 0, S, n, r_.p, r0_.p, p_.p, q_.p, st);
  MS_CHECK_LAUNCH();
  hs = read_state<double>(st);
  R.res = hs.res;
  if (hs.done) return R;
  double *v = q_.p;
  const int *done = &st->done;
  int enq = 0, chunk = 2;
  while (true) {
    int m = std::min(chunk, maxit - enq);
    for (int k = 0; k < m; k++) {
      hipLaunchKernelGGL(k_dot<double>, G, B, 0, S, n, r_.p, r0_.p, st, part);
      finish<OpSum, 1>(nblk_, part, FinBiRho{st});
      hipLaunchKernelGGL(k_bi_p<double>, G, B, 0, S, n, r_.p, v, p_.p, st);
      apply_pc<double>(pc, val, fac_.p, dg_.p, p_.p, ph_.p, done);
      if (coarse) coarse_apply<double>(p_.p, ph_.p, done, st);
      hipLaunchKernelGGL(k_bi_spmv_dot<double>, G, B, 0, S, n, A, val, ph_.p, r0_.p, v, st, part);
      finish<OpSum, 1>(nblk_, part, FinBiAlpha{st});
      hipLaunchKernelGGL(k_bi_s<double>, G, B, 0, S, n, r_.p, v, s_.p, st, part);
      finish<OpSum, 1>(nblk_, part, FinBiS{st});
      hipLaunchKernelGGL(k_bi_early<double>, G, B, 0, S, n, ph_.p, x, st);
      hipLaunchKernelGGL(k_done2_to_1, dim3(1), dim3(64), 0, S, st);
      apply_pc<double>(pc, val, fac_.p, dg_.p, s_.p, sh_.p, done);
      if (coarse) coarse_apply<double>(s_.p, sh_.p, done, st);
      hipLaunchKernelGGL(k_bi_spmv_omega<double>, G, B, 0, S, n, A, val, sh_.p, s_.p, t_.p, st, part);
      finish<OpSum, 2>(nblk_, part, FinBiOmega{st});
      hipLaunchKernelGGL(k_bi_xr<double>, G, B, 0, S, n, ph_.p, sh_.p, s_.p, t_.p, x, r_.p, st, part);
      finish<OpSum, 1>(nblk_, part, FinBiXR{st});
    }
    MS_CHECK_LAUNCH();
    enq += m;
    hs = read_state<double>(st);
    if (hs.done || enq >= maxit) break;
    chunk = std::min(chunk * 2, 32);
  }
  R.it = hs.it;
  R.res = hs.res;
  R.ierr = hs.ierr;
  if (!hs.done) {        // exhaustion: ierr = 1, it = maxit (mod_uns_linsolver_mpi.f90:224-227)
    R.it = maxit;
    R.ierr = 1;
  }
  return R;
}

// FP32 BiCGSTAB on the scaled residual.  An inner solve whose recursive
// residual has not halved over `window` iterations ends early (the caller's
// refinement restarts it from the FP64 residual).
SolveResult Krylov::bicg32(PcKind pc, int maxit, int window, bool coarse) {
  SolveResult R;
  Mat A = matof(pat);
  const dim3 G(nblk_), B(kBlock);
  hipStream_t S = stream();
  KS *st = st32_.p;
  float *part = part32_.p;
  float *v = q32_.p;
  const int *done = &st->done;
  hipLaunchKernelGGL(k_bi_init<float>, G, B, 0, S, n, r32_.p, r032_.p, p32_.p, v, st);
  int enq = 0, chunk = 4;
  KS hs{};
  double best = HUGE_VAL;
  int it_best = 0;
  while (true) {
    int m = std::min(chunk, maxit - enq);
    for (int k = 0; k < m; k++) {
      hipLaunchKernelGGL(k_dot<float>, G, B, 0, S, n, r32_.p, r032_.p, st, part);
      finish<OpSum, 1>(nblk_, part, FinBiRho{st});
      hipLaunchKernelGGL(k_bi_p<float>, G, B, 0, S, n, r32_.p, v, p32_.p, st);
      apply_pc<float>(pc, val32_.p, fac32_.p, dg32_.p, p32_.p, ph32_.p, done);
      if (coarse) coarse_apply<float>(p32_.p, ph32_.p, done, st);
      if (pat->has16)
        hipLaunchKernelGGL((k_bi_spmv_dot<float, 1>), G, B, 0, S, n, A, val32_.p, ph32_.p, r032_.p, v, st, part);
      else
        hipLaunchKernelGGL((k_bi_spmv_dot<float, 0>), G, B, 0, S, n, A, val32_.p, ph32_.p, r032_.p, v, st, part);
      finish<OpSum, 1>(nblk_, part, FinBiAlpha{st});
      hipLaunchKernelGGL(k_bi_s<float>, G, B, 0, S, n, r32_.p, v, s32_.p, st, part);
      finish<OpSum, 1>(nblk_, part, FinBiS{st});
      hipLaunchKernelGGL(k_bi_early<float>, G, B, 0, S, n, ph32_.p, x32_.p, st);
      hipLaunchKernelGGL(k_done2_to_1, dim3(1), dim3(64), 0, S, st);
      apply_pc<float>(pc, val32_.p, fac32_.p, dg32_.p, s32_.p, sh32_.p, done);
      if (coarse) coarse_apply<float>(s32_.p, sh32_.p, done, st);
      if (pat->has16)
        hipLaunchKernelGGL((k_bi_spmv_omega<float, 1>), G, B, 0, S, n, A, val32_.p, sh32_.p, s32_.p, t32_.p, st,
                           part);
      else
        hipLaunchKernelGGL((k_bi_spmv_omega<float, 0>), G, B, 0, S, n, A, val32_.p, sh32_.p, s32_.p, t32_.p, st,
                           part);
      finish<OpSum, 2>(nblk_, part, FinBiOmega{st});
      hipLaunchKernelGGL(k_bi_xr<float>, G, B, 0, S, n, ph32_.p, sh32_.p, s32_.p, t32_.p, x32_.p, r32_.p, st, part);
      finish<OpSum, 1>(nblk_, part, FinBiXR{st});
    }
    MS_CHECK_LAUNCH();
    enq += m;
    hs = read_state<float>(st);
    if (hs.done || enq >= maxit) break;
    if (hs.res < 0.5 * best) {
      best = hs.res;
      it_best = enq;
    } else if (enq - it_best >= window) {
      R.stalled = 1;
      break;
    }
    chunk = std::min(chunk * 2, 32);
  }
  R.it = hs.it;
  R.ierr = hs.ierr;
  R.res = hs.res;
  return R;
}

// ---- in-loop SpMV timing ---------------------------------------------------------
// With the probe on, the mixed PCG launches the last FP32 SpMV of every
// 16-iteration chunk through hipExtLaunchKernelGGL with a start / stop event
// pair, which take the kernel's own dispatch timestamps (that iteration is
// enqueued behind the captured graph of the other 15, so it runs back to
// back with them); the pairs are read after each chunk (the host waits for
// the chunk's state anyway).
namespace {
struct SpmvProbe {
  bool on = false;
  std::vector<hipEvent_t> ev;   // a ring of start / stop pairs
  int next = 0;
  long long launches = 0, rows = 0;
  double ms = 0.0;
};
constexpr int kProbePairs = 32;
SpmvProbe g_probe;
void probe_harvest(int pair) {
  float t = 0.f;
  MS_HIP(hipEventSynchronize(g_probe.ev[2 * pair + 1]));
  MS_HIP(hipEventElapsedTime(&t, g_probe.ev[2 * pair], g_probe.ev[2 * pair + 1]));
  g_probe.ms += t;
  g_probe.launches++;
}
}  // namespace

void set_spmv_probe(bool on) {
  if (on && g_probe.ev.empty()) {
    g_probe.ev.resize(2 * kProbePairs);
    for (auto &e : g_probe.ev) MS_HIP(hipEventCreate(&e));
  }
  g_probe.on = on;
  g_probe.next = 0;
  g_probe.launches = g_probe.rows = 0;
  g_probe.ms = 0.0;
}
void read_spmv_probe(long long *launches, double *ms, long long *rows) {
  *launches = g_probe.launches;
  *ms = g_probe.ms;
  *rows = g_probe.rows;
}

// ---- mixed precision ------------------------------------------------------------
static std::atomic<int> g_schur_mf{-1};   // -1: not yet read from the environment
bool schur_matfree_enabled() {
  int v = g_schur_mf.load();
  if (v < 0) {
    const char *e = std::getenv("MIXSOLVER_SCHUR_MF");
    v = (e && e[0] == '0') ? 0 : 1;
    g_schur_mf.store(v);
  }
  return v != 0;
}
void set_schur_matfree(bool on) { g_schur_mf.store(on ? 1 : 0); }
// E^-1 of a coarse space of at most 256 aggregates by one workgroup in FP32
// (MIXSOLVER_COARSE_INV32=1) instead of k launches in FP64; off by default:
// the one workgroup takes longer than the k launches over the grid (17.17
// against 17.64 PISO steps/s at 2M cells, 256 aggregates)
static bool coarse_inv32_enabled() {
  static const bool on = [] {
    const char *e = std::getenv("MIXSOLVER_COARSE_INV32");
    return e && e[0] == '1';
  }();
  return on;
}
static std::atomic<int> g_p_fused{-1};
bool p_fused_enabled() {
  int v = g_p_fused.load();
  if (v < 0) {
    const char *e = std::getenv("MIXSOLVER_P_FUSED");
    v = (e && e[0] == '1') ? 1 : 0;
    g_p_fused.store(v);
  }
  return v != 0;
}
void set_p_fused(bool on) { g_p_fused.store(on ? 1 : 0); }

static std::atomic<int> g_stall_window{1024};
void set_stall_window(int iters) { g_stall_window.store(iters > 0 ? iters : 1024); }
int stall_window() { return g_stall_window.load(); }

SolveResult Krylov::mixed(const double *val, const double *b, double *x, double tol, int maxit, PcKind pc,
                          int solver, double inner_tol, bool coarse, bool reuse) {
  // a system whose FP32 iteration stalled once is solved in FP64 from then on
  if (fp32_stalled_) {
    SolveResult F = solver == 0 ? pcg(val, b, x, tol, maxit, pc, coarse, reuse)
                                : bicgstab(val, b, x, tol, maxit, pc, coarse);
    F.stalled = 1;
    F.it64 = F.it;
    return F;
  }
  SolveResult R;
  Mat A = matof(pat);
  const dim3 G(nblk_), B(kBlock);
  hipStream_t S = stream();
  ensure32();
  if (val32_.n < (size_t)pat->nstore) val32_.alloc(pat->nstore);
  KS *st = st_.p;
  double *part = part_.p;
  KS h{};
  h.tol = tol;
  h.maxit = maxit;
  h.delta = inner_tol;
  MS_HIP(hipMemcpyAsync(st, &h, sizeof(h), hipMemcpyHostToDevice, S));
  hipLaunchKernelGGL(k_sumsq<double>, G, B, 0, S, n, b, part);
  finish<OpSum, 1>(nblk_, part, FinBnorm{st});
  MS_CHECK_LAUNCH();
  KS hs = read_state<double>(st);
  if (hs.bnrm == 0.0) {
    MS_HIP(hipMemsetAsync(x, 0, n * sizeof(double), S));
    return R;
  }
  coarse = coarse && pat->coarse.k > 0;
  const SetupKey key{val, (int)pc, 1, (int)coarse};
  const bool reused = reuse && last_ == key;
  if (!reused) {
    last_ = SetupKey{};
    hipLaunchKernelGGL(k_mat32_diff, G, B, 0, S, n, A, val, val32_.p);
    setup_pc32(pc, val);
    if (coarse && solver == 1) coarse_setup(val);
  }
  hipLaunchKernelGGL(k_residual<double>, G, B, 0, S, n, A, val, b, x, r_.p, part);
  finish<OpSum, 1>(nblk_, part, FinResidual{st});
  MS_CHECK_LAUNCH();
  hs = read_state<double>(st);
  R.res = hs.res;
  if (hs.done) return R;

  if (solver == 1) {
    // BiCGSTAB: FP32 inner solves inside FP64 iterative refinement.  An
    // inner solve ends at its tolerance or when its residual stops falling
    // (stall window); a refinement pass that does not halve the FP64
    // residual hands the system to FP64 BiCGSTAB (the remaining budget, from
    // the current x), and the system stays in FP64 from then on.
    int total = 0;
    double prev = R.res;
    while (true) {
      int imax = maxit - total;
      hipLaunchKernelGGL(k_mp_scale_down, G, B, 0, S, n, r_.p, st, r32_.p, x32_.p, st32_.p, kRefineTol, imax);
      SolveResult in = bicg32(pc, imax, stall_window(), coarse);
      total += std::max(in.it, 1);
      R.outer++;
      hipLaunchKernelGGL(k_mp_correct, G, B, 0, S, n, x32_.p, st, x);
      hipLaunchKernelGGL(k_residual<double>, G, B, 0, S, n, A, val, b, x, r_.p, part);
      finish<OpSum, 1>(nblk_, part, FinResidual{st});
      MS_CHECK_LAUNCH();
      hs = read_state<double>(st);
      R.res = hs.res;
      if (hs.done || total >= maxit || R.outer >= 200) break;
      if (in.ierr == 2 && in.it <= 1) {
        R.ierr = 2;
        break;
      }
      if (!(hs.res <= 0.5 * prev)) {
        fp32_stalled_ = true;
        const SolveResult F = bicgstab(val, b, x, tol, maxit - total, pc, coarse);
        R.it = total + F.it;
        R.it32 = total;
        R.it64 = F.it;
        R.res = F.res;
        R.ierr = F.ierr;
        R.stalled = 1;
        return R;
      }
      prev = hs.res;
    }
    R.it = total;
    R.it32 = total;
    if (R.ierr == 0 && !(R.res <= tol)) R.ierr = 1;
    return R;
  }

  // PCG with reliable updates: y accumulates alpha p in FP32 between updates.
  // The FP32 iteration pauses when its recursive residual meets the
  // tolerance; a drop below delta times its running maximum only flags an
  // update, which the host applies at the end of the chunk (after the
  // p update: r and p are rescaled, the recurrence continues).
  float *r = r32_.p, *z = z32_.p, *p = p32_.p, *q = q32_.p, *y = x32_.p;
  float *part32 = part32_.p;
  const float *v32 = val32_.p, *dg = dg32_.p, *fac = fac32_.p;
  const int *done = &st->done;
  const int jac = pc == PC_JACOBI;
  const int mode = jac ? 1 : (pc == PC_DILU ? 2 : 0);
  const int c0end = pat->color_ptr[1];
  // DILU with a coarse space: the coarse solve rides along the forward sweep
  const bool fused = coarse && pc == PC_DILU && pat->color_ptr.size() > 2;
  if (coarse) {
    h.coarse = fused ? 0 : 1;
    MS_HIP(hipMemcpyAsync(&st->coarse, &h.coarse, sizeof(int), hipMemcpyHostToDevice, S));
    if (!reused) coarse_setup(val);
  }
  last_ = key;
  const Rows RW = rows_of(pat, coarse);
  const dim3 GR(coarse ? pat->coarse.nchunks : nblk_);
  const CoDev co = codev_of(pat, cm_, coarse && !fused, st);
  const double *nu = coarse ? cm_.nu.p : nullptr;
  const CoWork<float> cw = cowork_of<float>(pat, cm_, fused, part32);
  float *prz = prz32_.p;
  MS_HIP(hipMemsetAsync(y, 0, n * sizeof(float), S));
  hipLaunchKernelGGL(k_set_sigma, dim3(1), dim3(64), 0, S, st);
  hipLaunchKernelGGL(k_scale_down, G, B, 0, S, n, r_.p, r, st);
  if (fused) {
    hipLaunchKernelGGL(k_rsum<float>, GR, B, 0, S, RW, r, st, part32);
    const auto &cp = pat->color_ptr;
    hipLaunchKernelGGL(k_dilu_first<float>, dim3(nblocks(cp[1] - cp[0])), B, 0, S, cp[0], cp[1], dg, r, z, done);
    const int np = dilu_sweeps<float>(pat, A, v32, dg, r, z, done, prz, cw);
    finish<OpSum, 1>(np, prz, FinRz0{st});
    MS_HIP(hipMemsetAsync(p, 0, n * sizeof(float), S));
    hipLaunchKernelGGL(k_pcg_p<float>, GR, B, 0, S, RW, r, z, dg, 0, p, nu, st);   // beta = 0
  } else {
    if (coarse) {
      hipLaunchKernelGGL(k_rsum<float>, GR, B, 0, S, RW, r, st, part32);
      finish_co<3>((int)GR.x, part32, FinNop{st}, co);
    }
    if (!jac) apply_pc<float>(pc, v32, fac, dg, r, z, done);
    hipLaunchKernelGGL(k_pcg_start<float>, GR, B, 0, S, RW, r, z, dg, jac, p, nu, st, part32);
    finish<OpSum, 1>((int)GR.x, part32, FinRz0{st});
  }
  // the part of an iteration after the update kernel: residual check,
  // preconditioner, r.z and p update.  DILU finishes r.r and r.z in one
  // kernel after the sweeps (trigger: reliable-update checks active).
  const bool merged = mode == 2 && (!coarse || fused);
  // Jacobi with a coarse space: the coarse solve is spread over the split
  // finisher's workgroups and beta is formed in the p update
  const bool split = coarse && !fused && mode == 1;
  const int nsplit = split ? (pat->coarse.k + kSplitRows - 1) / kSplitRows : 0;
  if (split && cnu_split_.n < (size_t)nsplit) cnu_split_.alloc(nsplit);
  // the split path's p update rides along the next matrix-free product
  const bool pfused = split && matfree_on() && mf_->agg1 && p_fused_enabled();
  const PUp pu{r, dg, mf_ ? mf_->agg1 : nullptr, nu, cnu_split_.p, nsplit};
  auto tail = [&](int trigger) {
    if (split) {
      hipLaunchKernelGGL((k_finish_split<3, float, FinUpdateSplit>), dim3(1 + nsplit), dim3(kFinThreads), 0, S,
                         (int)GR.x, part32, FinUpdateSplit{st, trigger}, co, cnu_split_.p);
      if (!pfused)
        hipLaunchKernelGGL(k_pcg_p_split<float>, GR, B, 0, S, RW, r, dg, p, nu, st, cnu_split_.p, nsplit, 0);
      return;
    }
    if (merged) {
      const int np = dilu_sweeps<float>(pat, A, v32, dg, r, z, done, prz, cw);
      hipLaunchKernelGGL((k_finish2<float, FinUpdRzRU>), dim3(1), dim3(kFinThreads), 0, S, (int)GR.x, part32, np,
                         prz, FinUpdRzRU{st, trigger});
    } else {
      finish_co<3>((int)GR.x, part32, FinUpdateRU{st, jac, trigger}, co);
      if (mode == 2) {
        int np = dilu_sweeps<float>(pat, A, v32, dg, r, z, done, prz, cw);
        finish<OpSum, 1>(np, prz, FinRz{st});
      } else if (mode == 0) {
        apply_pc<float>(pc, v32, fac, dg, r, z, done);
        hipLaunchKernelGGL(k_dot<float>, G, B, 0, S, n, r, z, st, part32);
        finish<OpSum, 1>(nblk_, part32, FinRz{st});
      }
    }
    hipLaunchKernelGGL(k_pcg_p<float>, GR, B, 0, S, RW, r, z, dg, jac, p, nu, st);
  };
  auto reliable_update = [&](bool paused) {
    if (pfused) {   // apply a pending p update before r, p and z are rescaled
      hipLaunchKernelGGL(k_pcg_p_split<float>, GR, B, 0, S, RW, r, dg, p, nu, st, cnu_split_.p, nsplit, 1);
      hipLaunchKernelGGL(k_clear_pup, dim3(1), dim3(64), 0, S, st);
    }
    hipLaunchKernelGGL(k_fold, G, B, 0, S, n, y, x, st);
    hipLaunchKernelGGL(k_true_res, G, B, 0, S, n, A, val, b, x, r_.p, st, part);
    finish<OpSum, 1>(nblk_, part, FinRU{st});
    hipLaunchKernelGGL(k_ru_rescale, GR, B, 0, S, RW, r_.p, r, p, z, dg, mode, c0end, st, part32);
    if (paused) tail(0);   // complete the interrupted iteration on the new r
    MS_CHECK_LAUNCH();
  };
  constexpr int kChunk = 16;   // iterations between state reads
  const bool probe = g_probe.on;
  if (probe) g_probe.rows = n;
  int probe_pair = -1;   // the probe pair of the chunk last queued
  // split path: the SpMV's p.q partials stay apart from the update's, which
  // the update writes while other workgroups still read the SpMV's
  if (split && pq32_.n < (size_t)nblk_) pq32_.alloc(nblk_);
  float *spart = split ? pq32_.p : part32;
  auto chunk = [&](int m, bool probe_last) {
    for (int it = 0; it < m; it++) {
      if (probe_last && it == m - 1) {
        probe_pair = g_probe.next;
        g_probe.next = (g_probe.next + 1) % kProbePairs;
        hipEvent_t e0 = g_probe.ev[2 * probe_pair], e1 = g_probe.ev[2 * probe_pair + 1];
        spmv32(val, v32, p, q, st, spart, e0, e1, pfused ? &pu : nullptr);
      } else {
        spmv32(val, v32, p, q, st, spart, nullptr, nullptr, pfused ? &pu : nullptr);
      }
      if (split) {   // alpha reduced inside the update (the SpMV's partials go to a buffer of their own)
        hipLaunchKernelGGL(k_pcg_update_fa<float>, GR, B, 0, S, RW, nblk_, pq32_.p, p, q, y, r, dg, st, part32);
      } else {
        finish<OpSum, 1>(nblk_, part32, FinPap{st});
        hipLaunchKernelGGL((k_pcg_update<float, float>), GR, B, 0, S, RW, p, q, y, r, z, dg, mode, c0end, st,
                           part32);
      }
      tail(1);
    }
  };
  hs = read_state<double>(st);
  // Stall guard: when the FP32 recurrence cannot resolve the system (its
  // residual has not halved over kStall iterations), the iteration stops and
  // the FP64 PCG with the same preconditioner continues from the current x.
  // A solve that keeps converging never reaches it, so its result is
  // unchanged.
  const int kStall = stall_window();
  double best = hs.res;
  int it_best = hs.it;
  bool stalled = false;
  // Queue one chunk: the captured graph for a full chunk, else the
  // launches; then its state goes to pinned slot `slot` and an event marks it.
  auto queue_chunk = [&](int m, int slot) {
    probe_pair = -1;
    if (m == kChunk && graphs_enabled()) {
      // every argument of the chunk's launches is fixed by the key (device
      // state lives in st, the iteration count in st->it), so one captured
      // chunk replays for every full chunk of every solve with this set-up;
      // with the probe on, the chunk's last iteration runs after the graph
      const int glen = probe ? kChunk - 1 : kChunk;
      const ChunkKey ck{v32, fac, dg, nu, (int)pc, (int)coarse, (int)fused, glen, (int)matfree_on(), (int)pfused};
      if (!chunk_exec_ || !(chunk_key_ == ck)) {
        std::lock_guard<std::recursive_mutex> lk(device_alloc_mutex());
        if (chunk_exec_) MS_HIP(hipGraphExecDestroy(chunk_exec_));
        chunk_exec_ = nullptr;
        hipGraph_t g = nullptr;
        MS_HIP(hipStreamBeginCapture(S, hipStreamCaptureModeRelaxed));
        try {
          chunk(glen, false);
        } catch (...) {
          (void)hipStreamEndCapture(S, &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        MS_HIP(hipStreamEndCapture(S, &g));
        const hipError_t e = hipGraphInstantiate(&chunk_exec_, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
          chunk_exec_ = nullptr;
          fail(MS_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
        chunk_key_ = ck;
      }
      MS_HIP(hipGraphLaunch(chunk_exec_, S));
      if (probe) chunk(1, true);
    } else {
      chunk(m, probe);
    }
    MS_CHECK_LAUNCH();
    MS_HIP(hipMemcpyAsync(state_slot(slot), st, sizeof(KS), hipMemcpyDeviceToHost, S));
    MS_HIP(hipEventRecord(state_ev_[slot], S));
  };
  // Chunks run one ahead of the host: the next chunk is queued before the
  // state after the current one is read, so the device does not idle while
  // the host decides.  A solve that finished or paused makes the queued
  // chunk's kernels return at once; a reliable update flagged in a chunk is
  // applied after the chunk queued behind it (at most 2 x 16 iterations
  // late), as the state then read is the latest.
  struct Queued {
    int slot, m, pair;
  };
  Queued inq[2];
  int nq = 0, next_slot = 0, queued_it = hs.it;
  const bool ahead = pipeline_enabled();
  auto wait_front = [&]() {
    const Queued f = inq[0];
    inq[0] = inq[1];
    nq--;
    MS_HIP(hipEventSynchronize(state_ev_[f.slot]));
    KS h;
    std::memcpy(&h, state_slot(f.slot), sizeof(KS));
    if (f.pair >= 0) probe_harvest(f.pair);
    return h;
  };
  auto drain = [&]() {
    while (nq > 0) hs = wait_front();
  };
  while (true) {
    while (nq < (ahead ? 2 : 1) && queued_it < maxit) {
      const int m = std::min(kChunk, maxit - queued_it);
      queue_chunk(m, next_slot);
      inq[nq++] = Queued{next_slot, m, probe_pair};
      next_slot ^= 1;
      queued_it += m;
    }
    if (nq == 0) break;
    hs = wait_front();
    if (hs.done == 1) {
      drain();
      break;
    }
    if (hs.res < 0.5 * best) {
      best = hs.res;
      it_best = hs.it;
    } else if (hs.it - it_best >= kStall) {
      drain();
      if (hs.done != 1) stalled = true;
      break;
    }
    if (hs.ru == 1) {   // reliable update flagged, or the solve paused for one
      drain();
      if (hs.done == 1) break;
      if (hs.ru == 1) {
        reliable_update(hs.done == 2);
        hs = read_state<double>(st);
      }
      queued_it = hs.it;
      if (hs.done == 1) break;
    }
  }
  if (hs.done != 1) {    // exhaustion: fold y and report the true residual
    KS force = hs;
    force.ru = 1;
    force.done = 0;
    MS_HIP(hipMemcpyAsync(st, &force, sizeof(KS), hipMemcpyHostToDevice, S));
    hipLaunchKernelGGL(k_fold, G, B, 0, S, n, y, x, st);
    hipLaunchKernelGGL(k_true_res, G, B, 0, S, n, A, val, b, x, r_.p, st, part);
    finish<OpSum, 1>(nblk_, part, FinRU{st});
    MS_CHECK_LAUNCH();
    hs = read_state<double>(st);
  }
  R.it = hs.it;
  R.it32 = hs.it;
  R.res = hs.res;
  R.ierr = hs.ierr;
  R.outer = hs.nru;
  if (stalled && !(hs.res <= tol) && hs.it < maxit) {
    fp32_stalled_ = true;
    // the FP64 PCG gets the remaining budget; on exhaustion it reports its
    // budget + 1, so the total is maxit + 1 as for any exhausted CG
    const SolveResult F = pcg(val, b, x, tol, maxit - hs.it, pc, coarse, false);
    R.it += F.it;
    R.it64 = F.it;
    R.res = F.res;
    R.ierr = F.ierr;
    R.stalled = 1;
    return R;
  }
  if (hs.done != 1) {     // exhaustion: it = maxit + 1, as the FP64 CG (Fortran DO-loop count)
    if (R.ierr == 0) R.ierr = 1;
    R.it = maxit + 1;
  }
  return R;
}

}  // namespace ms
