// ms_linalg.hip -- SELL-64 sparse kernels and device-resident Krylov solvers
// for gfx950.
//
// Recurrences follow the reference solvers at one partition
// (mod_uns_linsolver_mpi.f90: bicgstab_ilu0_mpi :155-229, cg_jacobi_mpi
// :235-300, cg_ic0_mpi :306-363; factorizations mod_uns_linsolver.f90:79-142
// and :328-399).  Scalars (alpha, beta, rho, omega, residuals, iteration
// count, breakdown flags) live in device memory and are updated by the last
// workgroup of each reduction kernel, so an iteration needs no host round
// trip; the host enqueues iterations in chunks and reads the state once per
// chunk.  Every iteration kernel returns at once when the state says done.
//
// Reductions are two-stage and deterministic: wave64 shuffles + LDS inside a
// workgroup, per-workgroup partials, then the last workgroup (ticket counter)
// sums the partials in a fixed order.
#include <cmath>
#include <cstring>

#include "ms_linalg.hpp"
#include "ms_reduce.hpp"

namespace ms {
namespace {

template <class T> struct Tiny;
template <> struct Tiny<double> { static constexpr double v = 2.2250738585072014e-308; };
template <> struct Tiny<float> { static constexpr float v = 1.17549435e-38f; };

template <class T>
__device__ __forceinline__ T row_dot(int r, const long long *off, const int *rowlen, const int *col,
                                     const T *val, const T *x) {
  const long long base = off[r >> 6] + (r & 63);
  const int len = rowlen[r];
  T s = T(0);
  for (int k = 0; k < len; k++) {
    const long long q = base + (long long)k * kSlice;
    s = s + val[q] * x[col[q]];
  }
  return s;
}

struct Mat {
  const long long *off;
  const int *rowlen, *col, *diagk;
};

// ------------------------------------------------------------ basic kernels
template <class T>
__global__ void k_spmv(int n, Mat A, const T *val, const T *x, T *y) {
  int r = blockIdx.x * kBlock + threadIdx.x;
  if (r < n) y[r] = row_dot(r, A.off, A.rowlen, A.col, val, x);
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

// bnrm = tsqrt(b.b)
template <class T>
__global__ void k_bnorm(int n, const T *b, KS<T> *st, T *part) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  T v[1] = {i < n ? b[i] * b[i] : T(0)};
  grid_reduce<OpSum, 1>(v, part, &st->ticket, [&](T *a) { st->bnrm = tsqrt(a[0]); });
}

// r = b - A x; res = ||r|| / bnrm; done if res <= tol
template <class T>
__global__ void k_residual(int n, Mat A, const T *val, const T *b, const T *x, T *r, KS<T> *st, T *part) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  T v[1] = {T(0)};
  if (i < n) {
    T ri = b[i] - row_dot(i, A.off, A.rowlen, A.col, val, x);
    r[i] = ri;
    v[0] = ri * ri;
  }
  grid_reduce<OpSum, 1>(v, part, &st->ticket, [&](T *a) {
    st->rr = a[0];
    st->res = tsqrt(a[0]) / st->bnrm;
    if (st->res <= st->tol) st->done = 1;
  });
}

// ------------------------------------------------------------------- PCG
// z = r/diag (Jacobi) fused variants carry a `jac` flag.
template <class T>
__global__ void k_pcg_start(int n, const T *r, const T *z, const T *diag, int jac, T *p, KS<T> *st, T *part) {
  if (st->done) return;
  int i = blockIdx.x * kBlock + threadIdx.x;
  T v[1] = {T(0)};
  if (i < n) {
    T zi = jac ? r[i] / diag[i] : z[i];
    p[i] = zi;
    v[0] = r[i] * zi;
  }
  grid_reduce<OpSum, 1>(v, part, &st->ticket, [&](T *a) { st->rz = a[0]; });
}

template <class T>
__global__ void k_pcg_spmv(int n, Mat A, const T *val, const T *p, T *q, KS<T> *st, T *part) {
  if (st->done) return;
  int i = blockIdx.x * kBlock + threadIdx.x;
  T v[1] = {T(0)};
  if (i < n) {
    T qi = row_dot(i, A.off, A.rowlen, A.col, val, p);
    q[i] = qi;
    v[0] = p[i] * qi;
  }
  grid_reduce<OpSum, 1>(v, part, &st->ticket, [&](T *a) {
    st->it = st->it + 1;
    st->pap = a[0];
    if (tabs(a[0]) < Tiny<T>::v) {
      st->ierr = 2;
      st->done = 1;
    } else {
      st->alpha = st->rz / a[0];
    }
  });
}

// x += alpha p; r -= alpha q; res check; Jacobi: rz_new = r.(r/diag), beta
template <class T>
__global__ void k_pcg_update(int n, const T *p, const T *q, T *x, T *r, const T *diag, int jac, KS<T> *st,
                             T *part) {
  if (st->done) return;
  int i = blockIdx.x * kBlock + threadIdx.x;
  T v[2] = {T(0), T(0)};
  if (i < n) {
    const T alpha = st->alpha;
    x[i] = x[i] + alpha * p[i];
    T ri = r[i] - alpha * q[i];
    r[i] = ri;
    v[0] = ri * ri;
    if (jac) v[1] = ri * (ri / diag[i]);
  }
  grid_reduce<OpSum, 2>(v, part, &st->ticket, [&](T *a) {
    st->rr = a[0];
    st->res = tsqrt(a[0]) / st->bnrm;
    if (st->res <= st->tol) {
      st->done = 1;
      return;
    }
    if (jac) {
      if (tabs(st->rz) < Tiny<T>::v) {
        st->ierr = 2;
        st->done = 1;
      } else {
        st->beta = a[1] / st->rz;
        st->rz = a[1];
      }
    }
  });
}

template <class T>
__global__ void k_pcg_rz(int n, const T *r, const T *z, KS<T> *st, T *part) {
  if (st->done) return;
  int i = blockIdx.x * kBlock + threadIdx.x;
  T v[1] = {i < n ? r[i] * z[i] : T(0)};
  grid_reduce<OpSum, 1>(v, part, &st->ticket, [&](T *a) {
    if (tabs(st->rz) < Tiny<T>::v) {
      st->ierr = 2;
      st->done = 1;
    } else {
      st->beta = a[0] / st->rz;
      st->rz = a[0];
    }
  });
}

template <class T>
__global__ void k_pcg_p(int n, const T *r, const T *z, const T *diag, int jac, T *p, const KS<T> *st) {
  if (st->done) return;
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) {
    T zi = jac ? r[i] / diag[i] : z[i];
    p[i] = zi + st->beta * p[i];
  }
}

// -------------------------------------------------------------- BiCGSTAB
template <class T>
__global__ void k_bi_init(int n, const T *r, T *r0, T *p, T *v, KS<T> *st) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) {
    r0[i] = r[i];
    p[i] = T(0);
    v[i] = T(0);
  }
  if (i == 0) {
    st->rho_old = T(1);
    st->alpha = T(1);
    st->omega = T(1);
  }
}

template <class T>
__global__ void k_bi_rho(int n, const T *r, const T *r0, KS<T> *st, T *part) {
  if (st->done) return;
  int i = blockIdx.x * kBlock + threadIdx.x;
  T v[1] = {i < n ? r[i] * r0[i] : T(0)};
  grid_reduce<OpSum, 1>(v, part, &st->ticket, [&](T *a) {
    st->it = st->it + 1;
    st->rho = a[0];
    if (tabs(a[0]) < Tiny<T>::v) {
      st->ierr = 2;
      st->done = 1;
    } else {
      st->beta = (a[0] / st->rho_old) * (st->alpha / st->omega);
    }
  });
}

template <class T>
__global__ void k_bi_p(int n, const T *r, const T *v, T *p, const KS<T> *st) {
  if (st->done) return;
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) p[i] = r[i] + st->beta * (p[i] - st->omega * v[i]);
}

// v = A ph ; alpha = rho / (r0.v)
template <class T>
__global__ void k_bi_spmv_alpha(int n, Mat A, const T *val, const T *ph, const T *r0, T *v, KS<T> *st, T *part) {
  if (st->done) return;
  int i = blockIdx.x * kBlock + threadIdx.x;
  T w[1] = {T(0)};
  if (i < n) {
    T vi = row_dot(i, A.off, A.rowlen, A.col, val, ph);
    v[i] = vi;
    w[0] = r0[i] * vi;
  }
  grid_reduce<OpSum, 1>(w, part, &st->ticket, [&](T *a) { st->alpha = st->rho / a[0]; });
}

// s = r - alpha v ; early exit when ||s||/bnrm <= tol (done = 2: x += alpha ph pending)
template <class T>
__global__ void k_bi_s(int n, const T *r, const T *v, T *s, KS<T> *st, T *part) {
  if (st->done) return;
  int i = blockIdx.x * kBlock + threadIdx.x;
  T w[1] = {T(0)};
  if (i < n) {
    T si = r[i] - st->alpha * v[i];
    s[i] = si;
    w[0] = si * si;
  }
  grid_reduce<OpSum, 1>(w, part, &st->ticket, [&](T *a) {
    st->snrm = tsqrt(a[0]);
    if (st->snrm / st->bnrm <= st->tol) {
      st->res = st->snrm / st->bnrm;
      st->done = 2;
    }
  });
}

// the pending early-exit update x += alpha ph (runs once: done 2 -> 1)
template <class T>
__global__ void k_bi_early(int n, const T *ph, T *x, KS<T> *st, T *part) {
  if (st->done != 2) return;
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] = x[i] + st->alpha * ph[i];
  T w[1] = {T(0)};
  grid_reduce<OpSum, 1>(w, part, &st->ticket, [&](T *) { st->done = 1; });
}

// t = A sh ; omega = (t.s)/(t.t) with breakdown checks
template <class T>
__global__ void k_bi_spmv_omega(int n, Mat A, const T *val, const T *sh, const T *s, T *t, KS<T> *st, T *part) {
  if (st->done) return;
  int i = blockIdx.x * kBlock + threadIdx.x;
  T w[2] = {T(0), T(0)};
  if (i < n) {
    T ti = row_dot(i, A.off, A.rowlen, A.col, val, sh);
    t[i] = ti;
    w[0] = ti * ti;
    w[1] = ti * s[i];
  }
  grid_reduce<OpSum, 2>(w, part, &st->ticket, [&](T *a) {
    st->tt = a[0];
    if (a[0] < Tiny<T>::v) {
      st->ierr = 2;
      st->done = 1;
      return;
    }
    st->omega = a[1] / a[0];
    if (tabs(st->omega) < Tiny<T>::v) {
      st->ierr = 2;
      st->done = 1;
    }
  });
}

template <class T>
__global__ void k_bi_xr(int n, const T *ph, const T *sh, const T *s, const T *t, T *x, T *r, KS<T> *st, T *part) {
  if (st->done) return;
  int i = blockIdx.x * kBlock + threadIdx.x;
  T w[1] = {T(0)};
  if (i < n) {
    x[i] = x[i] + st->alpha * ph[i] + st->omega * sh[i];
    T ri = s[i] - st->omega * t[i];
    r[i] = ri;
    w[0] = ri * ri;
  }
  grid_reduce<OpSum, 1>(w, part, &st->ticket, [&](T *a) {
    st->rr = a[0];
    st->res = tsqrt(a[0]) / st->bnrm;
    if (st->res <= st->tol) {
      st->done = 1;
      return;
    }
    st->rho_old = st->rho;
  });
}

// ------------------------------------------------ preconditioner kernels
template <class T>
__global__ void k_jacobi(int n, const T *in, const T *diag, T *out, const int *done) {
  if (done && *done) return;
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) out[i] = in[i] / diag[i];
}

template <class T>
__global__ void k_copy(int n, const T *in, T *out, const int *done) {
  if (done && *done) return;
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) out[i] = in[i];
}

// natural IC(0), one level of rows (mod_uns_linsolver.f90:328-365)
template <class T>
__global__ void k_ic0_factor(const int *rows, int nr, Mat A, T *F, int *flag) {
  int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= nr) return;
  const int i = rows[t];
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

// natural ILU(0), IKJ variant, one level of rows (mod_uns_linsolver.f90:79-115)
template <class T>
__global__ void k_ilu0_factor(const int *rows, int nr, Mat A, T *F, int *flag) {
  int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= nr) return;
  const int i = rows[t];
  const long long bi = A.off[i >> 6] + (i & 63);
  const int dpi = A.diagk[i], leni = A.rowlen[i];
  for (int kk = 0; kk < dpi; kk++) {
    const long long q = bi + (long long)kk * kSlice;
    const int k = A.col[q];
    const long long bk = A.off[k >> 6] + (k & 63);
    const int dpk = A.diagk[k], lenk = A.rowlen[k];
    const T piv = F[bk + (long long)dpk * kSlice];
    if (tabs(piv) < Tiny<T>::v) {
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

// forward sweep rows of one level: ILU0 (unit L, in place) or IC0 (L, / diag)
template <class T>
__global__ void k_lower_level(const int *rows, int nr, Mat A, const T *F, const T *r, T *z, int ic0, const int *done) {
  if (done && *done) return;
  int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= nr) return;
  const int i = rows[t];
  const long long bi = A.off[i >> 6] + (i & 63);
  const int dpi = A.diagk[i];
  T s = r[i];
  for (int k = 0; k < dpi; k++) {
    const long long q = bi + (long long)k * kSlice;
    s = s - F[q] * z[A.col[q]];
  }
  z[i] = ic0 ? s / F[bi + (long long)dpi * kSlice] : s;
}

// backward sweep rows of one level.  ILU0: U z = y with U from row i.
// IC0: L^T z = y, entries L(j,i) read through the transpose slot map in
// descending j (the column-oriented order of ic0_apply, :390-397).
template <class T>
__global__ void k_upper_level(const int *rows, int nr, Mat A, const long long *tpos, const T *F, T *z, int ic0,
                              const int *done) {
  if (done && *done) return;
  int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= nr) return;
  const int i = rows[t];
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

// multicolour DILU: d_i = a_ii - sum_{j<i} a_ij a_ji / d_j for one colour block
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

template <class T>
__global__ void k_dilu_fwd(int r0, int r1, Mat A, const T *val, const T *d, const T *r, T *z, const int *done) {
  if (done && *done) return;
  int i = r0 + blockIdx.x * kBlock + threadIdx.x;
  if (i >= r1) return;
  const long long bi = A.off[i >> 6] + (i & 63);
  const int dpi = A.diagk[i];
  T s = r[i];
  for (int k = 0; k < dpi; k++) {
    const long long q = bi + (long long)k * kSlice;
    s = s - val[q] * z[A.col[q]];
  }
  z[i] = s / d[i];
}

template <class T>
__global__ void k_dilu_bwd(int r0, int r1, Mat A, const T *val, const T *d, T *z, const int *done) {
  if (done && *done) return;
  int i = r0 + blockIdx.x * kBlock + threadIdx.x;
  if (i >= r1) return;
  const long long bi = A.off[i >> 6] + (i & 63);
  const int dpi = A.diagk[i], len = A.rowlen[i];
  T s = T(0);
  for (int k = dpi + 1; k < len; k++) {
    const long long q = bi + (long long)k * kSlice;
    s = s + val[q] * z[A.col[q]];
  }
  z[i] = z[i] - s / d[i];
}

// --------------------------------------------------- mixed precision glue
// r = b - A x (FP64), rr; then r32 = r / ||r|| and reset the inner state
__global__ void k_mp_residual(int n, Mat A, const double *val, const double *b, const double *x, double *r,
                              KS<double> *st, double *part) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  double v[1] = {0.0};
  if (i < n) {
    double ri = b[i] - row_dot(i, A.off, A.rowlen, A.col, val, x);
    r[i] = ri;
    v[0] = ri * ri;
  }
  grid_reduce<OpSum, 1>(v, part, &st->ticket, [&](double *a) {
    st->rr = a[0];
    st->res = tsqrt(a[0]) / st->bnrm;
    if (st->res <= st->tol) st->done = 1;
  });
}

__global__ void k_mp_scale_down(int n, const double *r, const KS<double> *st, float *r32, float *x32,
                                KS<float> *s32, float inner_tol, int inner_max) {
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
    s32->bnrm = 1.0f;
    s32->tol = inner_tol;
    s32->maxit = inner_max;
    s32->ticket = 0u;
  }
}

__global__ void k_mp_correct(int n, const float *x32, const KS<double> *st, double *x) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] = x[i] + sqrt(st->rr) * (double)x32[i];
}

}  // namespace

// ============================================================ host side
static Mat matof(const SellPattern *p) { return Mat{p->d_off.p, p->d_rowlen.p, p->d_col.p, p->d_diagk.p}; }

Krylov::Krylov(SellPattern *pat_) : pat(pat_), n(pat_->n) {
  nblk_ = nblocks(n);
  for (DBuf<double> *b : {&r_, &r0_, &p_, &q_, &z_, &s_, &t_, &ph_, &sh_, &dinv_}) b->alloc(n);
  part_.alloc(2 * (size_t)nblk_ + 8);
  st_.alloc(1);
  flag_.alloc(1);
  MS_HIP(hipHostMalloc(&hpin_, 4096, hipHostMallocDefault));
}

Krylov::~Krylov() {
  if (hpin_) (void)hipHostFree(hpin_);
}

template <class T>
KS<T> Krylov::read_state(KS<T> *d) {
  MS_HIP(hipMemcpyAsync(hpin_, d, sizeof(KS<T>), hipMemcpyDeviceToHost, stream()));
  MS_HIP(hipStreamSynchronize(stream()));
  KS<T> h;
  std::memcpy(&h, hpin_, sizeof(KS<T>));
  return h;
}

void Krylov::spmv(const double *val, const double *x, double *y) {
  hipLaunchKernelGGL(k_spmv<double>, dim3(nblk_), dim3(kBlock), 0, stream(), n, matof(pat), val, x, y);
  MS_CHECK_LAUNCH();
}

double Krylov::true_residual(const double *val, const double *b, const double *x) {
  KS<double> h{};
  h.tol = -1.0;
  MS_HIP(hipMemcpyAsync(st_.p, &h, sizeof(h), hipMemcpyHostToDevice, stream()));
  hipLaunchKernelGGL(k_bnorm<double>, dim3(nblk_), dim3(kBlock), 0, stream(), n, b, st_.p, part_.p);
  hipLaunchKernelGGL(k_residual<double>, dim3(nblk_), dim3(kBlock), 0, stream(), n, matof(pat), val, b, x, r_.p,
                     st_.p, part_.p);
  MS_CHECK_LAUNCH();
  return read_state(st_.p).res;
}

// ---- preconditioners ------------------------------------------------------
int Krylov::setup_pc64(PcKind pc, const double *val) {
  Mat A = matof(pat);
  if (pc == PC_JACOBI) {
    hipLaunchKernelGGL(k_extract_diag<double>, dim3(nblk_), dim3(kBlock), 0, stream(), n, A, val, dinv_.p);
    MS_CHECK_LAUNCH();
    return 0;
  }
  if ((pc == PC_DILU || pc == PC_NATURAL_IC0) && !pat->symmetric)
    fail(MS_EINVAL, "IC0 / DILU need a structurally symmetric pattern");
  if (pc == PC_DILU) {
    for (size_t c = 0; c + 1 < pat->color_ptr.size(); c++) {
      int r0 = pat->color_ptr[c], r1 = pat->color_ptr[c + 1];
      if (r1 > r0)
        hipLaunchKernelGGL(k_dilu_setup<double>, dim3(nblocks(r1 - r0)), dim3(kBlock), 0, stream(), r0, r1, A,
                           pat->d_tpos.p, val, dinv_.p);
    }
    MS_CHECK_LAUNCH();
    return 0;
  }
  // natural-order IC0 / ILU0 on a private copy (the reference factors a copy per call)
  if (!pat->has_levels) fail(MS_EINVAL, "natural-order factorization needs level schedules");
  if (fac_.n < (size_t)pat->nstore) fac_.alloc(pat->nstore);
  MS_HIP(hipMemcpyAsync(fac_.p, val, pat->nstore * sizeof(double), hipMemcpyDeviceToDevice, stream()));
  MS_HIP(hipMemsetAsync(flag_.p, 0, sizeof(int), stream()));
  for (size_t l = 0; l + 1 < pat->lev_ptr_f.size(); l++) {
    int a = pat->lev_ptr_f[l], b = pat->lev_ptr_f[l + 1];
    if (pc == PC_NATURAL_IC0)
      hipLaunchKernelGGL(k_ic0_factor<double>, dim3(nblocks(b - a)), dim3(kBlock), 0, stream(),
                         pat->d_lev_rows_f.p + a, b - a, A, fac_.p, flag_.p);
    else
      hipLaunchKernelGGL(k_ilu0_factor<double>, dim3(nblocks(b - a)), dim3(kBlock), 0, stream(),
                         pat->d_lev_rows_f.p + a, b - a, A, fac_.p, flag_.p);
  }
  MS_CHECK_LAUNCH();
  int h = 0;
  MS_HIP(hipMemcpyAsync(hpin_, flag_.p, sizeof(int), hipMemcpyDeviceToHost, stream()));
  MS_HIP(hipStreamSynchronize(stream()));
  std::memcpy(&h, hpin_, sizeof(int));
  return h;
}

void Krylov::apply_pc64(PcKind pc, const double *val, const double *in, double *out) {
  Mat A = matof(pat);
  const int *done = &st_.p->done;
  if (pc == PC_JACOBI) {
    hipLaunchKernelGGL(k_jacobi<double>, dim3(nblk_), dim3(kBlock), 0, stream(), n, in, dinv_.p, out, done);
  } else if (pc == PC_DILU) {
    const auto &cp = pat->color_ptr;
    int nc = (int)cp.size() - 1;
    for (int c = 0; c < nc; c++)
      if (cp[c + 1] > cp[c])
        hipLaunchKernelGGL(k_dilu_fwd<double>, dim3(nblocks(cp[c + 1] - cp[c])), dim3(kBlock), 0, stream(), cp[c],
                           cp[c + 1], A, val, dinv_.p, in, out, done);
    for (int c = nc - 1; c >= 0; c--)
      if (cp[c + 1] > cp[c])
        hipLaunchKernelGGL(k_dilu_bwd<double>, dim3(nblocks(cp[c + 1] - cp[c])), dim3(kBlock), 0, stream(), cp[c],
                           cp[c + 1], A, val, dinv_.p, out, done);
  } else {
    const int ic0 = pc == PC_NATURAL_IC0;
    for (size_t l = 0; l + 1 < pat->lev_ptr_f.size(); l++) {
      int a = pat->lev_ptr_f[l], b = pat->lev_ptr_f[l + 1];
      hipLaunchKernelGGL(k_lower_level<double>, dim3(nblocks(b - a)), dim3(kBlock), 0, stream(),
                         pat->d_lev_rows_f.p + a, b - a, A, fac_.p, ic0 ? in : in, out, ic0, done);
    }
    for (size_t l = 0; l + 1 < pat->lev_ptr_b.size(); l++) {
      int a = pat->lev_ptr_b[l], b = pat->lev_ptr_b[l + 1];
      hipLaunchKernelGGL(k_upper_level<double>, dim3(nblocks(b - a)), dim3(kBlock), 0, stream(),
                         pat->d_lev_rows_b.p + a, b - a, A, pat->d_tpos.p, fac_.p, out, ic0, done);
    }
  }
  MS_CHECK_LAUNCH();
}

void Krylov::setup_pc32(PcKind pc, const double *val) {
  Mat A = matof(pat);
  if (pc == PC_JACOBI) {
    hipLaunchKernelGGL(k_extract_diag<float>, dim3(nblk_), dim3(kBlock), 0, stream(), n, A, val32_.p, dinv32_.p);
  } else if (pc == PC_DILU) {
    for (size_t c = 0; c + 1 < pat->color_ptr.size(); c++) {
      int r0 = pat->color_ptr[c], r1 = pat->color_ptr[c + 1];
      if (r1 > r0)
        hipLaunchKernelGGL(k_dilu_setup<float>, dim3(nblocks(r1 - r0)), dim3(kBlock), 0, stream(), r0, r1, A,
                           pat->d_tpos.p, val32_.p, dinv32_.p);
    }
  } else {
    // factor in FP64, keep the factor in FP32
    int e = setup_pc64(pc, val);
    if (e) fail(MS_EINVAL, "factorization failed in mixed-precision setup");
    if (fac32_.n < (size_t)pat->nstore) fac32_.alloc(pat->nstore);
    hipLaunchKernelGGL(k_to_float, dim3(nblocks(pat->nstore)), dim3(kBlock), 0, stream(), pat->nstore, fac_.p,
                       fac32_.p);
  }
  MS_CHECK_LAUNCH();
}

void Krylov::apply_pc32(PcKind pc, const float *in, float *out) {
  Mat A = matof(pat);
  const int *done = &st32_.p->done;
  if (pc == PC_JACOBI) {
    hipLaunchKernelGGL(k_jacobi<float>, dim3(nblk_), dim3(kBlock), 0, stream(), n, in, dinv32_.p, out, done);
  } else if (pc == PC_DILU) {
    const auto &cp = pat->color_ptr;
    int nc = (int)cp.size() - 1;
    for (int c = 0; c < nc; c++)
      if (cp[c + 1] > cp[c])
        hipLaunchKernelGGL(k_dilu_fwd<float>, dim3(nblocks(cp[c + 1] - cp[c])), dim3(kBlock), 0, stream(), cp[c],
                           cp[c + 1], A, val32_.p, dinv32_.p, in, out, done);
    for (int c = nc - 1; c >= 0; c--)
      if (cp[c + 1] > cp[c])
        hipLaunchKernelGGL(k_dilu_bwd<float>, dim3(nblocks(cp[c + 1] - cp[c])), dim3(kBlock), 0, stream(), cp[c],
                           cp[c + 1], A, val32_.p, dinv32_.p, out, done);
  } else {
    const int ic0 = pc == PC_NATURAL_IC0;
    for (size_t l = 0; l + 1 < pat->lev_ptr_f.size(); l++) {
      int a = pat->lev_ptr_f[l], b = pat->lev_ptr_f[l + 1];
      hipLaunchKernelGGL(k_lower_level<float>, dim3(nblocks(b - a)), dim3(kBlock), 0, stream(),
                         pat->d_lev_rows_f.p + a, b - a, A, fac32_.p, in, out, ic0, done);
    }
    for (size_t l = 0; l + 1 < pat->lev_ptr_b.size(); l++) {
      int a = pat->lev_ptr_b[l], b = pat->lev_ptr_b[l + 1];
      hipLaunchKernelGGL(k_upper_level<float>, dim3(nblocks(b - a)), dim3(kBlock), 0, stream(),
                         pat->d_lev_rows_b.p + a, b - a, A, pat->d_tpos.p, fac32_.p, out, ic0, done);
    }
  }
  MS_CHECK_LAUNCH();
}

// ---- PCG ---------------------------------------------------------------------
SolveResult Krylov::pcg(const double *val, const double *b, double *x, double tol, int maxit, PcKind pc) {
  SolveResult R;
  Mat A = matof(pat);
  const dim3 G(nblk_), B(kBlock);
  hipStream_t S = stream();
  KS<double> h{};
  h.tol = tol;
  h.maxit = maxit;
  MS_HIP(hipMemcpyAsync(st_.p, &h, sizeof(h), hipMemcpyHostToDevice, S));
  hipLaunchKernelGGL(k_bnorm<double>, G, B, 0, S, n, b, st_.p, part_.p);
  MS_CHECK_LAUNCH();
  KS<double> hs = read_state(st_.p);
  if (hs.bnrm == 0.0) {
    MS_HIP(hipMemsetAsync(x, 0, n * sizeof(double), S));
    return R;
  }
  int fe = setup_pc64(pc, val);
  if (fe) {
    R.ierr = fe + 10;
    return R;
  }
  hipLaunchKernelGGL(k_residual<double>, G, B, 0, S, n, A, val, b, x, r_.p, st_.p, part_.p);
  MS_CHECK_LAUNCH();
  hs = read_state(st_.p);
  R.res = hs.res;
  if (hs.done) return R;
  const int jac = pc == PC_JACOBI;
  if (!jac) apply_pc64(pc, val, r_.p, z_.p);
  hipLaunchKernelGGL(k_pcg_start<double>, G, B, 0, S, n, r_.p, z_.p, dinv_.p, jac, p_.p, st_.p, part_.p);
  int enq = 0, chunk = 4;
  while (true) {
    int m = std::min(chunk, maxit - enq);
    for (int k = 0; k < m; k++) {
      hipLaunchKernelGGL(k_pcg_spmv<double>, G, B, 0, S, n, A, val, p_.p, q_.p, st_.p, part_.p);
      hipLaunchKernelGGL(k_pcg_update<double>, G, B, 0, S, n, p_.p, q_.p, x, r_.p, dinv_.p, jac, st_.p, part_.p);
      if (!jac) {
        apply_pc64(pc, val, r_.p, z_.p);
        hipLaunchKernelGGL(k_pcg_rz<double>, G, B, 0, S, n, r_.p, z_.p, st_.p, part_.p);
      }
      hipLaunchKernelGGL(k_pcg_p<double>, G, B, 0, S, n, r_.p, z_.p, dinv_.p, jac, p_.p, st_.p);
    }
    MS_CHECK_LAUNCH();
    enq += m;
    hs = read_state(st_.p);
    if (hs.done || enq >= maxit) break;
    chunk = std::min(chunk * 2, 64);
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
SolveResult Krylov::bicgstab(const double *val, const double *b, double *x, double tol, int maxit, PcKind pc) {
  SolveResult R;
  Mat A = matof(pat);
  const dim3 G(nblk_), B(kBlock);
  hipStream_t S = stream();
  KS<double> h{};
  h.tol = tol;
  h.maxit = maxit;
  MS_HIP(hipMemcpyAsync(st_.p, &h, sizeof(h), hipMemcpyHostToDevice, S));
  hipLaunchKernelGGL(k_bnorm<double>, G, B, 0, S, n, b, st_.p, part_.p);
  MS_CHECK_LAUNCH();
  KS<double> hs = read_state(st_.p);
  if (hs.bnrm == 0.0) {
    MS_HIP(hipMemsetAsync(x, 0, n * sizeof(double), S));
    return R;
  }
  int fe = setup_pc64(pc, val);
  if (fe) {
    R.ierr = fe + 10;
    return R;
  }
  hipLaunchKernelGGL(k_residual<double>, G, B, 0, S, n, A, val, b, x, r_.p, st_.p, part_.p);
  hipLaunchKernelGGL(k_bi_init<double>, G, B, 0, S, n, r_.p, r0_.p, p_.p, q_.p, st_.p);
  MS_CHECK_LAUNCH();
  hs = read_state(st_.p);
  R.res = hs.res;
  if (hs.done) return R;
  double *v = q_.p;
  int enq = 0, chunk = 2;
  while (true) {
    int m = std::min(chunk, maxit - enq);
    for (int k = 0; k < m; k++) {
      hipLaunchKernelGGL(k_bi_rho<double>, G, B, 0, S, n, r_.p, r0_.p, st_.p, part_.p);
      hipLaunchKernelGGL(k_bi_p<double>, G, B, 0, S, n, r_.p, v, p_.p, st_.p);
      apply_pc64(pc, val, p_.p, ph_.p);
      hipLaunchKernelGGL(k_bi_spmv_alpha<double>, G, B, 0, S, n, A, val, ph_.p, r0_.p, v, st_.p, part_.p);
      hipLaunchKernelGGL(k_bi_s<double>, G, B, 0, S, n, r_.p, v, s_.p, st_.p, part_.p);
      hipLaunchKernelGGL(k_bi_early<double>, G, B, 0, S, n, ph_.p, x, st_.p, part_.p);
      apply_pc64(pc, val, s_.p, sh_.p);
      hipLaunchKernelGGL(k_bi_spmv_omega<double>, G, B, 0, S, n, A, val, sh_.p, s_.p, t_.p, st_.p, part_.p);
      hipLaunchKernelGGL(k_bi_xr<double>, G, B, 0, S, n, ph_.p, sh_.p, s_.p, t_.p, x, r_.p, st_.p, part_.p);
    }
    MS_CHECK_LAUNCH();
    enq += m;
    hs = read_state(st_.p);
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

// ---- FP32 inner solves ------------------------------------------------------------
SolveResult Krylov::inner32(int solver, PcKind pc, double inner_tol, int maxit) {
  SolveResult R;
  Mat A = matof(pat);
  const dim3 G(nblk_), B(kBlock);
  hipStream_t S = stream();
  KS<float> *st = st32_.p;
  float *part = part32_.p;
  const int jac = pc == PC_JACOBI;
  int enq = 0, chunk = 8;
  KS<float> hs{};
  if (solver == 0) {
    // x32 = 0, r32 = scaled residual: z = M^-1 r, p = z, rz
    if (!jac) apply_pc32(pc, r32_.p, z32_.p);
    hipLaunchKernelGGL(k_pcg_start<float>, G, B, 0, S, n, r32_.p, z32_.p, dinv32_.p, jac, p32_.p, st, part);
    while (true) {
      int m = std::min(chunk, maxit - enq);
      for (int k = 0; k < m; k++) {
        hipLaunchKernelGGL(k_pcg_spmv<float>, G, B, 0, S, n, A, val32_.p, p32_.p, q32_.p, st, part);
        hipLaunchKernelGGL(k_pcg_update<float>, G, B, 0, S, n, p32_.p, q32_.p, x32_.p, r32_.p, dinv32_.p, jac, st,
                           part);
        if (!jac) {
          apply_pc32(pc, r32_.p, z32_.p);
          hipLaunchKernelGGL(k_pcg_rz<float>, G, B, 0, S, n, r32_.p, z32_.p, st, part);
        }
        hipLaunchKernelGGL(k_pcg_p<float>, G, B, 0, S, n, r32_.p, z32_.p, dinv32_.p, jac, p32_.p, st);
      }
      MS_CHECK_LAUNCH();
      enq += m;
      hs = read_state(st);
      if (hs.done || enq >= maxit) break;
      chunk = std::min(chunk * 2, 64);
    }
  } else {
    float *v = q32_.p;
    hipLaunchKernelGGL(k_bi_init<float>, G, B, 0, S, n, r32_.p, r032_.p, p32_.p, v, st);
    while (true) {
      int m = std::min(chunk, maxit - enq);
      for (int k = 0; k < m; k++) {
        hipLaunchKernelGGL(k_bi_rho<float>, G, B, 0, S, n, r32_.p, r032_.p, st, part);
        hipLaunchKernelGGL(k_bi_p<float>, G, B, 0, S, n, r32_.p, v, p32_.p, st);
        apply_pc32(pc, p32_.p, ph32_.p);
        hipLaunchKernelGGL(k_bi_spmv_alpha<float>, G, B, 0, S, n, A, val32_.p, ph32_.p, r032_.p, v, st, part);
        hipLaunchKernelGGL(k_bi_s<float>, G, B, 0, S, n, r32_.p, v, s32_.p, st, part);
        hipLaunchKernelGGL(k_bi_early<float>, G, B, 0, S, n, ph32_.p, x32_.p, st, part);
        apply_pc32(pc, s32_.p, sh32_.p);
        hipLaunchKernelGGL(k_bi_spmv_omega<float>, G, B, 0, S, n, A, val32_.p, sh32_.p, s32_.p, t32_.p, st, part);
        hipLaunchKernelGGL(k_bi_xr<float>, G, B, 0, S, n, ph32_.p, sh32_.p, s32_.p, t32_.p, x32_.p, r32_.p, st,
                           part);
      }
      MS_CHECK_LAUNCH();
      enq += m;
      hs = read_state(st);
      if (hs.done || enq >= maxit) break;
      chunk = std::min(chunk * 2, 32);
    }
  }
  R.it = hs.it;
  R.ierr = hs.ierr;
  R.res = hs.res;
  return R;
}

SolveResult Krylov::mixed(const double *val, const double *b, double *x, double tol, int maxit, PcKind pc,
                          int solver, double inner_tol) {
  SolveResult R;
  Mat A = matof(pat);
  const dim3 G(nblk_), B(kBlock);
  hipStream_t S = stream();
  if (!have32_) {
    for (DBuf<float> *v : {&r32_, &r032_, &p32_, &q32_, &z32_, &s32_, &t32_, &ph32_, &sh32_, &x32_, &dinv32_})
      v->alloc(n);
    part32_.alloc(2 * (size_t)nblk_ + 8);
    st32_.alloc(1);
    have32_ = true;
  }
  if (val32_.n < (size_t)pat->nstore) val32_.alloc(pat->nstore);
  KS<double> h{};
  h.tol = tol;
  h.maxit = maxit;
  MS_HIP(hipMemcpyAsync(st_.p, &h, sizeof(h), hipMemcpyHostToDevice, S));
  hipLaunchKernelGGL(k_bnorm<double>, G, B, 0, S, n, b, st_.p, part_.p);
  MS_CHECK_LAUNCH();
  KS<double> hs = read_state(st_.p);
  if (hs.bnrm == 0.0) {
    MS_HIP(hipMemsetAsync(x, 0, n * sizeof(double), S));
    return R;
  }
  hipLaunchKernelGGL(k_to_float, dim3(nblocks(pat->nstore)), dim3(kBlock), 0, S, pat->nstore, val, val32_.p);
  setup_pc32(pc, val);
  int total = 0;
  while (true) {
    hipLaunchKernelGGL(k_mp_residual, G, B, 0, S, n, A, val, b, x, r_.p, st_.p, part_.p);
    MS_CHECK_LAUNCH();
    hs = read_state(st_.p);
    R.res = hs.res;
    if (hs.done || total >= maxit) break;
    int imax = maxit - total;
    hipLaunchKernelGGL(k_mp_scale_down, G, B, 0, S, n, r_.p, st_.p, r32_.p, x32_.p, st32_.p, (float)inner_tol, imax);
    SolveResult in = inner32(solver, pc, inner_tol, imax);
    total += std::max(in.it, 1);
    R.outer++;
    hipLaunchKernelGGL(k_mp_correct, G, B, 0, S, n, x32_.p, st_.p, x);
    if (in.ierr == 2 && in.it <= 1) {   // inner breakdown without progress
      R.ierr = 2;
      break;
    }
  }
  R.it = total;
  if (R.ierr == 0 && !(R.res <= tol)) R.ierr = 1;
  return R;
}

}  // namespace ms
