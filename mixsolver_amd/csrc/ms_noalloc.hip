// ms_noalloc.hip -- PCG with Jacobi preconditioning on a caller CSR that
// allocates nothing (ms_cg_jacobi_noalloc).
//
// Same argument meaning, recurrences and it/res/ierr conventions as
// ms_cg_jacobi (cg_jacobi, mod_uns_linsolver.f90:230-291, and its
// one-partition twin): res = sqrt(r.r)/||b||, z = r/diag, breakdown when
// |p.Ap| or |r.z| falls below tiny, it = maxit+1 with ierr = 1 on exhaustion,
// x = 0 and it = 0 when ||b|| = 0.  A row without a diagonal entry is an
// argument error (MS_EINVAL), as for ms_cg_jacobi.
//
// Every device byte this path touches beyond what the driver itself holds
// lives in `scratch`, one file-scope device array per GPU: the staged CSR
// (1-based, as the caller passes it), the right-hand side, the iterate, the
// Krylov vectors, the reduction partials and the solver scalars.  Its
// address is taken with hipGetSymbolAddress once per device when the
// library is initialised on it (ms_init), and only that address reaches
// kernels and copies.  Calls on one GPU share the array: calls that overlap
// on one GPU race, and each call overwrites what the last one left.
#include <algorithm>
#include <mutex>

#include "ms_reduce.hpp"

namespace ms {

namespace {

constexpr int kNaMaxRows = 131072;          // largest n
constexpr long long kNaMaxNnz = 1048576;    // largest nnz
constexpr int kNaMaxBlocks = (kNaMaxRows + kBlock - 1) / kBlock;
constexpr int kNaChunk = 8;                 // iterations queued between state reads
constexpr int kNaMaxDevices = 64;

struct NaState {
  double bnrm, rz, pap, alpha, beta, res, tol;
  int it, ierr, done, maxit;   // done: 0 running, 1 finished, 3 ||b|| = 0
};

constexpr size_t al256(size_t x) { return (x + 255) & ~size_t(255); }
constexpr size_t kOffRp = 0;
constexpr size_t kOffCi = al256(kOffRp + sizeof(int) * (kNaMaxRows + 1));
constexpr size_t kOffVal = al256(kOffCi + sizeof(int) * kNaMaxNnz);
constexpr size_t kOffVec = al256(kOffVal + sizeof(double) * kNaMaxNnz);   // b x r z p q diag
constexpr size_t kVecBytes = al256(sizeof(double) * kNaMaxRows);
constexpr size_t kOffPart = kOffVec + 7 * kVecBytes;
constexpr size_t kOffState = al256(kOffPart + sizeof(double) * 2 * kNaMaxBlocks);
constexpr size_t kScratchBytes = al256(kOffState + sizeof(NaState));

struct NaBufs {
  const int *rp, *ci;   // 1-based CSR as passed by the caller
  const double *val, *b;
  double *x, *r, *z, *p, *q, *diag, *part;
  NaState *st;
  int n;
};

}  // namespace

// The scratch array (about 20 MB: 131072 rows, 1048576 nonzeros).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_na_mu;
void *g_na_ptr[kNaMaxDevices] = {};

NaBufs bufs_at(unsigned char *base, int n) {
  NaBufs B;
  B.rp = reinterpret_cast<const int *>(base + kOffRp);
  B.ci = reinterpret_cast<const int *>(base + kOffCi);
  B.val = reinterpret_cast<const double *>(base + kOffVal);
  double *v = reinterpret_cast<double *>(base + kOffVec);
  const size_t s = kVecBytes / sizeof(double);
  B.b = v;
  B.x = v + s;
  B.r = v + 2 * s;
  B.z = v + 3 * s;
  B.p = v + 4 * s;
  B.q = v + 5 * s;
  B.diag = v + 6 * s;
  B.part = reinterpret_cast<double *>(base + kOffPart);
  B.st = reinterpret_cast<NaState *>(base + kOffState);
  B.n = n;
  return B;
}

__device__ __forceinline__ double row_dot(const NaBufs &B, int i, const double *v) {
  double s = 0.0;
  for (int k = B.rp[i] - 1; k < B.rp[i + 1] - 1; k++) s += B.val[k] * v[B.ci[k] - 1];
  return s;
}

// diagonal of every row (present: the host checked) and b.b
__global__ __launch_bounds__(kBlock) void k_na_setup(NaBufs B) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (i < B.n) {
    double d = 0.0;
    for (int k = B.rp[i] - 1; k < B.rp[i + 1] - 1; k++)
      if (B.ci[k] == i + 1) { d = B.val[k]; break; }
    B.diag[i] = d;
    const double bi = B.b[i];
    v[0] = bi * bi;
  }
  block_partials<OpSum, 2>(v, B.part);
}

// r = b - A x, r.r
__global__ __launch_bounds__(kBlock) void k_na_residual(NaBufs B) {
  if (B.st->done) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (i < B.n) {
    const double ri = B.b[i] - row_dot(B, i, B.x);
    B.r[i] = ri;
    v[0] = ri * ri;
  }
  block_partials<OpSum, 2>(v, B.part);
}

// z = r / diag, p = z, r.z
__global__ __launch_bounds__(kBlock) void k_na_start(NaBufs B) {
  if (B.st->done) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (i < B.n) {
    const double ri = B.r[i], zi = ri / B.diag[i];
    B.z[i] = zi;
    B.p[i] = zi;
    v[0] = ri * zi;
  }
  block_partials<OpSum, 2>(v, B.part);
}

// q = A p, p.q
__global__ __launch_bounds__(kBlock) void k_na_spmv(NaBufs B) {
  if (B.st->done) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (i < B.n) {
    const double pi = B.p[i], qi = row_dot(B, i, B.p);
    B.q[i] = qi;
    v[0] = pi * qi;
  }
  block_partials<OpSum, 2>(v, B.part);
}

// x += alpha p, r -= alpha q, z = r / diag; r.r and r.z
__global__ __launch_bounds__(kBlock) void k_na_update(NaBufs B) {
  if (B.st->done) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (i < B.n) {
    const double a = B.st->alpha, pi = B.p[i], qi = B.q[i], xi = B.x[i], r0 = B.r[i];
    B.x[i] = xi + a * pi;
    const double ri = r0 - a * qi, zi = ri / B.diag[i];
    B.r[i] = ri;
    B.z[i] = zi;
    v[0] = ri * ri;
    v[1] = ri * zi;
  }
  block_partials<OpSum, 2>(v, B.part);
}

// p = z + beta p
__global__ __launch_bounds__(kBlock) void k_na_pupd(NaBufs B) {
  if (B.st->done) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < B.n) B.p[i] = B.z[i] + B.st->beta * B.p[i];
}

enum NaStage { NA_SETUP = 0, NA_RESIDUAL = 1, NA_START = 2, NA_PAP = 3, NA_STEP = 4 };

// second stage of the reductions plus the solver's scalar recurrences
__global__ __launch_bounds__(kFinThreads) void k_na_finish(int nparts, NaBufs B, int stage) {
  NaState *st = B.st;
  if (stage != NA_SETUP && st->done) return;
  double acc[2];
  fin_gather<OpSum, 2, double>(nparts, B.part, acc);
  __shared__ double lds[2 * (kFinThreads / 64)];
  block_reduce<OpSum, 2, double, kFinThreads>(acc, lds);
  if (threadIdx.x != 0) return;
  const double tiny = 2.2250738585072014e-308;
  switch (stage) {
    case NA_SETUP:
      st->bnrm = sqrt(acc[0]);
      if (st->bnrm == 0.0) st->done = 3;
      break;
    case NA_RESIDUAL:
      st->res = sqrt(acc[0]) / st->bnrm;
      if (st->res <= st->tol) st->done = 1;
      break;
    case NA_START:
      st->rz = acc[0];
      break;
    case NA_PAP:
      st->it += 1;
      st->pap = acc[0];
      if (fabs(st->pap) < tiny) { st->ierr = 2; st->done = 1; break; }
      st->alpha = st->rz / st->pap;
      break;
    default:   // NA_STEP: acc = (r.r, r.z) after the update
      st->res = sqrt(acc[0]) / st->bnrm;
      if (st->res <= st->tol) { st->done = 1; break; }
      if (fabs(st->rz) < tiny) { st->ierr = 2; st->done = 1; break; }
      st->beta = acc[1] / st->rz;
      st->rz = acc[1];
      break;
  }
}

void *scratch_of_current_device() {
  int dev = -1;
  MS_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> g(g_na_mu);
  if (dev < 0 || dev >= kNaMaxDevices || !g_na_ptr[dev])
    fail(MS_EINVAL, "ms_cg_jacobi_noalloc: call ms_init on this device first");
  return g_na_ptr[dev];
}

}  // namespace

// Called by the library's initialisation with `device` current: takes the
// device address of this GPU's scratch array once.
void noalloc_register_device(int device) {
  std::lock_guard<std::mutex> g(g_na_mu);
  if (device < 0 || device >= kNaMaxDevices) fail(MS_EINVAL, "device index beyond the scratch table");
  if (g_na_ptr[device]) return;
  void *ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    fail(MS_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_na_ptr[device] = ptr;
}

void cg_jacobi_noalloc(int n, const int *row_ptr, const int *col_idx, const double *val, const double *b, double *x,
                       double tol, int maxit, int *it, double *res, int *ierr) {
  if (n <= 0 || !row_ptr || !col_idx || !val || !b || !x) fail(MS_EINVAL, "ms_cg_jacobi_noalloc: bad arguments");
  if (n > kNaMaxRows)
    fail(MS_EINVAL, "ms_cg_jacobi_noalloc: n = " + std::to_string(n) + " exceeds the scratch limit of " +
                        std::to_string(kNaMaxRows) + " rows");
  if (row_ptr[0] != 1) fail(MS_EINVAL, "ms_cg_jacobi_noalloc: row_ptr(1) must be 1");
  for (int i = 0; i < n; i++)
    if (row_ptr[i + 1] < row_ptr[i]) fail(MS_EINVAL, "ms_cg_jacobi_noalloc: row_ptr must not decrease");
  const long long nnz = (long long)row_ptr[n] - 1;
  if (nnz > kNaMaxNnz)
    fail(MS_EINVAL, "ms_cg_jacobi_noalloc: nnz = " + std::to_string(nnz) + " exceeds the scratch limit of " +
                        std::to_string(kNaMaxNnz) + " nonzeros");
  for (int i = 0; i < n; i++) {
    bool diag = false;
    for (int k = row_ptr[i] - 1; k < row_ptr[i + 1] - 1; k++) {
      if (col_idx[k] < 1 || col_idx[k] > n) fail(MS_EINVAL, "ms_cg_jacobi_noalloc: column index out of range");
      if (k > row_ptr[i] - 1 && col_idx[k] <= col_idx[k - 1])
        fail(MS_EINVAL, "ms_cg_jacobi_noalloc: CSR columns must be strictly ascending per row");
      diag = diag || col_idx[k] == i + 1;
    }
    if (!diag) fail(MS_EINVAL, "ms_cg_jacobi_noalloc: matrix row without diagonal entry");
  }
  unsigned char *base = static_cast<unsigned char *>(scratch_of_current_device());
  NaBufs B = bufs_at(base, n);
  hipStream_t S = stream();
  const int nb = nblocks(n);
  NaState h{};
  h.tol = tol;
  h.maxit = maxit;
  MS_HIP(hipMemcpyAsync(const_cast<int *>(B.rp), row_ptr, sizeof(int) * (n + 1), hipMemcpyHostToDevice, S));
  if (nnz > 0) {
    MS_HIP(hipMemcpyAsync(const_cast<int *>(B.ci), col_idx, sizeof(int) * nnz, hipMemcpyHostToDevice, S));
    MS_HIP(hipMemcpyAsync(const_cast<double *>(B.val), val, sizeof(double) * nnz, hipMemcpyHostToDevice, S));
  }
  MS_HIP(hipMemcpyAsync(const_cast<double *>(B.b), b, sizeof(double) * n, hipMemcpyHostToDevice, S));
  MS_HIP(hipMemcpyAsync(B.x, x, sizeof(double) * n, hipMemcpyHostToDevice, S));
  MS_HIP(hipMemcpyAsync(B.st, &h, sizeof(h), hipMemcpyHostToDevice, S));
  hipLaunchKernelGGL(k_na_setup, dim3(nb), dim3(kBlock), 0, S, B);
  hipLaunchKernelGGL(k_na_finish, dim3(1), dim3(kFinThreads), 0, S, nb, B, (int)NA_SETUP);
  hipLaunchKernelGGL(k_na_residual, dim3(nb), dim3(kBlock), 0, S, B);
  hipLaunchKernelGGL(k_na_finish, dim3(1), dim3(kFinThreads), 0, S, nb, B, (int)NA_RESIDUAL);
  hipLaunchKernelGGL(k_na_start, dim3(nb), dim3(kBlock), 0, S, B);
  hipLaunchKernelGGL(k_na_finish, dim3(1), dim3(kFinThreads), 0, S, nb, B, (int)NA_START);
  MS_CHECK_LAUNCH();
  MS_HIP(hipMemcpyAsync(&h, B.st, sizeof(h), hipMemcpyDeviceToHost, S));
  MS_HIP(hipStreamSynchronize(S));
  int enq = 0;
  while (!h.done && enq < maxit) {
    const int m = std::min(kNaChunk, maxit - enq);
    for (int k = 0; k < m; k++) {
      hipLaunchKernelGGL(k_na_spmv, dim3(nb), dim3(kBlock), 0, S, B);
      hipLaunchKernelGGL(k_na_finish, dim3(1), dim3(kFinThreads), 0, S, nb, B, (int)NA_PAP);
      hipLaunchKernelGGL(k_na_update, dim3(nb), dim3(kBlock), 0, S, B);
      hipLaunchKernelGGL(k_na_finish, dim3(1), dim3(kFinThreads), 0, S, nb, B, (int)NA_STEP);
      hipLaunchKernelGGL(k_na_pupd, dim3(nb), dim3(kBlock), 0, S, B);
    }
    MS_CHECK_LAUNCH();
    enq += m;
    MS_HIP(hipMemcpyAsync(&h, B.st, sizeof(h), hipMemcpyDeviceToHost, S));
    MS_HIP(hipStreamSynchronize(S));
  }
  if (h.done == 3) {   // ||b|| = 0: x = 0, it = 0, res = 0
    for (int i = 0; i < n; i++) x[i] = 0.0;
    *it = 0;
    *res = 0.0;
    *ierr = 0;
    return;
  }
  MS_HIP(hipMemcpyAsync(x, B.x, sizeof(double) * n, hipMemcpyDeviceToHost, S));
  MS_HIP(hipStreamSynchronize(S));
  if (!h.done) {   // the DO loop ran out: it = maxit + 1 (mod_uns_linsolver_mpi.f90:297)
    h.it = std::max(maxit, 0) + 1;
    h.ierr = 1;
  }
  *it = h.it;
  *res = h.res;
  *ierr = h.ierr;
}

}  // namespace ms
