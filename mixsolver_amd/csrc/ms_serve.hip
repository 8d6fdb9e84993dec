// ms_serve.hip -- continuous-batching server for pressure-correction solves.
//
// Requests are SPD solves A x = b (x0 = 0) on registered sparsity patterns
// (one pattern per mesh).  A pattern keeps its device layout resident: the
// colour-blocked numbering, the SELL-64 pattern, the aggregate chunks and the
// slot map from the caller's CSR order.  A request brings only values and a
// right-hand side, which are scattered into the device layout by kernels.
//
// Every active request is packed into every launch.  A work item is one chunk
// (at most kBlock rows of one colour and one aggregate) of one request; the
// per-request state (vector pointers, FP32/FP64 matrices, solver scalars) is
// looked up from the item.  Each request's reductions are finished by its own
// finisher workgroup over its own contiguous item range, in a fixed order, so
// a request's iterates do not depend on which other requests share the
// launches: a result is the same alone and in a mix, bit for bit.
//
// Set-up of admitted requests runs on its own stream and overlaps the
// rounds; a request joins the iterating set at the first round boundary
// after its set-up event fired.  Values and right-hand sides are copied into
// pinned memory at submission; device buffers of finished requests and the
// pinned stages are pooled per pattern.
//
// The iteration is the mixed-precision PCG of ms_linalg.hip (FP32 matrix in
// difference form, FP32 preconditioner and vectors on r/sigma, FP64 solution
// and reliable updates) with Jacobi or multicolour DILU and the optional
// aggregate coarse correction.  A worker thread runs rounds of a fixed number
// of iterations; between rounds it applies reliable updates, retires finished
// requests and admits waiting ones while the device-memory budget and the
// slot count allow (continuous batching).  All device work of the server is
// issued by the worker thread.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <atomic>
#include <map>
#include <mutex>
#include <numeric>
#include <thread>

#include "ms_kernels.hpp"
#include "ms_serve.hpp"

namespace ms {
namespace {

struct Item {
  int sys, row0, row1, agg;
  int slot;   // index of this chunk's partial in the list it belongs to
  int fin;    // DILU sweep lists: 1 when this sweep gives the rows their final z
};
struct FinRange {
  int sys, start, end;   // partials [start, end) of request slot `sys`
};

struct BSys {
  Mat A;
  const float *v32;
  const double *v64;
  const float *dg;
  float *r, *z, *p, *q, *y;
  double *x, *b, *r64;
  KS *st;
  int c0end;
  int k;
  const int *agg_cptr, *agg_cidx;
  const float *Einv;
  double *nu;
};

struct FinStart0 {   // ||b||; x0 = 0 so r = b and sigma = ||b||
  KS *st;
  __device__ bool skip() const { return false; }
  template <class T> __device__ void operator()(const T *a) const {
    st->rr = a[0];
    st->bnrm = sqrt((double)a[0]);
    st->sigma = st->bnrm;
    st->rmax = 1.0;
    st->res = 1.0;
    if (st->bnrm == 0.0) {   // x = 0 is the solution
      st->res = 0.0;
      st->done = 1;
    }
  }
};
struct FinStartJ {   // initial r.z of the Jacobi preconditioner (with the coarse part)
  KS *st;
  int jac;
  __device__ bool skip() const { return st->done != 0; }
  template <class T> __device__ void operator()(const T *a) const {
    if (jac) st->rz = rz_total(st, (double)a[1]);
  }
};

template <class F> __device__ F mk(KS *st, int a0, int a1);
template <> __device__ FinPap mk<FinPap>(KS *st, int, int) { return FinPap{st}; }
template <> __device__ FinUpdateRU mk<FinUpdateRU>(KS *st, int a0, int a1) { return FinUpdateRU{st, a0, a1}; }
template <> __device__ FinRz mk<FinRz>(KS *st, int, int) { return FinRz{st}; }
template <> __device__ FinRz0 mk<FinRz0>(KS *st, int, int) { return FinRz0{st}; }
template <> __device__ FinRU mk<FinRU>(KS *st, int, int) { return FinRU{st}; }
template <> __device__ FinStart0 mk<FinStart0>(KS *st, int, int) { return FinStart0{st}; }
template <> __device__ FinStartJ mk<FinStartJ>(KS *st, int a0, int) { return FinStartJ{st, a0}; }

// ---------------------------------------------------------- request set-up
__global__ void k_s_scatter(long long nnz, const int *slot_of, const double *val, double *v64) {
  long long k = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (k < nnz) v64[slot_of[k]] = val[k];
}
__global__ void k_s_perm(int n, const int *o2n, const double *src, double *dst) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) dst[o2n[i]] = src[i];
}
__global__ void k_s_unperm(int n, const int *o2n, const double *src, double *dst) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) dst[i] = src[o2n[i]];
}

// ------------------------------------------------------ batched iteration
__global__ void k_b_sumsq(const Item *L, const BSys *S, double *part) {
  const int ii = xcd_block(blockIdx.x, gridDim.x);   // item of this workgroup
  const Item it = L[ii];
  const BSys &s = S[it.sys];
  const int i = it.row0 + (int)threadIdx.x;
  double v[1] = {0.0};
  if (i < it.row1) v[0] = s.b[i] * s.b[i];
  block_partials<OpSum, 1>(v, part, ii);
}

// x = y = p = 0, r = b / sigma; partials (r.r, Jacobi r.(r/d), sum r); DILU: z = r/d on colour 0
__global__ void k_b_init(const Item *L, const BSys *S, int mode, float *part) {
  const int ii = xcd_block(blockIdx.x, gridDim.x);   // item of this workgroup
  const Item it = L[ii];
  const BSys &s = S[it.sys];
  if (s.st->done) return;
  const int i = it.row0 + (int)threadIdx.x;
  float v[3] = {0.0f, 0.0f, 0.0f};
  if (i < it.row1) {
    s.y[i] = 0.0f;
    s.p[i] = 0.0f;
    const float ri = (float)(s.b[i] / s.st->sigma);
    s.r[i] = ri;
    v[0] = ri * ri;
    v[2] = ri;
    if (mode == 1) v[1] = ri * (ri / s.dg[i]);
    else if (mode == 2 && i < s.c0end) s.z[i] = ri / s.dg[i];
  }
  block_partials<OpSum, 3>(v, part, ii);
}

__global__ void k_b_spmv(const Item *L, const BSys *S, float *part) {
  const int ii = xcd_block(blockIdx.x, gridDim.x);   // item of this workgroup
  const Item it = L[ii];
  const BSys &s = S[it.sys];
  if (s.st->done) return;
  const int i = it.row0 + (int)threadIdx.x;
  float v[1] = {0.0f};
  if (i < it.row1) {
    const float qi = row_apply(i, s.A, s.v32, s.p);
    s.q[i] = qi;
    v[0] = s.p[i] * qi;
  }
  block_partials<OpSum, 1>(v, part, ii);
}

__global__ void k_b_update(const Item *L, const BSys *S, int mode, float *part) {
  const int ii = xcd_block(blockIdx.x, gridDim.x);   // item of this workgroup
  const Item it = L[ii];
  const BSys &s = S[it.sys];
  if (s.st->done) return;
  const int i = it.row0 + (int)threadIdx.x;
  float v[3] = {0.0f, 0.0f, 0.0f};
  if (i < it.row1) {
    const float alpha = (float)s.st->alpha;
    s.y[i] = s.y[i] + alpha * s.p[i];
    const float ri = s.r[i] - alpha * s.q[i];
    s.r[i] = ri;
    v[0] = ri * ri;
    v[2] = ri;
    if (mode == 1) v[1] = ri * (ri / s.dg[i]);
    else if (mode == 2 && i < s.c0end) s.z[i] = ri / s.dg[i];
  }
  block_partials<OpSum, 3>(v, part, ii);
}

// p = z (Jacobi: r/d) + nu of the chunk's aggregate + beta p
__global__ void k_b_p(const Item *L, const BSys *S, int jac) {
  const int ii = xcd_block(blockIdx.x, gridDim.x);   // item of this workgroup
  const Item it = L[ii];
  const BSys &s = S[it.sys];
  if (s.st->done) return;
  const int i = it.row0 + (int)threadIdx.x;
  if (i < it.row1) {
    float zi = jac ? s.r[i] / s.dg[i] : s.z[i];
    if (s.k > 0) zi = zi + (float)s.nu[it.agg];
    s.p[i] = zi + (float)s.st->beta * s.p[i];
  }
}

// DILU sweeps over the chunks of one colour; r.z partial at prz[slot] for the
// chunks whose z is final after this sweep
__global__ void k_b_fwd(const Item *L, const BSys *S, float *prz) {
  const int ii = xcd_block(blockIdx.x, gridDim.x);   // item of this workgroup
  const Item it = L[ii];
  const BSys &s = S[it.sys];
  if (s.st->done) return;
  const int i = it.row0 + (int)threadIdx.x;
  float v[1] = {0.0f};
  if (i < it.row1) {
    const long long bi = s.A.off[i >> 6] + (i & 63);
    const int dpi = s.A.diagk[i];
    float acc = s.r[i];
    for (int k = 0; k < dpi; k++) {
      const long long q = bi + (long long)k * kSlice;
      acc = acc - s.v32[q] * s.z[col_f32(s.A, q, i, i >> 6)];
    }
    const float zi = acc / s.dg[i];
    s.z[i] = zi;
    v[0] = s.r[i] * zi;
  }
  __shared__ float lds[kBlock / 64];
  block_reduce<OpSum, 1, float>(v, lds);
  if (it.fin && threadIdx.x == 0) prz[it.slot] = v[0];
}

__global__ void k_b_bwd(const Item *L, const BSys *S, float *prz) {
  const int ii = xcd_block(blockIdx.x, gridDim.x);   // item of this workgroup
  const Item it = L[ii];
  const BSys &s = S[it.sys];
  if (s.st->done) return;
  const int i = it.row0 + (int)threadIdx.x;
  float v[1] = {0.0f};
  if (i < it.row1) {
    const long long bi = s.A.off[i >> 6] + (i & 63);
    const int dpi = s.A.diagk[i], len = s.A.rowlen[i];
    float acc = 0.0f;
    for (int k = dpi + 1; k < len; k++) {
      const long long q = bi + (long long)k * kSlice;
      acc = acc + s.v32[q] * s.z[col_f32(s.A, q, i, i >> 6)];
    }
    const float zi = s.z[i] - acc / s.dg[i];
    s.z[i] = zi;
    v[0] = s.r[i] * zi;
  }
  __shared__ float lds[kBlock / 64];
  block_reduce<OpSum, 1, float>(v, lds);
  if (threadIdx.x == 0) prz[it.slot] = v[0];
}

// reliable update: x += sigma y, y = 0; r64 = b - A64 x with its r.r partials;
// then r = r64 / sigma_new, p *= scale with the partials of k_b_update
__global__ void k_b_fold(const Item *L, const BSys *S) {
  const int ii = xcd_block(blockIdx.x, gridDim.x);   // item of this workgroup
  const Item it = L[ii];
  const BSys &s = S[it.sys];
  if (s.st->done == 1 || s.st->ru != 1) return;
  const int i = it.row0 + (int)threadIdx.x;
  if (i < it.row1) {
    s.x[i] = s.x[i] + s.st->sigma * (double)s.y[i];
    s.y[i] = 0.0f;
  }
}
__global__ void k_b_true_res(const Item *L, const BSys *S, double *part) {
  const int ii = xcd_block(blockIdx.x, gridDim.x);   // item of this workgroup
  const Item it = L[ii];
  const BSys &s = S[it.sys];
  if (s.st->done == 1 || s.st->ru != 1) return;
  const int i = it.row0 + (int)threadIdx.x;
  double v[1] = {0.0};
  if (i < it.row1) {
    const double ri = s.b[i] - row_dot(i, s.A, s.v64, s.x);
    s.r64[i] = ri;
    v[0] = ri * ri;
  }
  block_partials<OpSum, 1>(v, part, ii);
}
__global__ void k_b_rescale(const Item *L, const BSys *S, int mode, float *part) {
  const int ii = xcd_block(blockIdx.x, gridDim.x);   // item of this workgroup
  const Item it = L[ii];
  const BSys &s = S[it.sys];
  if (s.st->done || s.st->ru != 2) return;
  const int i = it.row0 + (int)threadIdx.x;
  float v[3] = {0.0f, 0.0f, 0.0f};
  if (i < it.row1) {
    const float ri = (float)(s.r64[i] / s.st->sigma);
    s.r[i] = ri;
    s.p[i] = s.p[i] * (float)s.st->scale;
    v[0] = ri * ri;
    v[2] = ri;
    if (mode == 1) v[1] = ri * (ri / s.dg[i]);
    else if (mode == 2 && i < s.c0end) s.z[i] = ri / s.dg[i];
  }
  block_partials<OpSum, 3>(v, part, ii);
}

// one finisher workgroup per request: its partials [start, end) in a fixed
// order, the coarse solve when requested, then the functor on thread 0
template <int NV, class T, class Fin>
__global__ __launch_bounds__(kFinThreads) void k_b_fin(const FinRange *R, const T *part, const BSys *S, int coarse,
                                                      int a0, int a1) {
  const FinRange fr = R[blockIdx.x];
  const BSys &s = S[fr.sys];
  Fin fin = mk<Fin>(s.st, a0, a1);
  if (fin.skip()) return;
  T acc[NV];
  fin_gather<OpSum, NV, T>(fr.end - fr.start, part + (size_t)fr.start * NV, acc);
  __shared__ T lds[NV * (kFinThreads / 64)];
  block_reduce<OpSum, NV, T, kFinThreads>(acc, lds);
  if (coarse && s.k > 0)
    coarse_solve<NV, T>(part + (size_t)fr.start * NV, CoDev{s.k, s.agg_cptr, s.agg_cidx, s.Einv, s.nu, s.st});
  if (threadIdx.x == 0) fin(acc);
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

// ================================================================ host side
struct Slot;

// A request's values and right-hand side, copied at submission into pinned
// host memory (val at [0, nnz), b at [nnz, nnz + n)).
struct Stage {
  double *p = nullptr;
  size_t n = 0;
};

struct Pattern {
  int n = 0;
  long long nnz = 0;
  int ncol = 1, k = 0;
  SellPattern pat;             // device numbering; pat.coarse holds the chunks
  std::vector<int> o2n, slot_of, col_of_chunk, agg_dev;
  DBuf<int> d_o2n, d_slot_of;
  bool uploaded = false;
  // device buffers of retired requests and pinned stages, reused by the
  // next requests on this pattern
  std::vector<std::unique_ptr<Slot>> pool;
  std::vector<Stage> stages;
};

struct Req {
  long long ticket = 0;
  Pattern *P = nullptr;
  Stage stage;
  double tol = 0.0;
  int maxit = 0;
  double t_submit = 0.0, t_admit = -1.0, t_done = -1.0;
  std::vector<double> x;
  int it = 0, ierr = 0;
  double res = 0.0;
  bool done = false;
};

struct Slot {
  Req *req = nullptr;
  Pattern *P = nullptr;
  int id = 0;
  size_t bytes = 0;
  DBuf<double> v64, x64, b64, r64, tmp, E, nu;
  DBuf<float> v32, dg, r, z, p, q, y, Einv;
  BSys hsys{};
  KS hks{};
  // stall guard: the best residual so far (halvings only) and its iteration
  double best = HUGE_VAL;
  int it_best = 0;
  bool stalled = false;
};

// requests whose set-up runs on the set-up stream; ready when `done` fires
struct Setup {
  std::vector<Slot *> slots;
  hipEvent_t done = nullptr;
};

struct List {
  std::vector<Item> items;
  std::vector<FinRange> ranges;
  std::vector<std::vector<Item>> fwd, bwd;   // per colour
  DBuf<Item> d_items;
  DBuf<FinRange> d_ranges;
  std::vector<DBuf<Item>> d_fwd, d_bwd;
  int maxcol = 0;

  // whole requests, partial slots numbered in list order
  void build(const std::vector<Slot *> &slots) {
    items.clear();
    ranges.clear();
    for (Slot *sl : slots) {
      const Coarse &C = sl->P->pat.coarse;
      const int start = (int)items.size();
      for (int c = 0; c < C.nchunks; c++) {
        Item it{sl->id, C.chunk_ptr[c], C.chunk_ptr[c + 1], C.chunk_agg[c], (int)items.size(), 0};
        items.push_back(it);
      }
      ranges.push_back(FinRange{sl->id, start, (int)items.size()});
    }
    colour_lists(slots, nullptr);
    upload();
  }
  // the requests `keep` of list L, with L's partial slots and ranges
  void subset(const List &L, const std::vector<Slot *> &slots, const std::vector<Slot *> &keep) {
    items.clear();
    ranges.clear();
    for (size_t j = 0; j < slots.size(); j++) {
      if (std::find(keep.begin(), keep.end(), slots[j]) == keep.end()) continue;
      const FinRange fr = L.ranges[j];
      ranges.push_back(fr);
      for (int t = fr.start; t < fr.end; t++) items.push_back(L.items[t]);
    }
    colour_lists(keep, &L);
    upload();
  }

 private:
  void colour_lists(const std::vector<Slot *> &slots, const List *) {
    maxcol = 0;
    for (Slot *sl : slots) maxcol = std::max(maxcol, sl->P->ncol);
    fwd.assign(maxcol, {});
    bwd.assign(maxcol, {});
    std::map<int, Slot *> by_id;
    for (Slot *sl : slots) by_id[sl->id] = sl;
    for (const Item &it : items) {
      const Pattern *P = by_id[it.sys]->P;
      const int chunk = chunk_index(P, it.row0);
      const int c = P->col_of_chunk[chunk];
      if (c >= 1) {
        Item f = it;
        f.fin = c == P->ncol - 1;
        fwd[c].push_back(f);
      }
      if (c <= std::max(P->ncol - 2, 0)) {
        Item g = it;
        g.fin = 1;
        bwd[c].push_back(g);
      }
    }
  }
  static int chunk_index(const Pattern *P, int row0) {
    const auto &cp = P->pat.coarse.chunk_ptr;
    return (int)(std::upper_bound(cp.begin(), cp.end(), row0) - cp.begin()) - 1;
  }
  void upload() {
    if (!items.empty()) d_items.upload(items);
    if (!ranges.empty()) d_ranges.upload(ranges);
    d_fwd.resize(maxcol);
    d_bwd.resize(maxcol);
    for (int c = 0; c < maxcol; c++) {
      if (!fwd[c].empty()) d_fwd[c].upload(fwd[c]);
      if (!bwd[c].empty()) d_bwd[c].upload(bwd[c]);
    }
  }
};

class Server {
 public:
  explicit Server(const ServeOpts &o) : o_(o) {
    ensure_runtime(-1);
    MS_HIP(hipGetDevice(&device_));
    if (o_.max_active < 1) o_.max_active = 1;
    if (o_.round_iters < 1) o_.round_iters = 16;
    if (o_.precond != MS_PC_JACOBI && o_.precond != MS_PC_DILU)
      fail(MS_EINVAL, "server preconditioner must be MS_PC_JACOBI or MS_PC_DILU");
    if (o_.coarse < 0 || o_.coarse > kInvRegMax) fail(MS_EINVAL, "server coarse aggregates must be 0..256");
    if (o_.budget <= 0) {
      size_t fr = 0, tot = 0;
      MS_HIP(hipMemGetInfo(&fr, &tot));
      o_.budget = (long long)(0.5 * (double)fr);
    }
    d_ks_.alloc(o_.max_active);
    d_sys_.alloc(o_.max_active);
    MS_HIP(hipHostMalloc((void **)&h_ks_, sizeof(KS) * o_.max_active, hipHostMallocDefault));
    // request set-up (uploads, FP32 copy, preconditioner, coarse inverse)
    // overlaps the rounds on a second stream of the default priority
    MS_HIP(hipStreamCreateWithFlags(&setup_stream_, hipStreamNonBlocking));
    for (int i = o_.max_active - 1; i >= 0; i--) free_ids_.push_back(i);
    worker_ = std::thread([this] { loop(); });
  }

  ~Server() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_work_.notify_all();
    if (worker_.joinable()) worker_.join();
    if (round_exec_) (void)hipGraphExecDestroy(round_exec_);
    (void)hipStreamSynchronize(setup_stream_);
    for (Setup &su : setups_) (void)hipEventDestroy(su.done);
    (void)hipStreamDestroy(setup_stream_);
    for (auto &P : patterns_)
      for (Stage &st : P->stages) (void)hipHostFree(st.p);
    for (auto &kv : reqs_)
      if (kv.second->stage.p) (void)hipHostFree(kv.second->stage.p);
    if (h_ks_) (void)hipHostFree(h_ks_);
  }

  int add_pattern(int n, const int *rp1, const int *ci1, const double *xc) {
    if (n <= 0 || !rp1 || !ci1) fail(MS_EINVAL, "server pattern: empty matrix");
    std::unique_ptr<Pattern> P(new Pattern());
    P->n = n;
    P->nnz = rp1[n] - rp1[0];
    std::vector<int> rp(n + 1), ci(P->nnz);
    for (int i = 0; i <= n; i++) rp[i] = rp1[i] - rp1[0];
    for (long long k = 0; k < P->nnz; k++) {
      ci[k] = ci1[k] - 1;
      if (ci[k] < 0 || ci[k] >= n) fail(MS_EINVAL, "server pattern: column index out of range");
    }
    for (int i = 0; i < n; i++)
      for (int k = rp[i] + 1; k < rp[i + 1]; k++)
        if (ci[k] <= ci[k - 1]) fail(MS_EINVAL, "server pattern: columns must be strictly ascending per row");
    // cell graph, colouring and aggregates in the caller's numbering
    std::vector<int> adjp(n + 1, 0), adj;
    adj.reserve(P->nnz);
    for (int i = 0; i < n; i++) {
      for (int k = rp[i]; k < rp[i + 1]; k++)
        if (ci[k] != i) adj.push_back(ci[k]);
      adjp[i + 1] = (int)adj.size();
    }
    std::vector<int> color(n, 0), visit(n);
    std::iota(visit.begin(), visit.end(), 0);
    P->ncol = 1;
    if (o_.precond == MS_PC_DILU) color = greedy_colors(n, adjp, adj, visit, P->ncol);
    std::vector<int> agg(n, 0);
    if (o_.coarse > 0 && xc) {
      P->k = std::max(1, std::min(o_.coarse, n / 4096));   // about 4096 cells per aggregate, at most `coarse`
      std::vector<double> v(xc, xc + 3 * (size_t)n);
      agg = rcb_aggregates(n, v, P->k);
    }
    std::vector<int> n2o(n);
    std::iota(n2o.begin(), n2o.end(), 0);
    std::stable_sort(n2o.begin(), n2o.end(), [&](int a, int b) {
      return color[a] != color[b] ? color[a] < color[b] : agg[a] < agg[b];
    });
    P->o2n.assign(n, 0);
    for (int i = 0; i < n; i++) P->o2n[n2o[i]] = i;
    // device-numbered CSR; source position of every entry
    std::vector<int> drp(n + 1, 0), dci(P->nnz);
    std::vector<long long> src(P->nnz);
    {
      std::vector<std::pair<int, long long>> row;
      for (int r = 0; r < n; r++) {
        const int o = n2o[r];
        row.clear();
        for (int k = rp[o]; k < rp[o + 1]; k++) row.push_back({P->o2n[ci[k]], k});
        std::sort(row.begin(), row.end());
        drp[r + 1] = drp[r] + (int)row.size();
        for (size_t j = 0; j < row.size(); j++) {
          dci[drp[r] + j] = row[j].first;
          src[drp[r] + j] = row[j].second;
        }
      }
    }
    csr_to_sell(P->pat, n, drp, dci);
    if (o_.precond == MS_PC_DILU && !P->pat.symmetric)
      fail(MS_EINVAL, "server pattern: DILU needs a structurally symmetric pattern");
    P->pat.color_ptr.assign(P->ncol + 1, 0);
    for (int i = 0; i < n; i++) P->pat.color_ptr[color[i] + 1]++;
    for (int c = 0; c < P->ncol; c++) P->pat.color_ptr[c + 1] += P->pat.color_ptr[c];
    P->slot_of.assign(P->nnz, 0);
    if (P->pat.nstore > 0x7fffffffLL) fail(MS_EINVAL, "server pattern: matrix too large");
    for (int r = 0; r < n; r++)
      for (int j = 0; j < P->pat.rowlen[r]; j++) P->slot_of[src[drp[r] + j]] = (int)P->pat.slot(r, j);
    P->agg_dev.assign(n, 0);   // aggregates in the device numbering, used at upload
    for (int r = 0; r < n; r++) P->agg_dev[r] = agg[n2o[r]];
    std::lock_guard<std::mutex> lk(mu_);
    patterns_.push_back(std::move(P));
    return (int)patterns_.size() - 1;
  }

  // The values and the right-hand side are copied into pinned host memory
  // here, so the caller may reuse its arrays as soon as submit returns.
  long long submit(int pattern, const double *val, const double *b, double tol, int maxit) {
    if (!val || !b || maxit < 1 || !(tol > 0.0)) fail(MS_EINVAL, "server submit: bad arguments");
    std::unique_ptr<Req> r(new Req());
    r->tol = tol;
    r->maxit = maxit;
    r->t_submit = now_s();
    Pattern *P;
    {
      std::lock_guard<std::mutex> lk(mu_);
      if (!err_.empty()) fail(MS_EDEVICE, "server stopped: " + err_);
      if (pattern < 0 || pattern >= (int)patterns_.size()) fail(MS_EINVAL, "server submit: unknown pattern");
      P = patterns_[pattern].get();
      if (!P->stages.empty()) {
        r->stage = P->stages.back();
        P->stages.pop_back();
      }
    }
    const size_t need = (size_t)P->nnz + (size_t)P->n;
    if (!r->stage.p) {
      MS_HIP(hipHostMalloc((void **)&r->stage.p, need * sizeof(double), hipHostMallocDefault));
      r->stage.n = need;
    }
    std::memcpy(r->stage.p, val, (size_t)P->nnz * sizeof(double));
    std::memcpy(r->stage.p + P->nnz, b, (size_t)P->n * sizeof(double));
    r->P = P;
    long long t;
    {
      std::lock_guard<std::mutex> lk(mu_);
      t = r->ticket = next_ticket_++;
      pending_.push_back(r.get());
      reqs_[t] = std::move(r);
      n_sub_++;
    }
    cv_work_.notify_all();
    return t;
  }

  bool result(long long ticket, double timeout_ms, double *x, ServeResult *out) {
    std::unique_lock<std::mutex> lk(mu_);
    auto f = reqs_.find(ticket);
    if (f == reqs_.end()) fail(MS_EINVAL, "server result: unknown ticket");
    Req *r = f->second.get();
    auto ready = [&] { return r->done || !err_.empty(); };
    if (timeout_ms < 0) cv_done_.wait(lk, ready);
    else if (timeout_ms > 0)
      cv_done_.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), ready);
    if (!r->done) {
      if (!err_.empty()) fail(MS_EDEVICE, "server stopped: " + err_);
      return false;
    }
    if (x) std::copy(r->x.begin(), r->x.end(), x);
    if (out) {
      out->it = r->it;
      out->ierr = r->ierr;
      out->res = r->res;
      out->t_admit = r->t_admit - r->t_submit;
      out->t_done = r->t_done - r->t_submit;
    }
    reqs_.erase(f);
    return true;
  }

  void stats(long long *o8) {
    std::lock_guard<std::mutex> lk(mu_);
    o8[0] = n_sub_;
    o8[1] = n_done_;
    o8[2] = (long long)active_.size();
    o8[3] = (long long)pending_.size();
    o8[4] = rounds_;
    o8[5] = launches_;
    o8[6] = (long long)peak_;
    o8[7] = (long long)patterns_.size();
  }

 private:
  ServeOpts o_;
  int device_ = 0;
  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::vector<std::unique_ptr<Pattern>> patterns_;
  std::deque<Req *> pending_;
  std::map<long long, std::unique_ptr<Req>> reqs_;
  long long next_ticket_ = 1, n_sub_ = 0, n_done_ = 0;
  std::atomic<long long> rounds_{0}, launches_{0};
  hipGraphExec_t round_exec_ = nullptr;   // captured round of the current set
  long long round_launches_ = 0;          // kernel launches in that round
  int stable_rounds_ = 0;                 // rounds since the set last changed
  std::vector<std::unique_ptr<Slot>> active_;   // iterating
  std::vector<std::unique_ptr<Slot>> staged_;   // set-up in flight or done, not yet started
  std::vector<Setup> setups_;
  std::vector<int> free_ids_;
  size_t used_ = 0, pooled_ = 0, peak_ = 0;     // bytes held by requests / kept in pattern pools
  hipStream_t setup_stream_ = nullptr;
  DBuf<KS> d_ks_;
  DBuf<BSys> d_sys_;
  KS *h_ks_ = nullptr;
  DBuf<float> part32_, prz_;
  DBuf<double> part64_;
  List all_, aux_, tail_;
  bool dirty_ = true, stop_ = false;
  std::string err_;
  std::thread worker_;

  int mode() const { return o_.precond == MS_PC_JACOBI ? 1 : 2; }
  int jac() const { return o_.precond == MS_PC_JACOBI; }

  size_t request_bytes(const Pattern *P) const {
    const size_t n = P->n, ns = P->pat.nstore, k = P->k;
    return ns * 12 + n * (4 * 8 + 7 * 4) + k * k * 12 + k * 8 + (size_t)P->nnz * 8;
  }

  void upload_pattern(Pattern *P) {
    if (P->uploaded) return;
    P->pat.upload();
    build_coarse(P->pat, P->agg_dev, P->k);
    const Coarse &C = P->pat.coarse;
    P->col_of_chunk.assign(C.nchunks, 0);
    for (int c = 0; c < C.nchunks; c++) {
      int col = 0;
      while (col + 1 < (int)P->pat.color_ptr.size() - 1 && C.chunk_ptr[c] >= P->pat.color_ptr[col + 1]) col++;
      P->col_of_chunk[c] = col;
    }
    P->d_o2n.upload(P->o2n);
    P->d_slot_of.upload(P->slot_of);
    // these uploads run on the library stream, but the set-up kernels that
    // read them (scatter, FP32 copy, DILU set-up, coarse operator) run on the
    // set-up stream: complete them before any set-up is enqueued
    MS_HIP(hipStreamSynchronize(stream()));
    P->uploaded = true;
  }

  template <int NV, class T, class Fin>
  void fin(const List &L, const T *part, int coarse, int a0 = 0, int a1 = 0) {
    if (L.ranges.empty()) return;
    hipLaunchKernelGGL((k_b_fin<NV, T, Fin>), dim3((int)L.ranges.size()), dim3(kFinThreads), 0, stream(),
                       L.d_ranges.p, part, d_sys_.p, coarse, a0, a1);
    launches_++;
  }
  template <class K, class... A>
  void launch(K kern, const List &L, A... args) {
    if (L.items.empty()) return;
    hipLaunchKernelGGL(kern, dim3((int)L.items.size()), dim3(kBlock), 0, stream(), L.d_items.p, d_sys_.p, args...);
    launches_++;
  }
  void sweeps(const List &L) {
    for (int c = 1; c < L.maxcol; c++)
      if (!L.fwd[c].empty()) {
        hipLaunchKernelGGL(k_b_fwd, dim3((int)L.fwd[c].size()), dim3(kBlock), 0, stream(), L.d_fwd[c].p, d_sys_.p,
                           prz_.p);
        launches_++;
      }
    for (int c = std::max(L.maxcol - 2, 0); c >= 0; c--)
      if (!L.bwd[c].empty()) {
        hipLaunchKernelGGL(k_b_bwd, dim3((int)L.bwd[c].size()), dim3(kBlock), 0, stream(), L.d_bwd[c].p, d_sys_.p,
                           prz_.p);
        launches_++;
      }
  }
  // preconditioner, r.z and p update of an iteration
  void tail(const List &L) {
    if (!jac()) {
      sweeps(L);
      fin<1, float, FinRz>(L, prz_.p, 0);
    }
    launch(k_b_p, L, jac());
  }
  void ensure_partials(size_t nitems) {
    if (part32_.n < 3 * nitems) part32_.alloc(3 * nitems + 1024);
    if (prz_.n < nitems) prz_.alloc(nitems + 1024);
    if (part64_.n < nitems) part64_.alloc(nitems + 1024);
  }
  void read_states() {
    MS_HIP(hipMemcpyAsync(h_ks_, d_ks_.p, sizeof(KS) * o_.max_active, hipMemcpyDeviceToHost, stream()));
    MS_HIP(hipStreamSynchronize(stream()));
  }
  std::vector<Slot *> active_ptrs() {
    std::vector<Slot *> v;
    for (auto &s : active_) v.push_back(s.get());
    return v;
  }

  // A slot for a request on P: a pooled one, or new device buffers.  Called
  // with mu_ held; returns null when the budget does not allow it now.
  std::unique_ptr<Slot> take_slot(Pattern *P, bool nothing_running) {
    const size_t need = request_bytes(P);
    if (!P->pool.empty()) {
      std::unique_ptr<Slot> S = std::move(P->pool.back());
      P->pool.pop_back();
      pooled_ -= S->bytes;
      used_ += S->bytes;
      return S;
    }
    // make room by releasing pooled buffers of other patterns
    for (auto &Q : patterns_) {
      while (used_ + pooled_ + need > (size_t)o_.budget && !Q->pool.empty()) {
        pooled_ -= Q->pool.back()->bytes;
        Q->pool.pop_back();
      }
    }
    if (used_ + pooled_ + need > (size_t)o_.budget && !nothing_running) return nullptr;
    std::unique_ptr<Slot> S(new Slot());
    S->P = P;
    S->bytes = need;
    const int n = P->n;
    const long long ns = P->pat.nstore;
    S->v64.alloc(ns);
    S->v32.alloc(ns);
    S->tmp.alloc(std::max<long long>(P->nnz + n, n));
    for (DBuf<double> *d : {&S->x64, &S->b64, &S->r64}) d->alloc(n);
    for (DBuf<float> *d : {&S->dg, &S->r, &S->z, &S->p, &S->q, &S->y}) d->alloc(n);
    if (P->k > 0) {
      const long long kk = (long long)P->k * P->k;
      S->E.alloc(kk);
      S->Einv.alloc(kk);
      S->nu.alloc(P->k);
    }
    used_ += need;
    return S;
  }

  // Enqueue the set-up of the slots on the set-up stream: values and b from
  // the pinned stage into the device layout, FP32 copy, preconditioner
  // diagonal, coarse operator and its inverse, solver state.
  void begin_setup(std::vector<std::unique_ptr<Slot>> &slots) {
    hipStream_t st = setup_stream_;
    Setup su;
    for (auto &S : slots) {
      Req *r = S->req;
      Pattern *P = S->P;
      upload_pattern(P);
      const int n = P->n;
      const long long ns = P->pat.nstore;
      const Mat A = matof(&P->pat);
      MS_HIP(hipMemcpyAsync(S->tmp.p, r->stage.p, ((size_t)P->nnz + n) * sizeof(double), hipMemcpyHostToDevice,
                            st));
      hipLaunchKernelGGL(k_fill64, dim3(nblocks(ns)), dim3(kBlock), 0, st, ns, S->v64.p, 0.0);   // padding
      hipLaunchKernelGGL(k_s_scatter, dim3(nblocks(P->nnz)), dim3(kBlock), 0, st, P->nnz, P->d_slot_of.p, S->tmp.p,
                         S->v64.p);
      hipLaunchKernelGGL(k_s_perm, dim3(nblocks(n)), dim3(kBlock), 0, st, n, P->d_o2n.p, S->tmp.p + P->nnz,
                         S->b64.p);
      hipLaunchKernelGGL(k_fill64, dim3(nblocks(n)), dim3(kBlock), 0, st, (long long)n, S->x64.p, 0.0);
      hipLaunchKernelGGL(k_mat32_diff, dim3(nblocks(n)), dim3(kBlock), 0, st, n, A, S->v64.p, S->v32.p);
      if (jac()) {
        hipLaunchKernelGGL(k_extract_diag<double>, dim3(nblocks(n)), dim3(kBlock), 0, st, n, A, S->v64.p,
                           S->r64.p);
      } else {
        for (int c = 0; c < P->ncol; c++) {
          const int r0 = P->pat.color_ptr[c], r1 = P->pat.color_ptr[c + 1];
          if (r1 > r0)
            hipLaunchKernelGGL(k_dilu_setup<double>, dim3(nblocks(r1 - r0)), dim3(kBlock), 0, st, r0, r1, A,
                               P->pat.d_tpos.p, S->v64.p, S->r64.p);
        }
      }
      hipLaunchKernelGGL(k_to_float, dim3(nblocks(n)), dim3(kBlock), 0, st, (long long)n, S->r64.p, S->dg.p);
      const Coarse &C = P->pat.coarse;
      if (P->k > 0) {
        const long long kk = (long long)P->k * P->k;
        const int nE = (int)C.e_key.size();
        hipLaunchKernelGGL(k_fill64, dim3(nblocks(kk)), dim3(kBlock), 0, st, kk, S->E.p, 0.0);
        hipLaunchKernelGGL(k_coarse_E, dim3(nblocks(nE, kBlock / 64)), dim3(kBlock), 0, st, nE, C.d_e_ptr.p,
                           C.d_e_key.p, C.d_e_slot.p, S->v64.p, S->E.p);
        hipLaunchKernelGGL(k_coarse_inv, dim3(1), dim3(kInvThreads), 0, st, P->k, S->E.p, S->Einv.p);
      }
      MS_CHECK_LAUNCH();
      KS *dks = d_ks_.p + S->id;
      S->best = HUGE_VAL;
      S->it_best = 0;
      S->stalled = false;
      S->hks = KS{};
      S->hks.tol = r->tol;
      S->hks.maxit = r->maxit;
      S->hks.delta = o_.delta;
      S->hks.coarse = P->k > 0;
      MS_HIP(hipMemcpyAsync(dks, &S->hks, sizeof(KS), hipMemcpyHostToDevice, st));
      S->hsys = BSys{A, S->v32.p, S->v64.p, S->dg.p, S->r.p, S->z.p, S->p.p, S->q.p, S->y.p, S->x64.p, S->b64.p,
                     S->r64.p, dks, P->pat.color_ptr.size() > 1 ? P->pat.color_ptr[1] : n, P->k,
                     P->k > 0 ? C.d_agg_cptr.p : nullptr, P->k > 0 ? C.d_agg_cidx.p : nullptr,
                     P->k > 0 ? S->Einv.p : nullptr, P->k > 0 ? S->nu.p : nullptr};
      MS_HIP(hipMemcpyAsync(d_sys_.p + S->id, &S->hsys, sizeof(BSys), hipMemcpyHostToDevice, st));
      r->t_admit = now_s();
      su.slots.push_back(S.get());
      staged_.push_back(std::move(S));
    }
    MS_HIP(hipEventCreateWithFlags(&su.done, hipEventDisableTiming));
    MS_HIP(hipEventRecord(su.done, st));
    setups_.push_back(su);
  }

  // Set-ups whose event fired (or, with wait, the oldest one): their slots
  // are returned and their pinned stages go back to the pattern pools.
  std::vector<Slot *> finished_setups(bool wait) {
    std::vector<Slot *> ready;
    for (size_t i = 0; i < setups_.size();) {
      hipError_t e = (wait && i == 0) ? hipEventSynchronize(setups_[i].done) : hipEventQuery(setups_[i].done);
      if (e == hipErrorNotReady) {
        i++;
        continue;
      }
      MS_HIP(e);
      MS_HIP(hipEventDestroy(setups_[i].done));
      for (Slot *S : setups_[i].slots) ready.push_back(S);
      setups_.erase(setups_.begin() + (long)i);
      wait = false;
    }
    if (!ready.empty()) {
      std::lock_guard<std::mutex> lk(mu_);
      for (Slot *S : ready) {
        S->P->stages.push_back(S->req->stage);
        S->req->stage = Stage{};
      }
    }
    return ready;
  }

  // batched start of set-up requests on the compute stream
  void start(const std::vector<Slot *> &fresh) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      for (Slot *S : fresh)
        for (size_t i = 0; i < staged_.size(); i++)
          if (staged_[i].get() == S) {
            active_.push_back(std::move(staged_[i]));
            staged_.erase(staged_.begin() + (long)i);
            break;
          }
    }
    aux_.build(fresh);
    ensure_partials(aux_.items.size());
    launch(k_b_sumsq, aux_, part64_.p);
    fin<1, double, FinStart0>(aux_, part64_.p, 0);
    launch(k_b_init, aux_, mode(), part32_.p);
    fin<3, float, FinStartJ>(aux_, part32_.p, 1, jac());
    if (!jac()) {
      sweeps(aux_);
      fin<1, float, FinRz0>(aux_, prz_.p, 0);
    }
    launch(k_b_p, aux_, jac());
    MS_CHECK_LAUNCH();
    dirty_ = true;
  }

  void round_body(const List &L) {
    for (int it = 0; it < o_.round_iters; it++) {
      launch(k_b_spmv, L, part32_.p);
      fin<1, float, FinPap>(L, part32_.p, 0);
      launch(k_b_update, L, mode(), part32_.p);
      fin<3, float, FinUpdateRU>(L, part32_.p, 1, jac(), 1);
      tail(L);
    }
  }

  // One round: with graphs enabled, a set that stayed unchanged over one
  // round has its next round captured, and the graph is replayed while the
  // set stays unchanged (a set that changes every round never pays for a
  // capture).  Every argument of those launches (item lists, system table,
  // partial buffers) is fixed until the set changes; the per-request state
  // they read lives in device memory.  A change of the set (a request joins
  // or retires) drops the graph.
  void round() {
    const List &L = all_;
    if (graphs_enabled() && (round_exec_ || stable_rounds_ > 0)) {
      if (!round_exec_) {
        const long long l0 = launches_;
        hipGraph_t g = nullptr;
        std::lock_guard<std::recursive_mutex> lk(device_alloc_mutex());
        MS_HIP(hipStreamBeginCapture(stream(), hipStreamCaptureModeRelaxed));
        try {
          round_body(L);
        } catch (...) {
          (void)hipStreamEndCapture(stream(), &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        MS_HIP(hipStreamEndCapture(stream(), &g));
        const hipError_t e = hipGraphInstantiate(&round_exec_, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
          round_exec_ = nullptr;
          fail(MS_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
        round_launches_ = launches_ - l0;
        launches_ -= round_launches_;
      }
      MS_HIP(hipGraphLaunch(round_exec_, stream()));
      launches_ += round_launches_;
    } else {
      round_body(L);
    }
    MS_CHECK_LAUNCH();
    rounds_++;
    stable_rounds_++;
    read_states();
  }
  void drop_round_graph() {
    stable_rounds_ = 0;
    if (round_exec_) MS_HIP(hipGraphExecDestroy(round_exec_));
    round_exec_ = nullptr;
  }

  void after_round() {
    std::vector<Slot *> slots = active_ptrs(), ru, paused;
    std::vector<char> exhausted(o_.max_active, 0);
    // stall guard (as the mixed PCG of a system): a request whose FP32
    // residual has not halved over the stall window is retired like an
    // exhausted one, with its true residual and ierr = 3
    const int window = stall_window();
    for (Slot *s : slots) {
      const KS &h = h_ks_[s->id];
      if (h.done == 1) continue;
      if (h.res < 0.5 * s->best) {
        s->best = h.res;
        s->it_best = h.it;
      } else if (h.it - s->it_best >= window) {
        s->stalled = true;
      }
    }
    for (Slot *s : slots) {
      KS &h = h_ks_[s->id];
      const bool ex = h.done != 1 && (h.it >= s->req->maxit || s->stalled);
      if (((h.ru == 1 || h.done == 2) && h.done != 1) || ex) {
        ru.push_back(s);
        if (h.done == 2 && !ex) paused.push_back(s);
        if (ex) {
          exhausted[s->id] = 1;
          if (h.ru != 1 || h.done != 0) {   // force the final reliable update
            h.ru = 1;
            h.done = 0;
            MS_HIP(hipMemcpyAsync(d_ks_.p + s->id, &h, sizeof(KS), hipMemcpyHostToDevice, stream()));
          }
        } else if (h.ru != 1) {            // a pause always carries the update flag
          h.ru = 1;
          MS_HIP(hipMemcpyAsync(d_ks_.p + s->id, &h, sizeof(KS), hipMemcpyHostToDevice, stream()));
        }
      }
    }
    if (!ru.empty()) {
      aux_.build(ru);
      launch(k_b_fold, aux_);
      launch(k_b_true_res, aux_, part64_.p);
      fin<1, double, FinRU>(aux_, part64_.p, 0);
      launch(k_b_rescale, aux_, mode(), part32_.p);
      if (!paused.empty()) {     // complete the interrupted iterations on the new r
        tail_.subset(aux_, ru, paused);
        fin<3, float, FinUpdateRU>(tail_, part32_.p, 1, jac(), 0);
        tail(tail_);
      }
      MS_CHECK_LAUNCH();
      read_states();
    }
    // retire finished requests; their buffers go to the pattern pool
    std::vector<Slot *> fin_slots;
    for (Slot *s : slots) {
      const KS &h = h_ks_[s->id];
      if (h.done == 1 || exhausted[s->id]) fin_slots.push_back(s);
    }
    if (fin_slots.empty()) return;
    for (Slot *s : fin_slots) {
      const int n = s->P->n;
      hipLaunchKernelGGL(k_s_unperm, dim3(nblocks(n)), dim3(kBlock), 0, stream(), n, s->P->d_o2n.p, s->x64.p,
                         s->r64.p);
      s->req->x.resize(n);
      MS_HIP(hipMemcpyAsync(s->req->x.data(), s->r64.p, n * sizeof(double), hipMemcpyDeviceToHost, stream()));
    }
    MS_HIP(hipStreamSynchronize(stream()));
    {
      std::lock_guard<std::mutex> lk(mu_);
      const double t = now_s();
      for (Slot *s : fin_slots) {
        const KS &h = h_ks_[s->id];
        Req *r = s->req;
        r->it = h.it;
        r->res = h.res;
        r->ierr = h.ierr;
        if (r->ierr == 0 && !(h.done == 1 && h.res <= r->tol)) r->ierr = s->stalled ? 3 : 1;
        r->t_done = t;
        r->done = true;
        n_done_++;
        free_ids_.push_back(s->id);
      }
      for (size_t i = 0; i < active_.size();) {
        Slot *s = active_[i].get();
        if (std::find(fin_slots.begin(), fin_slots.end(), s) == fin_slots.end()) {
          i++;
          continue;
        }
        used_ -= s->bytes;
        pooled_ += s->bytes;
        s->req = nullptr;
        s->P->pool.push_back(std::move(active_[i]));
        active_.erase(active_.begin() + (long)i);
      }
    }
    cv_done_.notify_all();
    dirty_ = true;
  }

  void loop() {
    try {
      MS_HIP(hipSetDevice(device_));
      while (true) {
        std::vector<std::unique_ptr<Slot>> admit;
        {
          std::unique_lock<std::mutex> lk(mu_);
          cv_work_.wait(lk, [&] { return stop_ || !pending_.empty() || !active_.empty() || !staged_.empty(); });
          if (stop_) break;
          while (!pending_.empty() && (int)(active_.size() + staged_.size() + admit.size()) < o_.max_active) {
            Req *r = pending_.front();
            const bool idle = active_.empty() && staged_.empty() && admit.empty();
            std::unique_ptr<Slot> S = take_slot(r->P, idle);
            if (!S) break;
            pending_.pop_front();
            peak_ = std::max(peak_, used_);
            S->req = r;
            S->P = r->P;
            S->id = free_ids_.back();
            free_ids_.pop_back();
            admit.push_back(std::move(S));
          }
        }
        if (!admit.empty()) begin_setup(admit);
        // requests whose set-up finished join at this round boundary; with
        // nothing iterating, wait for the oldest set-up instead of spinning
        std::vector<Slot *> ready = finished_setups(active_.empty());
        if (!ready.empty()) start(ready);
        if (active_.empty()) continue;
        if (dirty_) {
          drop_round_graph();
          all_.build(active_ptrs());
          ensure_partials(all_.items.size());
          dirty_ = false;
        }
        round();
        after_round();
      }
    } catch (const Error &e) {
      std::lock_guard<std::mutex> lk(mu_);
      err_ = e.msg;
    } catch (...) {
      std::lock_guard<std::mutex> lk(mu_);
      err_ = "unknown error in the server worker";
    }
    cv_done_.notify_all();
  }
};

Server *server_create(const ServeOpts &o) { return new Server(o); }
void server_destroy(Server *s) { delete s; }
int server_add_pattern(Server *s, int n, const int *rp, const int *ci, const double *xc) {
  return s->add_pattern(n, rp, ci, xc);
}
long long server_submit(Server *s, int pattern, const double *val, const double *b, double tol, int maxit) {
  return s->submit(pattern, val, b, tol, maxit);
}
bool server_result(Server *s, long long ticket, double timeout_ms, double *x, ServeResult *r) {
  return s->result(ticket, timeout_ms, x, r);
}
void server_stats(Server *s, long long *out8) { s->stats(out8); }

}  // namespace ms
