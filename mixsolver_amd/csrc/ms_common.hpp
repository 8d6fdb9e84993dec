// ms_common.hpp -- shared declarations of libmixsolver_hip (gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mixsolver.h"

namespace ms {

// ----------------------------------------------------------------- errors
struct Error {
  int code;
  std::string msg;
};
void set_error(const std::string &m);
[[noreturn]] void fail(int code, const std::string &m);

#define MS_HIP(call)                                                            \
  do {                                                                          \
    hipError_t e_ = (call);                                                     \
    if (e_ != hipSuccess)                                                       \
      ::ms::fail(MS_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define MS_CHECK_LAUNCH() MS_HIP(hipGetLastError())

hipStream_t stream();           // the calling thread's stream (default: the library stream)
void ensure_runtime(int device); // select the device, create the library stream
int runtime_device();            // device selected by ensure_runtime (-1 before)
// A worker thread that drives systems of its own (the sweep server) gives
// itself a stream of its own: every library call made on that thread then
// launches and synchronises there.  Default priority; nullptr restores the
// library stream.
void set_thread_stream(hipStream_t s);
// device bytes allocated minus freed by DBuf on the calling thread (the
// sweep server measures a case's footprint with it)
long long thread_device_bytes();
void count_device_bytes(long long delta);
// Held around device allocations and frees and around stream captures, so
// that no thread allocates or frees (which may synchronise the device) while
// another thread captures a graph.
std::recursive_mutex &device_alloc_mutex();
// replay the mixed-precision PCG chunk (16 iterations) from a captured HIP
// graph instead of launching its kernels one by one (default on; the
// environment variable MIXSOLVER_GRAPHS=0 or ms_set_graphs(0) turns it off)
bool graphs_enabled();
void set_graphs(bool on);
// the drivers assemble the three momentum components' interior faces in one
// pass where that is exact (default on; MIXSOLVER_MOM_FUSED=0 or
// ms_set_fused_momentum(0) turns it off)
bool momentum_fused_enabled();
void set_momentum_fused(bool on);
// 16-bit column offsets with far columns read from the int32 array
// (MIXSOLVER_COL16_FAR=1 at start-up; default off: 16-bit offsets only for
// patterns without far columns, since the far branch and its dependent load
// cost more than the 2 bytes per entry they save, 15.98 against 17.80 PISO
// steps/s at 2M cells)
bool col16_far_enabled();
// the mixed PCG queues its next chunk before reading the current chunk's
// state (MIXSOLVER_PIPELINE=0 turns it off)
bool pipeline_enabled();

// --------------------------------------------------------- device buffers
template <class T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  bool own = true;   // false: a read-only view of another system's buffer (forked systems)
  DBuf() = default;
  explicit DBuf(size_t count) { alloc(count); }
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  DBuf(DBuf &&o) noexcept : p(o.p), n(o.n), own(o.own) { o.p = nullptr; o.n = 0; o.own = true; }
  DBuf &operator=(DBuf &&o) noexcept {
    if (this != &o) {
      release();
      p = o.p; n = o.n; own = o.own;
      o.p = nullptr; o.n = 0; o.own = true;
    }
    return *this;
  }
  ~DBuf() { release(); }
  void alloc(size_t count) {
    release();
    n = count;
    if (count) {
      std::lock_guard<std::recursive_mutex> g(device_alloc_mutex());
      hipError_t e = hipMalloc(&p, count * sizeof(T));
      if (e != hipSuccess) { p = nullptr; n = 0; fail(MS_ENOMEM, "hipMalloc failed"); }
      count_device_bytes((long long)(count * sizeof(T)));
    }
  }
  // view of o's memory; never freed through this buffer
  void alias(const DBuf &o) {
    release();
    p = o.p;
    n = o.n;
    own = false;
  }
  void release() {
    if (p && own) {
      std::lock_guard<std::recursive_mutex> g(device_alloc_mutex());
      (void)hipFree(p);
      count_device_bytes(-(long long)(n * sizeof(T)));
    }
    p = nullptr;
    n = 0;
    own = true;
  }
  // bytes this buffer owns (0 for a view)
  size_t owned_bytes() const { return own ? n * sizeof(T) : 0; }
  void upload(const T *h, size_t count) {
    if (count > n) alloc(count);
    if (count) MS_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, stream()));
  }
  void upload(const std::vector<T> &v) { upload(v.data(), v.size()); }
  void download(T *h, size_t count) const {
    if (count) MS_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, stream()));
  }
  void zero() {
    if (n) MS_HIP(hipMemsetAsync(p, 0, n * sizeof(T), stream()));
  }
  size_t bytes() const { return n * sizeof(T); }
};

constexpr int kBlock = 256;   // 4 wave64 per workgroup
constexpr int kSlice = 64;    // SELL slice height = one wavefront
constexpr int kSegThreads = 1024;   // workgroup of a run of narrow levels
constexpr int kNarrowRows = 2048;   // widest level such a run takes

inline int nblocks(long long n, int b = kBlock) { return (int)((n + b - 1) / b); }

// ------------------------------------------------------ coarse space
// Piecewise-constant coarse space for the pressure-correction PCG: cells are
// grouped into k compact aggregates (recursive coordinate bisection of the
// cell centres) and the preconditioner becomes M^-1 r + Z E^-1 Z^T r with
// Z the aggregate indicator matrix and E = Z^T A Z.  The device numbering
// keeps every aggregate contiguous inside each colour, so the rows are cut
// into chunks of at most kBlock rows that each lie in one aggregate; the
// cell-wise PCG kernels run one workgroup per chunk, which turns Z^T r into
// per-workgroup partial sums and Z nu into a broadcast.
constexpr int kMaxAgg = 1024;
constexpr int kInvRegMax = 256;   // largest E of the one-workgroup register inverter (request server)
struct Coarse {
  int k = 0;                          // number of aggregates (0 = none)
  int nchunks = 0;
  std::vector<int> chunk_ptr;         // nchunks+1 row boundaries
  std::vector<int> chunk_agg;         // aggregate of each chunk
  std::vector<int> agg_cptr, agg_cidx; // chunks of each aggregate (CSR)
  std::vector<int> e_ptr, e_key;      // E gather map: entry key a*k+b, slots [e_ptr[i], e_ptr[i+1])
  std::vector<int> e_slot;            // matrix slots, grouped by key, ascending
  DBuf<int> d_chunk_ptr, d_chunk_agg, d_agg_cptr, d_agg_cidx, d_e_ptr, d_e_key, d_e_slot;
  // (the matrix-dependent E, its inverse and the coarse solution belong to
  // each solver: CoarseMat)
  void upload();
  size_t device_bytes() const;
};

// Values of the coarse space for one matrix: dense E (k x k), FP64
// Gauss-Jordan work matrix and pivot snapshots, E^-1 in FP32, and the
// coarse solution nu (k).  Owned by the solver, so systems that share a
// pattern (forked sweep cases) keep their own.
struct CoarseMat {
  DBuf<double> E, nu, W, rc, Eparts;
  DBuf<float> Einv;
  void ensure(int k) {
    const size_t kk = (size_t)k * k;
    if (E.n < kk) E.alloc(kk);
    if (Einv.n < kk) Einv.alloc(kk);
    if (W.n < kk) W.alloc(kk);
    if (rc.n < 4 * (size_t)k) rc.alloc(4 * (size_t)k);
    if (nu.n < (size_t)k) nu.alloc(k);
  }
};

// ------------------------------------------------- SELL-64 sparse pattern
// Row r lives in slice s = r/64, lane l = r%64; its k-th entry (columns in
// ascending order, as csr_t) sits at off[s] + k*64 + l.  Padding entries use
// col = r, val = 0.  diagk[r] = k of the diagonal entry.
struct SellPattern {
  int n = 0, nslices = 0;
  long long nnz = 0, nstore = 0;
  std::vector<long long> off;       // nslices+1
  std::vector<int> rowlen, col, diagk;
  std::vector<long long> tpos;       // stored slot of (c,r) for each stored (r,c); -1 pad
  DBuf<long long> d_off, d_tpos;
  DBuf<int> d_rowlen, d_col, d_diagk;
  // 16-bit column offsets for the FP32 kernels: col = cbase[slice] + col16,
  // 0xFFFF = the row itself (diagonal and padding), 0xFFFE = a far column
  // outside the slice's window, read from col (has16; nfar such entries)
  std::vector<unsigned short> col16;
  long long nfar = 0;
  std::vector<int> cbase;
  DBuf<unsigned short> d_col16;
  DBuf<int> d_cbase;
  bool has16 = false;
  // level schedules (natural-order triangular sweeps / factorizations)
  std::vector<int> lev_ptr_f, lev_ptr_b;   // rows grouped per level
  DBuf<int> d_lev_rows_f, d_lev_rows_b, d_lev_ptr_f, d_lev_ptr_b;
  // launch plan of a sweep: runs of consecutive narrow levels (at most
  // kNarrowRows rows each) go to one workgroup that steps through them with
  // a barrier between levels; every wider level is a launch of its own
  struct LevSeg {
    int l0, l1;
    bool narrow;
  };
  std::vector<LevSeg> seg_f, seg_b;
  // colour classes: rows of colour c are [col_ptr[c], col_ptr[c+1])
  std::vector<int> color_ptr;
  bool has_levels = false;
  bool symmetric = true;             // structurally symmetric (tpos complete)
  Coarse coarse;                     // coarse space of the pressure-correction solves
  long long slot(int r, int k) const { return off[r / kSlice] + (long long)k * kSlice + (r % kSlice); }
  void build_from_csr(int n, const std::vector<int> &rp0, const std::vector<int> &ci0); // 0-based CSR
  void build_levels();
  void upload();
  size_t device_bytes() const;
};



}  // namespace ms
