// ms_setup.cpp -- host-side construction of a device system.
//
// Restates the setup chain of the reference for one partition:
//   compute_geometry  (mod_uns_geometry.f90:40-174; Newell area vector,
//                      fan-triangulated face centroid, pyramid volumes and
//                      centroids about the node-average point, c0->c1
//                      orientation)
//   build_connectivity(mod_uns_connectivity.f90:28-108)
//   build_cell_csr    (mod_uns_simple.f90:835-904)
//   build_bc          (mod_uns_bc.f90:155-265)
//   init_fields lf    (mod_uns_fields.f90:147-157)
// with the same evaluation order, so every geometric coefficient handed to
// the device equals the reference's bit for bit.  On top of that it builds
// the device layout: optional colour-blocked / RCM renumbering of cells,
// the SELL-64 matrix pattern, per-cell face lists with matrix slot maps,
// level schedules for natural-order factorizations and per-face constants.
#include <algorithm>
#include <atomic>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <queue>

#include "ms_system.hpp"

namespace ms {

// ------------------------------------------------------------- runtime glue
static thread_local std::string g_err;
static hipStream_t g_stream = nullptr;
static int g_device = -1;

void set_error(const std::string &m) { g_err = m; }
void fail(int code, const std::string &m) { throw Error{code, m}; }
const char *last_error() { return g_err.c_str(); }
void noalloc_register_device(int device);   // ms_noalloc.hip


void ensure_runtime(int device) {
  if (g_stream && (device < 0 || device == g_device)) return;
  int ndev = 0;
  MS_HIP(hipGetDeviceCount(&ndev));
  if (ndev <= 0) fail(MS_EDEVICE, "no HIP device visible");
  if (device < 0) device = 0;
  if (device >= ndev) fail(MS_EINVAL, "device index out of range");
  MS_HIP(hipSetDevice(device));
  if (g_stream) MS_HIP(hipStreamDestroy(g_stream));
  MS_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  g_device = device;
  noalloc_register_device(device);
}

static thread_local hipStream_t t_stream = nullptr;
static thread_local long long t_dev_bytes = 0;

hipStream_t stream() {
  if (t_stream) return t_stream;
  if (!g_stream) ensure_runtime(-1);
  return g_stream;
}
void set_thread_stream(hipStream_t s) { t_stream = s; }
bool col16_far_enabled() {
  static const bool on = [] {
    const char *e = std::getenv("MIXSOLVER_COL16_FAR");
    return e && e[0] == '1';
  }();
  return on;
}
std::recursive_mutex &device_alloc_mutex() {
  static std::recursive_mutex m;
  return m;
}
int runtime_device() { return g_device; }
long long thread_device_bytes() { return t_dev_bytes; }
void count_device_bytes(long long delta) { t_dev_bytes += delta; }

static std::atomic<int> g_graphs{-1};   // -1: not yet read from the environment
bool graphs_enabled() {
  int v = g_graphs.load();
  if (v < 0) {
    const char *e = std::getenv("MIXSOLVER_GRAPHS");
    v = (e && e[0] == '0') ? 0 : 1;
    g_graphs.store(v);
  }
  return v == 1;
}
void set_graphs(bool on) { g_graphs.store(on ? 1 : 0); }
static std::atomic<int> g_mom_fused{-1};   // -1: not yet read from the environment
bool momentum_fused_enabled() {
  int v = g_mom_fused.load();
  if (v < 0) {
    const char *e = std::getenv("MIXSOLVER_MOM_FUSED");
    v = (e && e[0] == '0') ? 0 : 1;
    g_mom_fused.store(v);
  }
  return v != 0;
}
void set_momentum_fused(bool on) { g_mom_fused.store(on ? 1 : 0); }
bool pipeline_enabled() {
  static const bool on = [] {
    const char *e = std::getenv("MIXSOLVER_PIPELINE");
    return !(e && e[0] == '0');
  }();
  return on;
}

// ------------------------------------------------------------------- norms
double norm2_ref(const double *v, int n) {
  double mx = 0.0, s = 0.0;
  for (int i = 0; i < n; i++) {
    double ax = std::fabs(v[i]);
    if (mx == 0.0) {
      mx = ax;
    } else if (ax > mx) {
      double t = mx / ax;
      double tsq = t * t;
      s = s * tsq;
      s = s + tsq;
      mx = ax;
    } else {
      double t = ax / mx;
      s = s + t * t;
    }
  }
  return mx * std::sqrt(1.0 + s);
}

static inline double dot3(const double *a, const double *b) {
  double s = 0.0;
  s = s + a[0] * b[0];
  s = s + a[1] * b[1];
  s = s + a[2] * b[2];
  return s;
}

// -------------------------------------------------------------- SELL pattern
void csr_to_sell(SellPattern &P, int n, const std::vector<int> &rp, const std::vector<int> &ci) {
  P.n = n;
  P.nslices = (n + kSlice - 1) / kSlice;
  P.nnz = rp[n];
  P.rowlen.assign(n, 0);
  P.diagk.assign(n, -1);
  P.off.assign(P.nslices + 1, 0);
  for (int r = 0; r < n; r++) P.rowlen[r] = rp[r + 1] - rp[r];
  for (int s = 0; s < P.nslices; s++) {
    int w = 0;
    for (int r = s * kSlice; r < std::min(n, (s + 1) * kSlice); r++) w = std::max(w, P.rowlen[r]);
    P.off[s + 1] = P.off[s] + (long long)w * kSlice;
  }
  P.nstore = P.off[P.nslices];
  P.col.assign(P.nstore, 0);
  for (int s = 0; s < P.nslices; s++) {   // padding: col = row itself
    long long w = (P.off[s + 1] - P.off[s]) / kSlice;
    for (long long k = 0; k < w; k++)
      for (int l = 0; l < kSlice; l++) {
        int r = s * kSlice + l;
        P.col[P.off[s] + k * kSlice + l] = r < n ? r : 0;
      }
  }
  for (int r = 0; r < n; r++)
    for (int k = 0; k < P.rowlen[r]; k++) {
      int c = ci[rp[r] + k];
      P.col[P.slot(r, k)] = c;
      if (c == r) P.diagk[r] = k;
    }
  for (int r = 0; r < n; r++)
    if (P.diagk[r] < 0) fail(MS_EINVAL, "matrix row without diagonal entry");
  // transpose slot map (pattern is structurally symmetric)
  P.tpos.assign(P.nstore, -1);
  P.symmetric = true;
  for (int r = 0; r < n; r++)
    for (int k = 0; k < P.rowlen[r]; k++) {
      int c = P.col[P.slot(r, k)];
      const int *b = &ci[rp[c]], *e = &ci[rp[c + 1]];
      const int *f = std::lower_bound(b, e, r);
      if (f == e || *f != r) {
        P.symmetric = false;
        continue;
      }
      P.tpos[P.slot(r, k)] = P.slot(c, (int)(f - b));
    }
  P.color_ptr = {0, n};
  // 16-bit offsets per slice: each slice takes the window of kCol16Span
  // columns that holds most of its off-diagonal columns; columns outside it
  // (far: neighbours across a periodic seam or an aggregate split plane)
  // are marked and read from the int32 column array.  Used when at most a
  // quarter of the off-diagonal entries are far.
  constexpr int kCol16Span = 0xFFFD;   // offsets 0 .. 0xFFFD; 0xFFFE far, 0xFFFF the row itself
  const bool allow_far = col16_far_enabled();
  P.cbase.assign(P.nslices, 0);
  P.has16 = true;
  P.nfar = 0;
  long long noff = 0;
  std::vector<int> v;
  std::vector<unsigned char> far_slice(P.nslices, 0);
  for (int s = 0; s < P.nslices && P.has16; s++) {
    v.clear();
    for (int r = s * kSlice; r < std::min(n, (s + 1) * kSlice); r++)
      for (int k = 0; k < P.rowlen[r]; k++) {
        const int c = P.col[P.slot(r, k)];
        if (c != r) v.push_back(c);
      }
    noff += (long long)v.size();
    if (v.empty()) continue;
    std::sort(v.begin(), v.end());
    size_t best = 0, bi = 0;
    for (size_t i = 0, j = 0; j < v.size(); j++) {
      while ((long long)v[j] - v[i] > kCol16Span) i++;
      if (j - i + 1 > best) {
        best = j - i + 1;
        bi = i;
      }
    }
    P.cbase[s] = v[bi];
    const long long nf = (long long)(v.size() - best);
    if (nf > 0 && !allow_far) P.has16 = false;
    P.nfar += nf;
  }
  if (P.has16 && 4 * P.nfar > noff) P.has16 = false;
  P.col16.clear();
  if (P.has16) {
    P.col16.assign(P.nstore, 0xFFFF);
    for (int r = 0; r < n; r++)
      for (int k = 0; k < P.rowlen[r]; k++) {
        const long long q = P.slot(r, k);
        const int c = P.col[q];
        if (c == r) continue;
        const long long d = (long long)c - P.cbase[r / kSlice];
        P.col16[q] = (d >= 0 && d <= kCol16Span) ? (unsigned short)d : (unsigned short)0xFFFE;
      }
  } else {
    P.nfar = 0;
  }
}

void SellPattern::build_levels() {
  // forward: level(i) = 1 + max level of lower neighbours; backward on upper
  std::vector<int> lf(n, 0), lb(n, 0);
  int nlf = 0, nlb = 0;
  for (int r = 0; r < n; r++) {
    int l = 0;
    for (int k = 0; k < diagk[r]; k++) l = std::max(l, lf[col[slot(r, k)]] + 1);
    lf[r] = l;
    nlf = std::max(nlf, l + 1);
  }
  for (int r = n - 1; r >= 0; r--) {
    int l = 0;
    for (int k = diagk[r] + 1; k < rowlen[r]; k++) l = std::max(l, lb[col[slot(r, k)]] + 1);
    lb[r] = l;
    nlb = std::max(nlb, l + 1);
  }
  auto group = [&](const std::vector<int> &lv, int nl, std::vector<int> &ptr, std::vector<int> &rows) {
    ptr.assign(nl + 1, 0);
    for (int r = 0; r < n; r++) ptr[lv[r] + 1]++;
    for (int l = 0; l < nl; l++) ptr[l + 1] += ptr[l];
    rows.assign(n, 0);
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (int r = 0; r < n; r++) rows[fill[lv[r]]++] = r;
  };
  std::vector<int> rf, rb;
  group(lf, nlf, lev_ptr_f, rf);
  group(lb, nlb, lev_ptr_b, rb);
  auto plan = [](const std::vector<int> &ptr, std::vector<LevSeg> &seg) {
    seg.clear();
    const int nl = (int)ptr.size() - 1;
    for (int l = 0; l < nl;) {
      if (ptr[l + 1] - ptr[l] > kNarrowRows) {
        seg.push_back({l, l + 1, false});
        l++;
        continue;
      }
      int e = l;
      while (e < nl && ptr[e + 1] - ptr[e] <= kNarrowRows) e++;
      seg.push_back({l, e, true});
      l = e;
    }
  };
  plan(lev_ptr_f, seg_f);
  plan(lev_ptr_b, seg_b);
  d_lev_rows_f.upload(rf);
  d_lev_rows_b.upload(rb);
  d_lev_ptr_f.upload(lev_ptr_f);
  d_lev_ptr_b.upload(lev_ptr_b);
  has_levels = true;
}

void SellPattern::upload() {
  d_off.upload(off);
  d_rowlen.upload(rowlen);
  d_col.upload(col);
  d_diagk.upload(diagk);
  d_tpos.upload(tpos);
  if (has16) {
    d_col16.upload(col16);
    d_cbase.upload(cbase);
  }
}

size_t SellPattern::device_bytes() const {
  return d_off.bytes() + d_rowlen.bytes() + d_col.bytes() + d_diagk.bytes() + d_tpos.bytes() +
         d_lev_rows_f.bytes() + d_lev_rows_b.bytes() + d_lev_ptr_f.bytes() + d_lev_ptr_b.bytes() +
         coarse.device_bytes() + d_col16.bytes() + d_cbase.bytes();
}

// ------------------------------------------------------------ orderings
// Reverse Cuthill-McKee on the cell graph, starting from a minimum-degree
// cell of each component; ties broken by natural index (deterministic).
static std::vector<int> rcm_order(int n, const std::vector<int> &ptr, const std::vector<int> &adj) {
  std::vector<int> deg(n), order;
  order.reserve(n);
  for (int i = 0; i < n; i++) deg[i] = ptr[i + 1] - ptr[i];
  std::vector<char> seen(n, 0);
  std::vector<int> byDeg(n);
  std::iota(byDeg.begin(), byDeg.end(), 0);
  std::stable_sort(byDeg.begin(), byDeg.end(), [&](int a, int b) { return deg[a] < deg[b]; });
  std::vector<int> nb;
  for (int start : byDeg) {
    if (seen[start]) continue;
    std::queue<int> q;
    q.push(start);
    seen[start] = 1;
    while (!q.empty()) {
      int v = q.front();
      q.pop();
      order.push_back(v);
      nb.clear();
      for (int k = ptr[v]; k < ptr[v + 1]; k++)
        if (!seen[adj[k]]) nb.push_back(adj[k]);
      std::sort(nb.begin(), nb.end(), [&](int a, int b) { return deg[a] != deg[b] ? deg[a] < deg[b] : a < b; });
      nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
      for (int w : nb) {
        seen[w] = 1;
        q.push(w);
      }
    }
  }
  std::reverse(order.begin(), order.end());
  return order;   // order[new] = old
}

// Greedy colouring visiting cells in `visit` order; returns colour per cell.
std::vector<int> greedy_colors(int n, const std::vector<int> &ptr, const std::vector<int> &adj,
                               const std::vector<int> &visit, int &ncol);

// The device numbering of the cells (n2o[new] = old): natural or RCM visiting
// order; with colouring, the greedy colour classes in ascending order, and
// with k aggregates (agg[old]), each aggregate contiguous inside each class;
// the visiting order decides ties.  ptr/adj: cell graph without self loops.  color[old] is
// the colour of each cell (all 0 without colouring).  Deterministic.
void cell_ordering(int nc, const std::vector<int> &ptr, const std::vector<int> &adj, const std::vector<double> &xc,
                   bool rcm, bool colored, int nagg, std::vector<int> &n2o, std::vector<int> &color, int &ncolors,
                   std::vector<int> &agg) {
  std::vector<int> visit(nc);
  std::iota(visit.begin(), visit.end(), 0);
  if (rcm) visit = rcm_order(nc, ptr, adj);
  color.assign(nc, 0);
  ncolors = 1;
  if (colored) color = greedy_colors(nc, ptr, adj, visit, ncolors);
  agg.assign(nc, 0);
  if (nagg > 0) agg = rcb_aggregates(nc, xc, nagg);
  n2o = visit;
  if (colored || nagg > 0) {
    std::vector<int> pos(nc);
    for (int i = 0; i < nc; i++) pos[visit[i]] = i;
    std::stable_sort(n2o.begin(), n2o.end(), [&](int a, int b) {
      if (color[a] != color[b]) return color[a] < color[b];
      if (agg[a] != agg[b]) return agg[a] < agg[b];
      return pos[a] < pos[b];
    });
  }
}

// Greedy colouring visiting cells in `visit` order; returns colour per cell.
std::vector<int> greedy_colors(int n, const std::vector<int> &ptr, const std::vector<int> &adj,
                                      const std::vector<int> &visit, int &ncol) {
  std::vector<int> color(n, -1);
  std::vector<int> mark(64, -1);
  ncol = 0;
  for (int v : visit) {
    for (int k = ptr[v]; k < ptr[v + 1]; k++) {
      int c = color[adj[k]];
      if (c >= 0) {
        if (c >= (int)mark.size()) mark.resize(c + 1, -1);
        mark[c] = v;
      }
    }
    int c = 0;
    while (c < (int)mark.size() && mark[c] == v) c++;
    if (c >= (int)mark.size()) mark.resize(c + 1, -1);
    color[v] = c;
    ncol = std::max(ncol, c + 1);
  }
  return color;
}

// Recursive coordinate bisection of the cell centres into k aggregates of
// equal size (+-1): each step splits the cell set at the count-weighted
// position along the longest side of its bounding box.  The order used for
// the split is (coordinate, cell index), a strict total order, so the result
// does not depend on the selection algorithm.  Returns the aggregate of each
// cell (natural numbering).
std::vector<int> rcb_aggregates(int n, const std::vector<double> &xc, int k) {
  std::vector<int> agg(n, 0), ids(n);
  std::iota(ids.begin(), ids.end(), 0);
  struct Task { int lo, hi, a0, cnt; };
  std::vector<Task> todo{{0, n, 0, k}};
  while (!todo.empty()) {
    Task t = todo.back();
    todo.pop_back();
    if (t.cnt <= 1) {
      for (int i = t.lo; i < t.hi; i++) agg[ids[i]] = t.a0;
      continue;
    }
    double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
    for (int i = t.lo; i < t.hi; i++)
      for (int d = 0; d < 3; d++) {
        lo[d] = std::min(lo[d], xc[3 * (size_t)ids[i] + d]);
        hi[d] = std::max(hi[d], xc[3 * (size_t)ids[i] + d]);
      }
    int dir = 0;
    for (int d = 1; d < 3; d++)
      if (hi[d] - lo[d] > hi[dir] - lo[dir]) dir = d;
    const int h = t.cnt / 2;
    const int cut = t.lo + (int)((long long)(t.hi - t.lo) * h / t.cnt);
    std::nth_element(ids.begin() + t.lo, ids.begin() + cut, ids.begin() + t.hi, [&](int a, int b) {
      double xa = xc[3 * (size_t)a + dir], xb = xc[3 * (size_t)b + dir];
      return xa != xb ? xa < xb : a < b;
    });
    todo.push_back({cut, t.hi, t.a0 + h, t.cnt - h});
    todo.push_back({t.lo, cut, t.a0, h});
  }
  return agg;
}

// Chunks (rows of one aggregate and one colour, at most kBlock), chunk lists
// per aggregate and the gather map of E = Z^T A Z over the device numbering
// (aggregate of device row r = agg_dev[r]).  k = 0 builds the chunks only.
void build_coarse(SellPattern &P, const std::vector<int> &agg_dev, int k) {
  Coarse &C = P.coarse;
  const int n = P.n;
  C.k = k;
  C.chunk_ptr.assign(1, 0);
  C.chunk_agg.clear();
  // colour of each row, so that no chunk straddles a colour boundary
  std::vector<int> col_of(n, 0);
  for (size_t c = 0; c + 1 < P.color_ptr.size(); c++)
    for (int r = P.color_ptr[c]; r < P.color_ptr[c + 1]; r++) col_of[r] = (int)c;
  for (int r = 0; r < n; r++) {
    const int start = C.chunk_ptr.back();
    if (r > start && (agg_dev[r] != agg_dev[start] || col_of[r] != col_of[start] || r - start == kBlock)) {
      C.chunk_agg.push_back(agg_dev[start]);
      C.chunk_ptr.push_back(r);
    }
  }
  C.chunk_agg.push_back(agg_dev[C.chunk_ptr.back()]);
  C.chunk_ptr.push_back(n);
  C.nchunks = (int)C.chunk_agg.size();
  if (k == 0) {          // chunks only (aggregate ids all 0): no coarse space
    C.d_chunk_ptr.upload(C.chunk_ptr);
    C.d_chunk_agg.upload(C.chunk_agg);
    return;
  }
  C.agg_cptr.assign(k + 1, 0);
  for (int c = 0; c < C.nchunks; c++) C.agg_cptr[C.chunk_agg[c] + 1]++;
  for (int a = 0; a < k; a++) C.agg_cptr[a + 1] += C.agg_cptr[a];
  C.agg_cidx.assign(C.nchunks, 0);
  {
    std::vector<int> fill(C.agg_cptr.begin(), C.agg_cptr.end() - 1);
    for (int c = 0; c < C.nchunks; c++) C.agg_cidx[fill[C.chunk_agg[c]]++] = c;
  }
  // counting sort of the stored entries by key agg(row)*k + agg(col)
  if (P.nstore > 0x7fffffffLL) fail(MS_EINVAL, "coarse space: matrix too large for 32-bit slots");
  std::vector<int> cnt((size_t)k * k + 1, 0);
  for (int r = 0; r < n; r++)
    for (int j = 0; j < P.rowlen[r]; j++) cnt[(size_t)agg_dev[r] * k + agg_dev[P.col[P.slot(r, j)]] + 1]++;
  for (size_t i = 0; i < (size_t)k * k; i++) cnt[i + 1] += cnt[i];
  C.e_slot.assign(cnt.back(), 0);
  {
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (int r = 0; r < n; r++)
      for (int j = 0; j < P.rowlen[r]; j++) {
        const long long q = P.slot(r, j);
        C.e_slot[fill[(size_t)agg_dev[r] * k + agg_dev[P.col[q]]]++] = (int)q;
      }
  }
  C.e_ptr.assign(1, 0);
  C.e_key.clear();
  for (size_t key = 0; key < (size_t)k * k; key++)
    if (cnt[key + 1] > cnt[key]) {
      std::sort(C.e_slot.begin() + cnt[key], C.e_slot.begin() + cnt[key + 1]);
      C.e_key.push_back((int)key);
      C.e_ptr.push_back(cnt[key + 1]);
    }
  C.upload();
}

void Coarse::upload() {
  d_chunk_ptr.upload(chunk_ptr);
  d_chunk_agg.upload(chunk_agg);
  d_agg_cptr.upload(agg_cptr);
  d_agg_cidx.upload(agg_cidx);
  d_e_ptr.upload(e_ptr);
  d_e_key.upload(e_key);
  d_e_slot.upload(e_slot);
}

size_t Coarse::device_bytes() const {
  return d_chunk_ptr.bytes() + d_chunk_agg.bytes() + d_agg_cptr.bytes() + d_agg_cidx.bytes() + d_e_ptr.bytes() +
         d_e_key.bytes() + d_e_slot.bytes();
}

// ---------------------------------------------------------------- system
// ---- natural-order geometry: compute_geometry (mod_uns_geometry.f90:40-174)
static void compute_geometry(System &s, const double *x, int nf, const long long *fptr, const int *fnodes,
                             const int *c0in, const int *c1orig, int nc) {
  s.xf.assign(3 * (size_t)nf, 0.0);
  s.sf.assign(3 * (size_t)nf, 0.0);
  s.area.assign(nf, 0.0);
  s.vol_nat.assign(nc, 0.0);
  s.xc_nat.assign(3 * (size_t)nc, 0.0);
  auto X = [&](int node, int d) { return x[3 * (size_t)(node - 1) + d]; };
  for (int f = 0; f < nf; f++) {
    int m = (int)(fptr[f + 1] - fptr[f]);
    const int *nd = &fnodes[fptr[f]];
    double sv[3] = {0, 0, 0}, cen[3] = {0, 0, 0}, at = 0.0;
    for (int j = 0; j < m; j++) {
      int jn = (j + 1) % m;
      double a[3] = {X(nd[j], 0), X(nd[j], 1), X(nd[j], 2)}, b[3] = {X(nd[jn], 0), X(nd[jn], 1), X(nd[jn], 2)};
      double c[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
      for (int d = 0; d < 3; d++) sv[d] = sv[d] + 0.5 * c[d];
    }
    double ar = norm2_ref(sv, 3);
    for (int j = 1; j < m - 1; j++) {
      double e1[3], e2[3], st[3], p1[3], p2[3], p3[3];
      for (int d = 0; d < 3; d++) {
        p1[d] = X(nd[0], d);
        p2[d] = X(nd[j], d);
        p3[d] = X(nd[j + 1], d);
        e1[d] = p2[d] - p1[d];
        e2[d] = p3[d] - p1[d];
      }
      st[0] = e1[1] * e2[2] - e1[2] * e2[1];
      st[1] = e1[2] * e2[0] - e1[0] * e2[2];
      st[2] = e1[0] * e2[1] - e1[1] * e2[0];
      for (int d = 0; d < 3; d++) st[d] = st[d] * 0.5;
      double ns = norm2_ref(st, 3);
      for (int d = 0; d < 3; d++) cen[d] = cen[d] + ns * (p1[d] + p2[d] + p3[d]) / 3.0;
      at = at + ns;
    }
    if (at > 0.0)
      for (int d = 0; d < 3; d++) cen[d] = cen[d] / at;
    for (int d = 0; d < 3; d++) {
      s.xf[3 * f + d] = cen[d];
      s.sf[3 * f + d] = sv[d];
    }
    s.area[f] = ar;
  }
  {
    std::vector<double> pcen(3 * (size_t)nc, 0.0);
    std::vector<int> cnt(nc, 0);
    for (int f = 0; f < nf; f++) {
      int a = c0in[f] - 1, b = c1orig[f] - 1;
      for (long long j = fptr[f]; j < fptr[f + 1]; j++) {
        for (int d = 0; d < 3; d++) pcen[3 * a + d] = pcen[3 * a + d] + X(fnodes[j], d);
        cnt[a]++;
        if (b >= 0) {
          for (int d = 0; d < 3; d++) pcen[3 * b + d] = pcen[3 * b + d] + X(fnodes[j], d);
          cnt[b]++;
        }
      }
    }
    for (int c = 0; c < nc; c++)
      if (cnt[c] > 0)
        for (int d = 0; d < 3; d++) pcen[3 * c + d] = pcen[3 * c + d] / (double)cnt[c];
    for (int f = 0; f < nf; f++) {
      double r[3];
      int a = c0in[f] - 1;
      for (int d = 0; d < 3; d++) r[d] = s.xf[3 * f + d] - pcen[3 * a + d];
      if (dot3(&s.sf[3 * f], r) < 0.0)
        for (int d = 0; d < 3; d++) s.sf[3 * f + d] = -s.sf[3 * f + d];
    }
    const double THIRD = 1.0 / 3.0, QRTR = 1.0 / 4.0;
    for (int f = 0; f < nf; f++)
      for (int side = 0; side < 2; side++) {
        int cell = (side == 0 ? c0in[f] : c1orig[f]) - 1;
        if (cell < 0) continue;
        const double *pc = &pcen[3 * cell];
        double r[3];
        for (int d = 0; d < 3; d++) r[d] = s.xf[3 * f + d] - pc[d];
        double dv = THIRD * dot3(&s.sf[3 * f], r) * (side == 0 ? 1.0 : -1.0);
        s.vol_nat[cell] = s.vol_nat[cell] + dv;
        for (int d = 0; d < 3; d++)
          s.xc_nat[3 * cell + d] = s.xc_nat[3 * cell + d] + dv * ((1.0 - QRTR) * s.xf[3 * f + d] + QRTR * pc[d]);
      }
    for (int c = 0; c < nc; c++) {
      if (s.vol_nat[c] > 0.0)
        for (int d = 0; d < 3; d++) s.xc_nat[3 * c + d] = s.xc_nat[3 * c + d] / s.vol_nat[c];
      else
        fail(MS_EINVAL, "non-positive cell volume");
    }
  }
}

// Periodic pairs (no reference counterpart).  Every boundary face of zone A
// is matched with the boundary face of zone B whose centroid, translated by
// the pair's shift, lies on its own (hash of the centroids on a grid of the
// smallest face length, search of the 27 neighbouring bins, tolerance 1e-3 of
// the face length).  The A face becomes an interior face whose c1 is B's
// cell, seen at xc + shift; the B face becomes its shadow and leaves every
// face list.  c1v receives the joined topology (1-based, 0 = boundary).
static void join_periodic(System &s, const ms_ctrl_t *ctrl, int nf, const int *c0in, const int *c1in,
                          const int *zone, std::vector<int> &c1v) {
  s.nper = ctrl->nperiodic;
  s.pidx.assign(nf, 0);
  s.shadow.assign(nf, -1);
  c1v.assign(c1in, c1in + nf);
  if (s.nper == 0) return;
  if (s.nper < 0 || s.nper > 4) fail(MS_EINVAL, "nperiodic must be 0..4");
  for (int k = 0; k < s.nper; k++) {
    const int zA = ctrl->periodic_zone[k][0], zB = ctrl->periodic_zone[k][1];
    const std::string tag = "periodic pair " + std::to_string(zA) + " " + std::to_string(zB) + ": ";
    if (zA == zB) fail(MS_EINVAL, tag + "the two zones must differ");
    std::vector<int> fa, fb;
    for (int f = 0; f < nf; f++)
      if (c1in[f] == 0) {
        if (zone[f] == zA) fa.push_back(f);
        else if (zone[f] == zB) fb.push_back(f);
      }
    if (fa.empty() || fa.size() != fb.size()) fail(MS_EINVAL, tag + "zones need equal, non-zero face counts");
    double sh[3];
    if (ctrl->periodic_auto[k]) {
      double ma[3] = {0, 0, 0}, mb[3] = {0, 0, 0};
      for (int f : fa)
        for (int d = 0; d < 3; d++) ma[d] += s.xf[3 * (size_t)f + d];
      for (int f : fb)
        for (int d = 0; d < 3; d++) mb[d] += s.xf[3 * (size_t)f + d];
      for (int d = 0; d < 3; d++) sh[d] = ma[d] / (double)fa.size() - mb[d] / (double)fb.size();
    } else {
      for (int d = 0; d < 3; d++) sh[d] = ctrl->periodic_shift[k][d];
    }
    double h = INFINITY;
    for (int f : fa) h = std::min(h, std::sqrt(s.area[f]));
    if (!(h > 0.0) || !std::isfinite(h)) fail(MS_EINVAL, tag + "degenerate faces");
    struct Bin {
      long long i, j, k;
      int f;
      bool operator<(const Bin &o) const {
        return i != o.i ? i < o.i : j != o.j ? j < o.j : k != o.k ? k < o.k : f < o.f;
      }
    };
    auto bin = [&](const double *p, int f) {
      return Bin{(long long)std::floor(p[0] / h), (long long)std::floor(p[1] / h), (long long)std::floor(p[2] / h), f};
    };
    std::vector<Bin> bins;
    bins.reserve(fa.size());
    for (int f : fa) bins.push_back(bin(&s.xf[3 * (size_t)f], f));
    std::sort(bins.begin(), bins.end());
    for (int g : fb) {
      double t[3];
      for (int d = 0; d < 3; d++) t[d] = s.xf[3 * (size_t)g + d] + sh[d];
      const Bin c = bin(t, -1);
      int best = -1;
      double bd = INFINITY;
      for (long long di = -1; di <= 1; di++)
        for (long long dj = -1; dj <= 1; dj++)
          for (long long dk = -1; dk <= 1; dk++) {
            Bin lo{c.i + di, c.j + dj, c.k + dk, -1};
            for (auto it = std::lower_bound(bins.begin(), bins.end(), lo);
                 it != bins.end() && it->i == lo.i && it->j == lo.j && it->k == lo.k; ++it) {
              double r[3];
              for (int d = 0; d < 3; d++) r[d] = s.xf[3 * (size_t)it->f + d] - t[d];
              const double dd = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
              if (dd < bd) bd = dd, best = it->f;
            }
          }
      const double lg = std::sqrt(s.area[g]);
      if (best < 0 || bd > 1.0e-3 * lg) fail(MS_EINVAL, tag + "a face of the second zone has no translated partner");
      if (s.pidx[best] != 0 || s.shadow[best] >= 0 || s.shadow[g] >= 0 || s.pidx[g] != 0)
        fail(MS_EINVAL, tag + "a face is matched twice (zones shared between pairs?)");
      double sum[3];
      for (int d = 0; d < 3; d++) sum[d] = s.sf[3 * (size_t)best + d] + s.sf[3 * (size_t)g + d];
      if (std::fabs(s.area[best] - s.area[g]) > 1.0e-6 * s.area[g] ||
          std::sqrt(sum[0] * sum[0] + sum[1] * sum[1] + sum[2] * sum[2]) > 1.0e-6 * s.area[g])
        fail(MS_EINVAL, tag + "partner faces differ in area or are not opposite");
      if (c0in[best] == c0in[g])
        fail(MS_EINVAL, tag + "a cell would neighbour itself (need two cells across the periodic direction)");
      c1v[best] = c0in[g];
      s.pidx[best] = (signed char)(k + 1);
      s.shadow[g] = best;
    }
    for (int d = 0; d < 3; d++) s.psh[k][d] = sh[d];
  }
}

static void check_faces(int nnodes, int nf, const long long *fptr, const int *fnodes, const int *c0in,
                        const int *c1orig, int nc) {
  for (int f = 0; f < nf; f++) {
    if (c0in[f] < 1 || c0in[f] > nc || c1orig[f] < 0 || c1orig[f] > nc) fail(MS_EINVAL, "face with invalid cell id");
    for (long long j = fptr[f]; j < fptr[f + 1]; j++)
      if (fnodes[j] < 1 || fnodes[j] > nnodes) fail(MS_EINVAL, "face with invalid node id");
    if (fptr[f + 1] - fptr[f] < 3) fail(MS_EINVAL, "face with fewer than 3 nodes");
  }
}

void periodic_match(int nnodes, const double *x, int nfaces, const long long *fptr, const int *fnodes,
                    const int *c0in, const int *c1orig, const int *zone, int ncells, const ms_ctrl_t *ctrl,
                    int *partner) {
  if (nnodes <= 0 || nfaces <= 0 || ncells <= 0 || !x || !fptr || !fnodes || !c0in || !c1orig || !zone || !ctrl ||
      !partner)
    fail(MS_EINVAL, "ms_periodic_match: empty mesh or null argument");
  check_faces(nnodes, nfaces, fptr, fnodes, c0in, c1orig, ncells);
  System s;
  s.nc = ncells;
  s.nf = nfaces;
  compute_geometry(s, x, nfaces, fptr, fnodes, c0in, c1orig, ncells);
  std::vector<int> c1v;
  join_periodic(s, ctrl, nfaces, c0in, c1orig, zone, c1v);
  for (int f = 0; f < nfaces; f++) partner[f] = -1;
  for (int f = 0; f < nfaces; f++)
    if (s.shadow[f] >= 0) {
      partner[f] = s.shadow[f];
      partner[s.shadow[f]] = f;
    }
}

// the boundary groups as the kernels see them
static void upload_bc(System &s) {
  std::vector<DevBC> dbc(s.nbc);
  for (int i = 0; i < s.nbc; i++) {
    const ms_bcgroup_t &g = s.groups[i];
    DevBC &b = dbc[i];
    b.btype = g.btype;
    b.prfl_axis = g.prfl_axis;
    b.has_lid = g.has_lid;
    b.lid_dir = g.lid_dir;
    for (int d = 0; d < 3; d++) {
      b.uvel[d] = g.uvel[d];
      b.lid_vel[d] = g.lid_vel[d];
    }
    b.pval = g.pval;
    b.uspeed = g.uspeed;
    b.prfl_mean = g.prfl_mean;
    b.prfl_span = g.prfl_span;
    b.prfl_origin = g.prfl_origin;
    b.lid_coord = g.lid_coord;
    b.lid_tol = g.lid_tol;
  }
  s.d_bc.upload(dbc);
}

// flux_ref (simple_run :118-130): rho times the largest boundary speed times
// the boundary area
static void set_flux_ref(System &s) {
  double usc = 0.0;
  for (int i = 0; i < s.nbc; i++) {
    const ms_bcgroup_t &g = s.groups[i];
    usc = std::max(usc, std::max(norm2_ref(g.lid_vel, 3), std::max(norm2_ref(g.uvel, 3), g.uspeed)));
  }
  if (usc <= 0.0) usc = 1.0;
  s.flux_ref = s.ctrl.rho * usc * (s.bnd_area > 0.0 ? s.bnd_area : 1.0);
}

System *create_system(int nnodes, const double *x, int nfaces, const long long *fptr, const int *fnodes,
                      const int *c0in, const int *c1orig, const int *zone, int ncells, const ms_ctrl_t *ctrl,
                      const ms_bcgroup_t *groups) {
  if (nnodes <= 0 || nfaces <= 0 || ncells <= 0 || !x || !fptr || !fnodes || !c0in || !c1orig || !zone || !ctrl)
    fail(MS_EINVAL, "ms_system_create: empty mesh or null argument");
  if (ctrl->nbc <= 0 || ctrl->nbc > 32 || !groups) fail(MS_EINVAL, "ms_system_create: need 1..32 bc groups");
  std::unique_ptr<System> S(new System());
  System &s = *S;
  s.nc = ncells;
  s.nf = nfaces;
  s.nn = nnodes;
  s.nbc = ctrl->nbc;
  s.ctrl = *ctrl;
  {
    double zmx = -DBL_MAX, zmn = DBL_MAX;
    for (int i = 0; i < nnodes; i++) {
      zmx = std::max(zmx, x[3 * (size_t)i + 2]);
      zmn = std::min(zmn, x[3 * (size_t)i + 2]);
    }
    s.zspan = zmx - zmn;
  }
  s.groups.assign(groups, groups + ctrl->nbc);
  const int nc = ncells, nf = nfaces;
  check_faces(nnodes, nf, fptr, fnodes, c0in, c1orig, nc);
  for (auto &g : s.groups)
    if (g.btype < 1 || g.btype > 10) fail(MS_EINVAL, "unsupported boundary type");

  compute_geometry(s, x, nf, fptr, fnodes, c0in, c1orig, nc);

  // periodic pairs join faces; from here on c1in is the joined topology
  std::vector<int> c1v;
  join_periodic(s, ctrl, nf, c0in, c1orig, zone, c1v);
  const int *c1in = c1v.data();
  auto is_shadow = [&](int f) { return s.shadow[f] >= 0; };

  // lf, mod_uns_fields.f90:147-157 (the c1 cell of a joined periodic face
  // is seen at its centre + shift)
  s.lf.assign(nf, 1.0);
  for (int f = 0; f < nf; f++) {
    if (c1in[f] == 0) continue;
    double r0[3], r1[3];
    for (int d = 0; d < 3; d++) {
      r0[d] = s.xf[3 * f + d] - s.xc_nat[3 * (c0in[f] - 1) + d];
      r1[d] = s.xc_nat[3 * (c1in[f] - 1) + d] - s.xf[3 * f + d];
      if (s.pidx[f]) r1[d] = (s.xc_nat[3 * (c1in[f] - 1) + d] + s.psh[s.pidx[f] - 1][d]) - s.xf[3 * f + d];
    }
    double d0 = norm2_ref(r0, 3), d1 = norm2_ref(r1, 3);
    s.lf[f] = d1 / (d0 + d1);
  }

  // ---- bc groups: build_bc -------------------------------------------------
  s.fgrp.assign(nf, 0);
  {
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int k = 0; k < nnodes; k++)
      for (int d = 0; d < 3; d++) {
        lo[d] = std::min(lo[d], x[3 * (size_t)k + d]);
        hi[d] = std::max(hi[d], x[3 * (size_t)k + d]);
      }
    double cr = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
    for (int i = 0; i < s.nbc; i++) {
      ms_bcgroup_t &g = s.groups[i];
      if (g.btype == MS_BC_MASSINLET) g.uspeed = g.mdot / ctrl->rho;
      if (g.has_lid) g.lid_tol = 1.0e-4 * cr;
      if (g.ntbc_plane > 0) g.tbc_tol = 1.0e-4 * cr;
      if (g.btype == MS_BC_VINLET_PARAB && (g.prfl_span <= 0.0 || g.prfl_axis < 1 || g.prfl_axis > 3))
        fail(MS_EINVAL, "velocity-inlet-parabolic needs span > 0 and axis 1|2|3");
      int nb = 0;
      for (int f = 0; f < nf; f++)
        if (c1in[f] == 0 && !is_shadow(f) && zone[f] == g.zone) {
          nb++;
          s.fgrp[f] = i + 1;
        }
      if (nb == 0) fail(MS_EINVAL, "bc zone has no boundary faces");
    }
    for (int f = 0; f < nf; f++)
      if (c1in[f] == 0 && !is_shadow(f) && s.fgrp[f] == 0) fail(MS_EINVAL, "boundary face without bc specification");
  }

  // ---- natural connectivity and reference CSR ---------------------------
  std::vector<int> cfp(nc + 1, 0), cfn, c2p(nc + 1, 0), c2c;
  for (int f = 0; f < nf; f++) {
    if (is_shadow(f)) continue;
    cfp[c0in[f]]++;
    if (c1in[f] > 0) cfp[c1in[f]]++;
  }
  for (int c = 0; c < nc; c++) cfp[c + 1] += cfp[c];
  cfn.resize(cfp[nc]);
  {
    std::vector<int> fill(cfp.begin(), cfp.end() - 1);
    for (int f = 0; f < nf; f++) {
      if (is_shadow(f)) continue;
      cfn[fill[c0in[f] - 1]++] = f;
      if (c1in[f] > 0) cfn[fill[c1in[f] - 1]++] = f;
    }
  }
  for (int f = 0; f < nf; f++)
    if (c1in[f] > 0) {
      c2p[c0in[f]]++;
      c2p[c1in[f]]++;
    }
  for (int c = 0; c < nc; c++) c2p[c + 1] += c2p[c];
  c2c.resize(c2p[nc]);
  {
    std::vector<int> fill(c2p.begin(), c2p.end() - 1);
    for (int f = 0; f < nf; f++)
      if (c1in[f] > 0) {
        c2c[fill[c0in[f] - 1]++] = c1in[f] - 1;
        c2c[fill[c1in[f] - 1]++] = c0in[f] - 1;
      }
  }
  s.ref_row_ptr.assign(nc + 1, 1);
  s.ref_col.clear();
  {
    std::vector<int> row;
    for (int c = 0; c < nc; c++) {
      row.assign(c2c.begin() + c2p[c], c2c.begin() + c2p[c + 1]);
      row.push_back(c);
      std::sort(row.begin(), row.end());
      row.erase(std::unique(row.begin(), row.end()), row.end());
      for (int v : row) s.ref_col.push_back(v + 1);
      s.ref_row_ptr[c + 1] = s.ref_row_ptr[c] + (int)row.size();
    }
  }

  // ---- device numbering -----------------------------------------------------
  const int ppc = ctrl->gpu_ppe_precond, mpc = ctrl->gpu_mom_precond;
  const bool want_schur = ctrl->gpu_ppe_solver == 2;
  const bool want_color = ppc == MS_PC_DILU || mpc == MS_PC_DILU || ppc == MS_PC_MC || mpc == MS_PC_MC || want_schur;
  const bool want_rcm = ctrl->gpu_reorder == 1;
  if (ctrl->gpu_coarse < 0) fail(MS_EINVAL, "gpu_coarse must be >= 0");
  if (ctrl->gpu_ppe_solver < 0 || ctrl->gpu_ppe_solver > 2) fail(MS_EINVAL, "gpu_ppe_solver must be 0, 1 or 2");
  const int nagg = std::min(std::min(ctrl->gpu_coarse, kMaxAgg), nc);
  std::vector<int> color, agg;
  cell_ordering(nc, c2p, c2c, s.xc_nat, want_rcm, want_color, nagg, s.n2o, color, s.ncolors, agg);
  s.o2n.assign(nc, 0);
  for (int i = 0; i < nc; i++) s.o2n[s.n2o[i]] = i;
  s.renumbered = want_color || want_rcm || nagg > 0;

  // device-numbered topology
  s.c0.resize(nf);
  s.c1.resize(nf);
  for (int f = 0; f < nf; f++) {
    s.c0[f] = s.o2n[c0in[f] - 1];
    s.c1[f] = c1in[f] > 0 ? s.o2n[c1in[f] - 1] : -1;
  }
  std::vector<int> rp(nc + 1, 0), ci;
  {
    std::vector<int> row;
    for (int r = 0; r < nc; r++) {
      int o = s.n2o[r];
      row.clear();
      for (int k = c2p[o]; k < c2p[o + 1]; k++) row.push_back(s.o2n[c2c[k]]);
      row.push_back(r);
      std::sort(row.begin(), row.end());
      row.erase(std::unique(row.begin(), row.end()), row.end());
      ci.insert(ci.end(), row.begin(), row.end());
      rp[r + 1] = (int)ci.size();
    }
  }
  csr_to_sell(s.pat, nc, rp, ci);
  if (want_color) {
    s.pat.color_ptr.assign(s.ncolors + 1, 0);
    for (int r = 0; r < nc; r++) s.pat.color_ptr[color[s.n2o[r]] + 1]++;
    for (int c = 0; c < s.ncolors; c++) s.pat.color_ptr[c + 1] += s.pat.color_ptr[c];
  }
  auto find_slot = [&](int r, int c) -> long long {
    const int *b = &ci[rp[r]], *e = &ci[rp[r + 1]];
    const int *f = std::lower_bound(b, e, c);
    return s.pat.slot(r, (int)(f - b));
  };
  s.cf_ptr.assign(nc + 1, 0);
  s.cf_ent.clear();
  s.cf_slot.clear();
  s.diag_slot.assign(nc, 0);
  for (int r = 0; r < nc; r++) {
    int o = s.n2o[r];
    for (int k = cfp[o]; k < cfp[o + 1]; k++) {
      int f = cfn[k];
      int side = (c0in[f] - 1 == o) ? 0 : 1;
      s.cf_ent.push_back(2 * f + side);
      if (c1in[f] > 0) {
        int other = side == 0 ? s.c1[f] : s.c0[f];
        s.cf_slot.push_back(find_slot(r, other));
      } else {
        s.cf_slot.push_back(-1);
      }
    }
    s.cf_ptr[r + 1] = (int)s.cf_ent.size();
    s.diag_slot[r] = s.pat.slot(r, s.pat.diagk[r]);
  }
  s.ref2dev.resize(s.ref_col.size());
  for (int i = 0; i < nc; i++)
    for (int k = s.ref_row_ptr[i] - 1; k < s.ref_row_ptr[i + 1] - 1; k++)
      s.ref2dev[k] = find_slot(s.o2n[i], s.o2n[s.ref_col[k] - 1]);
  s.pat.upload();
  if (ppc == MS_PC_NATURAL || mpc == MS_PC_NATURAL || ppc == MS_PC_MC || mpc == MS_PC_MC) s.pat.build_levels();
  std::vector<int> agg_dev(nc, 0);
  if (nagg > 0) {
    for (int r = 0; r < nc; r++) agg_dev[r] = agg[s.n2o[r]];
    build_coarse(s.pat, agg_dev, nagg);
  }
  if (want_schur) build_schur(s, agg_dev, nagg);

  // pin cell and pressure-Dirichlet flag (ppe_assembly :1612-1642)
  s.pin_cell = s.o2n[0];
  s.has_pdir = 0;
  s.has_outflow = 0;
  for (int f = 0; f < nf; f++)
    if (s.fgrp[f] > 0) {
      int bt = s.groups[s.fgrp[f] - 1].btype;
      if (bt == MS_BC_POUTLET || bt == MS_BC_FARFIELD) s.has_pdir = 1;
      if (bt == MS_BC_OUTFLOW) s.has_outflow = 1;
      if (bt == MS_BC_INTERFACE) s.has_iface = true;
    }
  if (s.has_iface) {   // build_bc allocates the interface arrays zeroed (:325-330)
    for (auto *v : {&s.iface_p, &s.iface_T, &s.iface_Ts, &s.iface_q, &s.iface_slip_C}) v->assign(nf, 0.0);
    s.iface_vel.assign(3 * (size_t)nf, 0.0);
    s.iface_peer_vel.assign(3 * (size_t)nf, 0.0);
    s.iface_T_mode.assign(nf, 0);
    s.iface_dirty = true;
  }
  // boundary area, summed group by group in face order (simple_run :118-130)
  {
    double ab = 0.0;
    for (int i = 0; i < s.nbc; i++)
      for (int f = 0; f < nf; f++)
        if (s.fgrp[f] == i + 1) ab = ab + s.area[f];
    s.bnd_area = ab;
    set_flux_ref(s);
  }

  // ---- device face order ----------------------------------------------------
  s.fd2o.resize(nf);
  for (int f = 0; f < nf; f++) s.fd2o[f] = f;
  {
    auto lo = [&](int f) { return s.c1[f] >= 0 ? std::min(s.c0[f], s.c1[f]) : s.c0[f]; };
    auto hi = [&](int f) { return s.c1[f] >= 0 ? std::max(s.c0[f], s.c1[f]) : s.c0[f]; };
    std::stable_sort(s.fd2o.begin(), s.fd2o.end(), [&](int a, int b) {
      const bool ba = s.c1[a] < 0, bb = s.c1[b] < 0;
      if (ba != bb) return bb;   // interior faces first
      const int la = lo(a), lb = lo(b);
      if (la != lb) return la < lb;
      return hi(a) < hi(b);
    });
  }
  s.fo2d.resize(nf);
  for (int f = 0; f < nf; f++) s.fo2d[s.fd2o[f]] = f;

  // ---- upload mesh data ---------------------------------------------------
  std::vector<double> vol(nc), xc(3 * (size_t)nc);
  for (int r = 0; r < nc; r++) {
    vol[r] = s.vol_nat[s.n2o[r]];
    for (int d = 0; d < 3; d++) xc[3 * r + d] = s.xc_nat[3 * (size_t)s.n2o[r] + d];
  }
  s.face_upload(s.d_c0, s.c0);
  s.face_upload(s.d_c1, s.c1);
  s.face_upload(s.d_fgrp, s.fgrp);
  if (s.nper > 0) s.face_upload(s.d_pidx, s.pidx);
  s.d_cf_ptr.upload(s.cf_ptr);
  {
    std::vector<int> ent(s.cf_ent.size());
    for (size_t k = 0; k < ent.size(); k++) ent[k] = 2 * s.fo2d[s.cf_ent[k] >> 1] + (s.cf_ent[k] & 1);
    s.d_cf_ent.upload(ent);
  }
  s.d_cf_slot.upload(s.cf_slot);
  s.d_diag_slot.upload(s.diag_slot);
  s.face_upload(s.d_xf, s.xf, 3);
  s.face_upload(s.d_sf, s.sf, 3);
  s.face_upload(s.d_area, s.area);
  s.face_upload(s.d_lf, s.lf);
  {
    std::vector<int> oth(s.cf_ent.size());
    for (size_t k = 0; k < oth.size(); k++) {
      const int f = s.cf_ent[k] >> 1;
      oth[k] = s.c1[f] < 0 ? -1 : ((s.cf_ent[k] & 1) == 0 ? s.c1[f] : s.c0[f]);
    }
    s.d_e_oth.upload(oth);
  }
  s.entry_upload(s.d_e_lf, s.lf);
  s.entry_upload(s.d_e_sf, s.sf, 3);
  s.d_vol.upload(vol);
  s.d_xc.upload(xc);
  upload_bc(s);

  // ---- fields: init_fields (mod_uns_fields.f90:86-159) ---------------------
  const size_t n3 = 3 * (size_t)nc;
  for (DBuf<double> *b : {&s.u, &s.gp, &s.u_old, &s.u_old_old, &s.gt, &s.gts, &s.dpp, &s.perm_dir_d})
    b->alloc(n3);
  for (DBuf<double> *b : {&s.p, &s.T, &s.T_s, &s.apc, &s.T_old, &s.T_old_old, &s.T_s_old, &s.T_s_old_old,
                          &s.porosity_d, &s.inertial_d, &s.k_s_d, &s.rho_s_d, &s.cp_s_d, &s.rhs, &s.ap, &s.pp,
                          &s.ucomp, &s.diag8})
    b->alloc(nc);
  s.tmpc.alloc(std::max(nc, 8));   // also holds the reduced scalars (up to 6) of tiny meshes
  s.gu.alloc(9 * (size_t)nc);
  s.psi.alloc(4 * (size_t)nc);
  s.flux.alloc(nf);
  s.af.alloc(nf);
  s.val.alloc(s.pat.nstore);
  s.part.alloc(8 * (size_t)nblocks(std::max(nc, nf)) + 64);
  s.ticket.alloc(4);
  for (DBuf<double> *b : {&s.u, &s.gp, &s.u_old, &s.u_old_old, &s.gt, &s.gts, &s.dpp, &s.perm_dir_d, &s.p, &s.T,
                          &s.T_s, &s.T_old, &s.T_old_old, &s.T_s_old, &s.T_s_old_old, &s.inertial_d, &s.k_s_d,
                          &s.rho_s_d, &s.cp_s_d, &s.rhs, &s.ap, &s.pp, &s.gu, &s.psi, &s.flux, &s.af, &s.val,
                          &s.ucomp, &s.diag8, &s.tmpc})
    b->zero();
  s.ticket.zero();
  std::vector<double> ones(nc, 1.0);
  s.apc.upload(ones);
  s.porosity_d.upload(ones);
  s.porosity.assign(nc, 1.0);
  s.perm_dir.assign(n3, 0.0);
  s.inertial.assign(nc, 0.0);
  s.k_s.assign(nc, 0.0);
  s.rho_s.assign(nc, 0.0);
  s.cp_s.assign(nc, 0.0);
  for (std::vector<double> *v : {&s.disp_l, &s.disp_t, &s.bj_alpha, &s.h_sf, &s.a_sf, &s.q_src}) v->assign(nc, 0.0);
  s.props_dirty = true;
  s.refresh_face_constants();
  s.kry.reset(new Krylov(&s.pat));
  MS_HIP(hipStreamSynchronize(stream()));
  return S.release();
}

// Per-face constants that depend only on geometry, mu, k_cond and the cell
// properties, evaluated exactly as the reference's face loops do
// (momentum_assembly :974-1060, temperature_assembly :1803-1940,
// solid_temperature_assembly :2125-2225).  Porous and solid zones add
// per-face flags: fluid/porous faces (kink face pressure, Beavers-Joseph
// slip), solid/fluid faces (conjugate heat transfer) and solid/solid faces.
void System::refresh_face_constants() {
  if ((props_dirty || iface_dirty) && (nforks.load() > 0 || base))
    fail(MS_EINVAL, "the zone properties or interface data of a system changed while sweep cases share its "
                    "face constants");
  if (iface_dirty) {   // interface data: device copies; thermal face modes depend on them
    face_upload(d_iface_vel, iface_vel, 3);
    face_upload(d_iface_p, iface_p);
    face_upload(d_iface_slip_C, iface_slip_C);
    face_upload(d_iface_peer_vel, iface_peer_vel, 3);
    iface_dirty = false;
    props_dirty = true;
  }
  if (!props_dirty) return;
  const double POROSITY_FLOOR = 1.0e-8;
  const ms_ctrl_t &c = ctrl;
  const bool ltne = c.thermal_model != 0, unified = c.porous_unified != 0;
  auto is_porous = [&](int o) {
    return perm_dir[3 * (size_t)o] > 0.0 || perm_dir[3 * (size_t)o + 1] > 0.0 || perm_dir[3 * (size_t)o + 2] > 0.0 ||
           inertial[o] > 0.0;
  };
  auto ztype = [&](int o) { return cztype.empty() ? (int)CZ_FLUID : cztype[o]; };
  auto is_solid = [&](int o) { return ztype(o) == CZ_SOLID; };
  has_solid = false;
  for (int o = 0; o < nc && !cztype.empty(); o++) has_solid = has_solid || cztype[o] == CZ_SOLID;
  has_disp = false;
  has_qsrc = false;
  for (int o = 0; o < nc; o++) {
    has_disp = has_disp || disp_l[o] > 0.0 || disp_t[o] > 0.0;
    has_qsrc = has_qsrc || q_src[o] != 0.0;
  }
  const bool solid_phase = ltne || has_solid || solid_consts;
  // the c1 cell of a joined periodic face sits at xc + shift
  auto xc1 = [&](int f, int o, int d) {
    return pidx[f] ? xc_nat[3 * (size_t)o + d] + psh[pidx[f] - 1][d] : xc_nat[3 * (size_t)o + d];
  };
  std::vector<double> dn(nf, 0.0), dvs(nf, 0.0), Dm(nf, 0.0), muf(nf, 0.0), tcor(3 * (size_t)nf, 0.0),
      dw(nf, 0.0), Dmb(nf, 0.0), Dt(nf, 0.0), kfv(nf, 0.0), Dtb(nf, 0.0), tv(nf, 0.0), qv(nf, 0.0),
      tvs(nf, 0.0);
  std::vector<signed char> tm(nf, 0), tms(nf, 0), ff(nf, 0);
  std::vector<double> Cbj(nf, 0.0), kch(nf, 0.0), Ds(nf, 0.0), ksfv(nf, 0.0), Dsb(nf, 0.0);
  bc_energy = false;
  has_kink = false;
  has_bj = false;
  bool any_flag = false;
  for (int f = 0; f < nf; f++) {
    const int a = n2o[c0[f]];
    const double *sfp = &sf[3 * (size_t)f];
    if (shadow[f] >= 0) continue;   // replaced by its joined periodic face
    if (c1[f] >= 0) {
      const int b = n2o[c1[f]];
      double dvec[3];
      for (int d = 0; d < 3; d++) dvec[d] = xc1(f, b, d) - xc_nat[3 * (size_t)a + d];
      dn[f] = norm2_ref(dvec, 3);
      dvs[f] = dot3(dvec, sfp);
      const double sd = dot3(sfp, dvec);
      const double mu0 = c.mu / std::max(porosity[a], POROSITY_FLOOR);
      const double mu1 = c.mu / std::max(porosity[b], POROSITY_FLOOR);
      const double l = lf[f];
      muf[f] = unified ? 1.0 / ((1.0 - l) / mu0 + l / mu1) : l * mu0 + (1.0 - l) * mu1;
      double k0, k1;
      if (ltne) {
        k0 = porosity[a] * c.k_cond;
        k1 = porosity[b] * c.k_cond;
      } else {
        k0 = porosity[a] * c.k_cond + (1.0 - porosity[a]) * k_s[a];
        k1 = porosity[b] * c.k_cond + (1.0 - porosity[b]) * k_s[b];
      }
      double kf = l * k0 + (1.0 - l) * k1;
      kfv[f] = kf;
      kf = kf + (l * 0.0 + (1.0 - l) * 0.0);   // + dispersion, 0 without dispersivities
      if (c.nonorth_corr > 0.0) {
        Dm[f] = muf[f] * (area[f] * area[f]) / sd;
        Dt[f] = kf * (area[f] * area[f]) / sd;
        for (int d = 0; d < 3; d++) tcor[3 * (size_t)f + d] = sfp[d] - dvec[d] * ((area[f] * area[f]) / sd);
      } else {
        Dm[f] = muf[f] * area[f] / dn[f];
        Dt[f] = kf * area[f] / dn[f];
      }
      // kink face pressure at fluid/porous faces (:1111-1113, fields :191-193)
      if (!unified && is_porous(a) != is_porous(b)) {
        ff[f] |= FF_KINK;
        has_kink = true;
      }
      // Beavers-Joseph slip (:1031-1059): static coefficient C per face
      if (!unified && ztype(a) != ztype(b)) {
        const int cp = ztype(a) == CZ_POROUS ? a : b, cf = cp == a ? b : a;
        const double bja = bj_alpha[cp];
        if (bja > 0.0) {
          double nh[3];
          for (int d = 0; d < 3; d++) nh[d] = sfp[d] / area[f];
          double Knn = 0.0;
          for (int d = 0; d < 3; d++) Knn = Knn + perm_dir[3 * (size_t)cp + d] * (nh[d] * nh[d]);
          if (Knn > 0.0) {
            double r[3];
            for (int d = 0; d < 3; d++) r[d] = xf[3 * (size_t)f + d] - (cf == b ? xc1(f, b, d) : xc_nat[3 * (size_t)a + d]);
            const double dPf = norm2_ref(r, 3);
            Cbj[f] = c.mu * area[f] * (bja / std::sqrt(Knn)) / (1.0 + bja * dPf / std::sqrt(Knn));
            ff[f] |= FF_BJ;
            has_bj = true;
          }
        }
      }
      // solid zones: conjugate heat transfer faces and solid/solid faces
      const bool s0 = is_solid(a), s1 = is_solid(b);
      if (s0 || s1) {
        if (s0 && s1) {
          ff[f] |= FF_SOLID2;
        } else {
          const int cs = s0 ? a : b, cfl = s0 ? b : a;
          double rf[3], rs[3];
          for (int d = 0; d < 3; d++) {
            rf[d] = xf[3 * (size_t)f + d] - (cfl == b ? xc1(f, b, d) : xc_nat[3 * (size_t)a + d]);
            rs[d] = xf[3 * (size_t)f + d] - (cs == b ? xc1(f, b, d) : xc_nat[3 * (size_t)a + d]);
          }
          const double k1c = porosity[cfl] * c.k_cond, k2c = k_s[cs];
          kch[f] = area[f] / (norm2_ref(rf, 3) / std::max(k1c, 1.0e-30) + norm2_ref(rs, 3) / std::max(k2c, 1.0e-30));
          ff[f] |= FF_CHT;
        }
      }
      // solid-phase diffusion (:2161-2175)
      if (solid_phase) {
        const double ks0 = (1.0 - porosity[a]) * k_s[a], ks1 = (1.0 - porosity[b]) * k_s[b];
        ksfv[f] = l * ks0 + (1.0 - l) * ks1;
        Ds[f] = c.nonorth_corr > 0.0 ? ksfv[f] * (area[f] * area[f]) / sd : ksfv[f] * area[f] / dn[f];
      }
      if (ff[f]) any_flag = true;
    } else {
      double r[3];
      for (int d = 0; d < 3; d++) r[d] = xf[3 * (size_t)f + d] - xc_nat[3 * (size_t)a + d];
      dw[f] = norm2_ref(r, 3);
      Dmb[f] = (c.mu / std::max(porosity[a], POROSITY_FLOOR)) * area[f] / dw[f];
      double kf = ltne ? porosity[a] * c.k_cond : porosity[a] * c.k_cond + (1.0 - porosity[a]) * k_s[a];
      kf = kf + 0.0;
      Dtb[f] = kf * area[f] / dw[f];
      // bc_face_T (mod_uns_bc.f90:560-703): static per face
      const ms_bcgroup_t &g = groups[fgrp[f] - 1];
      int tt = 0;
      double tvv = 0.0, qvv = 0.0;
      if (g.btype == MS_BC_WALL || g.btype == MS_BC_POUTLET || g.btype == MS_BC_OUTFLOW) {
        tt = g.ttype;
        tvv = g.tval;
        qvv = g.qval;
        for (int k = 0; k < g.ntbc_plane; k++)
          if (std::fabs(xf[3 * (size_t)f + g.tbc_pdir[k] - 1] - g.tbc_pcoord[k]) <= g.tbc_tol) {
            tt = g.tbc_pttype[k];
            tvv = g.tbc_ptval[k];
            qvv = g.tbc_pqval[k];
            break;
          }
        if (tt != 1 && tt != 2) tt = 0;
      } else if (g.btype == MS_BC_VINLET || g.btype == MS_BC_VINLET_PARAB || g.btype == MS_BC_MASSINLET) {
        tt = 1;
        tvv = g.tval;
      } else if (g.btype == MS_BC_INTERFACE) {   // :671-701
        if (iface_T_mode[f] == 1) {
          tt = 2;
          qvv = iface_q[f];
        } else {
          tt = 1;
          tvv = iface_T[f];
        }
        if (tt == 1 && iface_T_ltne) tvs[f] = iface_Ts[f];
      }
      tm[f] = (signed char)tt;
      tv[f] = tt == 1 ? tvv : 0.0;
      if (g.btype != MS_BC_INTERFACE || !(tt == 1 && iface_T_ltne)) tvs[f] = tv[f];
      qv[f] = tt == 2 ? qvv : 0.0;
      if ((tt == 1 && tvv != 0.0) || (tt == 2 && qvv != 0.0)) bc_energy = true;
      // the solid phase sees inlets as adiabatic (:2187-2193)
      const bool inlet = g.btype == MS_BC_VINLET || g.btype == MS_BC_VINLET_PARAB || g.btype == MS_BC_MASSINLET;
      tms[f] = inlet ? 0 : (signed char)tt;
      if (solid_phase) Dsb[f] = ((1.0 - porosity[a]) * k_s[a]) * area[f] / dw[f];
    }
  }
  has_face_flags = any_flag;
  face_upload(d_dn, dn);
  face_upload(d_dvs, dvs);
  face_upload(d_Dm, Dm);
  face_upload(d_muf, muf);
  face_upload(d_tcor, tcor, 3);
  entry_upload(d_e_Dm, Dm);
  entry_upload(d_e_muf, muf);
  entry_upload(d_e_tcor, tcor, 3);
  face_upload(d_dw, dw);
  face_upload(d_Dmb, Dmb);
  face_upload(d_Dt, Dt);
  face_upload(d_kf, kfv);
  face_upload(d_Dtb, Dtb);
  face_upload(d_tmode, tm);
  face_upload(d_tval, tv);
  face_upload(d_qval, qv);
  if (any_flag) {
    face_upload(d_fflag, ff);
    face_upload(d_Cbj, Cbj);
    face_upload(d_kch, kch);
  }
  // per-cell arrays the thermal kernels read, in device order
  auto dev_upload = [&](const std::vector<double> &h, DBuf<double> &d) {
    std::vector<double> v(nc);
    for (int r = 0; r < nc; r++) v[r] = h[n2o[r]];
    d.upload(v);
  };
  if (solid_phase || has_disp) {
    dev_upload(h_sf, h_sf_d);
    dev_upload(a_sf, a_sf_d);
    dev_upload(q_src, q_src_d);
    dev_upload(disp_l, disp_l_d);
    dev_upload(disp_t, disp_t_d);
  }
  if (has_disp) kdisp.alloc(3 * (size_t)nc);
  if (has_kink) gp_prev.alloc(3 * (size_t)nc);
  if (!cztype.empty()) {
    std::vector<signed char> t(nc);
    for (int r = 0; r < nc; r++) t[r] = (signed char)cztype[n2o[r]];
    d_cztype.upload(t);
  }
  if (solid_phase) {
    face_upload(d_Ds, Ds);
    face_upload(d_ksf, ksfv);
    face_upload(d_Dsb, Dsb);
    face_upload(d_tmode_s, tms);
    if (has_iface && iface_T_ltne) face_upload(d_tval_s, tvs);
  }
  props_dirty = false;
}

System *fork_system(System &b, const std::vector<ms_bcgroup_t> &groups) {
  if (b.base) fail(MS_EINVAL, "fork of a forked system");
  if (b.has_iface) fail(MS_EINVAL, "sweep cases of a system with coupling-interface data are not supported");
  if ((int)groups.size() != b.nbc) fail(MS_EINVAL, "sweep case: boundary group count differs from the system's");
  for (int i = 0; i < b.nbc; i++)
    if (groups[i].btype != b.groups[i].btype) fail(MS_EINVAL, "sweep case: boundary group types differ");
  b.refresh_face_constants();   // no-op unless properties changed since the last run
  std::unique_ptr<System> F(new System);
  System &s = *F;
  s.base = &b;
  s.P = b.P;
  s.nc = b.nc;
  s.nf = b.nf;
  s.nn = b.nn;
  s.nbc = b.nbc;
  s.ctrl = b.ctrl;
  s.groups = groups;
  s.renumbered = b.renumbered;
  s.ncolors = b.ncolors;
  s.nper = b.nper;
  for (int k = 0; k < 4; k++)
    for (int d = 0; d < 3; d++) s.psh[k][d] = b.psh[k][d];
  s.props_dirty = false;
  s.iface_dirty = false;
  s.has_face_flags = b.has_face_flags;
  s.has_solid = b.has_solid;
  s.has_disp = b.has_disp;
  s.has_kink = b.has_kink;
  s.has_bj = b.has_bj;
  s.solid_consts = b.solid_consts;
  s.bc_energy = b.bc_energy;
  s.has_qsrc = b.has_qsrc;
  s.T_nonzero = b.T_nonzero;
  s.ts_order = b.ts_order;
  s.inlet_ramp = b.inlet_ramp;
  s.pval_ramp = b.pval_ramp;
  s.bnd_area = b.bnd_area;
  set_flux_ref(s);
  s.steps_done = b.steps_done;
  s.ts_order_in = b.ts_order_in;
  s.pin_cell = b.pin_cell;
  s.has_pdir = b.has_pdir;
  s.has_outflow = b.has_outflow;
  s.ppe_serial = b.ppe_serial;
  s.zspan = b.zspan;
  s.minv13_ = b.min_vol13();
  // views of the base's read-only device data
  s.d_c0.alias(b.d_c0);
  s.d_c1.alias(b.d_c1);
  s.d_fgrp.alias(b.d_fgrp);
  s.d_cf_ptr.alias(b.d_cf_ptr);
  s.d_cf_ent.alias(b.d_cf_ent);
  s.d_cf_slot.alias(b.d_cf_slot);
  s.d_diag_slot.alias(b.d_diag_slot);
  s.d_e_oth.alias(b.d_e_oth);
  for (auto pr : {std::make_pair(&s.d_e_lf, &b.d_e_lf), std::make_pair(&s.d_e_sf, &b.d_e_sf),
                  std::make_pair(&s.d_e_Dm, &b.d_e_Dm), std::make_pair(&s.d_e_muf, &b.d_e_muf),
                  std::make_pair(&s.d_e_tcor, &b.d_e_tcor), std::make_pair(&s.d_xf, &b.d_xf),
                  std::make_pair(&s.d_sf, &b.d_sf), std::make_pair(&s.d_area, &b.d_area),
                  std::make_pair(&s.d_lf, &b.d_lf), std::make_pair(&s.d_vol, &b.d_vol),
                  std::make_pair(&s.d_xc, &b.d_xc), std::make_pair(&s.d_dn, &b.d_dn),
                  std::make_pair(&s.d_dvs, &b.d_dvs), std::make_pair(&s.d_Dm, &b.d_Dm),
                  std::make_pair(&s.d_muf, &b.d_muf), std::make_pair(&s.d_tcor, &b.d_tcor),
                  std::make_pair(&s.d_dw, &b.d_dw), std::make_pair(&s.d_Dmb, &b.d_Dmb),
                  std::make_pair(&s.d_Dt, &b.d_Dt), std::make_pair(&s.d_kf, &b.d_kf),
                  std::make_pair(&s.d_Dtb, &b.d_Dtb), std::make_pair(&s.d_tval, &b.d_tval),
                  std::make_pair(&s.d_qval, &b.d_qval), std::make_pair(&s.d_Cbj, &b.d_Cbj),
                  std::make_pair(&s.d_kch, &b.d_kch), std::make_pair(&s.d_Ds, &b.d_Ds),
                  std::make_pair(&s.d_ksf, &b.d_ksf), std::make_pair(&s.d_Dsb, &b.d_Dsb),
                  std::make_pair(&s.porosity_d, &b.porosity_d), std::make_pair(&s.perm_dir_d, &b.perm_dir_d),
                  std::make_pair(&s.inertial_d, &b.inertial_d), std::make_pair(&s.k_s_d, &b.k_s_d),
                  std::make_pair(&s.rho_s_d, &b.rho_s_d), std::make_pair(&s.cp_s_d, &b.cp_s_d),
                  std::make_pair(&s.disp_l_d, &b.disp_l_d), std::make_pair(&s.disp_t_d, &b.disp_t_d),
                  std::make_pair(&s.h_sf_d, &b.h_sf_d), std::make_pair(&s.a_sf_d, &b.a_sf_d),
                  std::make_pair(&s.q_src_d, &b.q_src_d)})
    pr.first->alias(*pr.second);
  s.d_tmode.alias(b.d_tmode);
  s.d_pidx.alias(b.d_pidx);
  s.d_fflag.alias(b.d_fflag);
  s.d_tmode_s.alias(b.d_tmode_s);
  s.d_cztype.alias(b.d_cztype);
  upload_bc(s);   // the case's own boundary groups
  // owned: the fields (copies of the base's) and the work arrays
  hipStream_t st = stream();
  auto own = [&](DBuf<double> &d, const DBuf<double> &src) {
    d.alloc(src.n);
    if (src.n) MS_HIP(hipMemcpyAsync(d.p, src.p, src.n * sizeof(double), hipMemcpyDeviceToDevice, st));
  };
  for (auto pr : {std::make_pair(&s.u, &b.u), std::make_pair(&s.p, &b.p), std::make_pair(&s.T, &b.T),
                  std::make_pair(&s.T_s, &b.T_s), std::make_pair(&s.flux, &b.flux), std::make_pair(&s.apc, &b.apc),
                  std::make_pair(&s.u_old, &b.u_old), std::make_pair(&s.u_old_old, &b.u_old_old),
                  std::make_pair(&s.T_old, &b.T_old), std::make_pair(&s.T_old_old, &b.T_old_old),
                  std::make_pair(&s.T_s_old, &b.T_s_old), std::make_pair(&s.T_s_old_old, &b.T_s_old_old),
                  std::make_pair(&s.gp, &b.gp), std::make_pair(&s.gu, &b.gu), std::make_pair(&s.gt, &b.gt),
                  std::make_pair(&s.gts, &b.gts), std::make_pair(&s.psi, &b.psi), std::make_pair(&s.kdisp, &b.kdisp),
                  std::make_pair(&s.gp_prev, &b.gp_prev), std::make_pair(&s.val, &b.val),
                  std::make_pair(&s.rhs, &b.rhs), std::make_pair(&s.ap, &b.ap), std::make_pair(&s.pp, &b.pp),
                  std::make_pair(&s.dpp, &b.dpp), std::make_pair(&s.af, &b.af), std::make_pair(&s.ucomp, &b.ucomp),
                  std::make_pair(&s.part, &b.part), std::make_pair(&s.diag8, &b.diag8),
                  std::make_pair(&s.tmpc, &b.tmpc)})
    own(*pr.first, *pr.second);
  s.ticket.alloc(b.ticket.n);
  s.ticket.zero();
  s.kry.reset(new Krylov(s.P));
  if (b.schur) {
    auto R = std::make_unique<SchurSystem>();
    R->n0 = b.schur->n0;
    R->n1 = b.schur->n1;
    R->P = b.schur->P;
    R->kry = std::make_unique<Krylov>(R->P);
    R->val.alloc(R->P->nstore);
    R->rhs.alloc(R->n1);
    R->d.alloc(R->n1);
    R->r64.alloc(s.nc);
    R->part.alloc(nblocks(s.nc) + 64);
    attach_matfree(*R, s.P);
    s.schur = std::move(R);
  }
  MS_HIP(hipStreamSynchronize(st));
  b.nforks++;
  return F.release();
}

DMesh System::dmesh() const {
  DMesh m{};
  m.nc = nc;
  m.nf = nf;
  m.nbc = nbc;
  m.pin_cell = pin_cell;
  m.iface_pdir = has_iface && iface_p_dir;
  m.iface_ppe = m.iface_pdir && ppe_serial;
  m.has_pdir = has_pdir || m.iface_ppe;
  m.iface_slip = iface_slip;
  m.iface_zhang = iface_zhang;
  m.iface_vel = has_iface ? d_iface_vel.p : nullptr;
  m.iface_p = has_iface ? d_iface_p.p : nullptr;
  m.iface_slip_C = has_iface ? d_iface_slip_C.p : nullptr;
  m.iface_peer_vel = has_iface ? d_iface_peer_vel.p : nullptr;
  m.nonorth = ctrl.nonorth_corr > 0.0;
  m.blend_on = ctrl.conv_blend > 0.0;
  m.c0 = d_c0.p;
  m.c1 = d_c1.p;
  m.fgrp = d_fgrp.p;
  m.xf = d_xf.p;
  m.sf = d_sf.p;
  m.area = d_area.p;
  m.lf = d_lf.p;
  m.vol = d_vol.p;
  m.xc = d_xc.p;
  m.cf_ptr = d_cf_ptr.p;
  m.cf_ent = d_cf_ent.p;
  m.pair_n0 = P->color_ptr.size() == 3 ? P->color_ptr[1] : 0;
  m.e_oth = d_e_oth.p;
  m.e_lf = d_e_lf.p;
  m.e_sf = d_e_sf.p;
  m.e_Dm = d_e_Dm.p;
  m.e_muf = d_e_muf.p;
  m.e_tcor = d_e_tcor.p;
  m.cf_slot = d_cf_slot.p;
  m.diag_slot = d_diag_slot.p;
  m.dn = d_dn.p;
  m.dvs = d_dvs.p;
  m.Dm = d_Dm.p;
  m.muf = d_muf.p;
  m.tcor = d_tcor.p;
  m.dw = d_dw.p;
  m.Dmb = d_Dmb.p;
  m.Dt = d_Dt.p;
  m.kf = d_kf.p;
  m.Dtb = d_Dtb.p;
  m.tmode = d_tmode.p;
  m.tval = d_tval.p;
  m.qval = d_qval.p;
  m.bc = d_bc.p;
  m.nper = nper;
  m.pidx = d_pidx.p;
  m.has_face_flags = has_face_flags;
  m.ltne = ctrl.thermal_model != 0;
  m.has_solid = has_solid;
  m.has_disp = has_disp;
  m.solid_phase = solid_phase();
  m.fflag = d_fflag.p;
  m.Cbj = d_Cbj.p;
  m.kch = d_kch.p;
  m.Ds = d_Ds.p;
  m.ksf = d_ksf.p;
  m.Dsb = d_Dsb.p;
  m.tmode_s = d_tmode_s.p;
  m.tval_s = (has_iface && iface_T_ltne && d_tval_s.p) ? d_tval_s.p : d_tval.p;
  for (int k = 0; k < 4; k++)
    for (int d = 0; d < 3; d++) m.psh[k][d] = psh[k][d];
  return m;
}

DFields System::dfields() {
  DFields F{};
  F.u = u.p;
  F.p = p.p;
  F.T = T.p;
  F.T_s = T_s.p;
  F.flux = flux.p;
  F.apc = apc.p;
  F.u_old = u_old.p;
  F.u_old_old = u_old_old.p;
  F.T_old = T_old.p;
  F.T_old_old = T_old_old.p;
  F.T_s_old = T_s_old.p;
  F.T_s_old_old = T_s_old_old.p;
  F.gp = gp.p;
  F.gu = gu.p;
  F.gt = gt.p;
  F.gts = gts.p;
  F.psi = psi.p;
  F.porosity = porosity_d.p;
  F.perm_dir = perm_dir_d.p;
  F.inertial = inertial_d.p;
  F.k_s = k_s_d.p;
  F.rho_s = rho_s_d.p;
  F.cp_s = cp_s_d.p;
  F.disp_l = disp_l_d.p;
  F.disp_t = disp_t_d.p;
  F.h_sf = h_sf_d.p;
  F.a_sf = a_sf_d.p;
  F.q_src = q_src_d.p;
  F.cztype = has_solid ? d_cztype.p : nullptr;
  F.kdisp = kdisp.p;
  F.gp_prev = gp_prev.p;
  return F;
}

size_t System::device_bytes() const {
  size_t b = pat.device_bytes();
  for (const DBuf<double> *v : {&d_xf, &d_sf, &d_area, &d_lf, &d_vol, &d_xc, &d_dn, &d_dvs, &d_Dm, &d_muf, &d_tcor,
                                &d_dw, &d_Dmb, &d_Dt, &d_kf, &d_Dtb, &d_tval, &d_qval, &u, &p, &T, &T_s, &flux, &apc,
                                &u_old, &u_old_old, &T_old, &T_old_old, &T_s_old, &T_s_old_old, &gp, &gu, &gt, &gts,
                                &psi, &porosity_d, &perm_dir_d, &inertial_d, &k_s_d, &rho_s_d, &cp_s_d, &val, &rhs,
                                &ap, &pp, &dpp, &af, &ucomp, &part, &diag8, &tmpc, &disp_l_d, &disp_t_d, &h_sf_d,
                                &a_sf_d, &q_src_d, &kdisp, &gp_prev, &d_Cbj, &d_kch, &d_Ds, &d_ksf, &d_Dsb})
    b += v->bytes();
  b += d_fflag.bytes() + d_tmode_s.bytes() + d_cztype.bytes() + d_iface_vel.bytes() + d_iface_p.bytes() +
       d_iface_slip_C.bytes() + d_iface_peer_vel.bytes() + d_tval_s.bytes();
  b += d_c0.bytes() + d_c1.bytes() + d_fgrp.bytes() + d_pidx.bytes() + d_cf_ptr.bytes() + d_cf_ent.bytes() + d_cf_slot.bytes() +
       d_diag_slot.bytes();
  b += d_e_oth.bytes() + d_e_lf.bytes() + d_e_sf.bytes() + d_e_Dm.bytes() + d_e_muf.bytes() + d_e_tcor.bytes();
  return b;
}

}  // namespace ms
