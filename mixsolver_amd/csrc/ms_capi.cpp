// ms_capi.cpp -- extern "C" entry points of libmixsolver_hip.so.
// Every call converts internal errors into MS_E* return codes; fields cross
// the boundary in the reference's natural cell order and Fortran layouts.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>

#include "ms_serve.hpp"
#include "ms_sweep.hpp"
#include "ms_system.hpp"

namespace ms {
void ensure_runtime(int device);
const char *last_error();
int run_simple(System &s, int nsteps, int *it_done);
int run_piso(System &s, int nsteps, bool resume);
int run_pimple(System &s, int nsteps, bool resume);
void call_gradients(System &s);
void call_momentum(System &s, int comp);
void call_energy(System &s);
void call_solid_energy(System &s);
void call_rhiechow(System &s);
void call_outflow(System &s);
void call_ppe(System &s);
void call_correct(System &s, const double *d_pp);
void cg_jacobi_noalloc(int n, const int *row_ptr, const int *col_idx, const double *val, const double *b, double *x,
                       double tol, int maxit, int *it, double *res, int *ierr);
}  // namespace ms

using namespace ms;

struct ms_system {
  std::unique_ptr<System> s;
};

template <class F>
static int guarded(F f) {
  try {
    f();
    return MS_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  } catch (const std::bad_alloc &) {
    set_error("host out of memory");
    return MS_ENOMEM;
  } catch (...) {
    set_error("unknown error");
    return MS_EINVAL;
  }
}

namespace {
enum Kind { CELL, FACE, PROP };
struct FieldRef {
  DBuf<double> *buf;
  int width;
  Kind kind;
  std::vector<double> *host;   // PROP: natural-order host copy
  bool writable;
};

FieldRef lookup(System &s, const char *name) {
  std::string n(name ? name : "");
  if (n == "u") return {&s.u, 3, CELL, nullptr, true};
  if (n == "p") return {&s.p, 1, CELL, nullptr, true};
  if (n == "T") return {&s.T, 1, CELL, nullptr, true};
  if (n == "T_s") return {&s.T_s, 1, CELL, nullptr, true};
  if (n == "apc") return {&s.apc, 1, CELL, nullptr, true};
  if (n == "u_old") return {&s.u_old, 3, CELL, nullptr, true};
  if (n == "u_old_old") return {&s.u_old_old, 3, CELL, nullptr, true};
  if (n == "T_old") return {&s.T_old, 1, CELL, nullptr, true};
  if (n == "T_old_old") return {&s.T_old_old, 1, CELL, nullptr, true};
  if (n == "T_s_old") return {&s.T_s_old, 1, CELL, nullptr, true};
  if (n == "T_s_old_old") return {&s.T_s_old_old, 1, CELL, nullptr, true};
  if (n == "gp") return {&s.gp, 3, CELL, nullptr, true};
  if (n == "gu") return {&s.gu, 9, CELL, nullptr, true};
  if (n == "gt") return {&s.gt, 3, CELL, nullptr, true};
  if (n == "gts") return {&s.gts, 3, CELL, nullptr, true};
  if (n == "flux") return {&s.flux, 1, FACE, nullptr, true};
  if (n == "porosity") return {&s.porosity_d, 1, PROP, &s.porosity, true};
  if (n == "perm_dir") return {&s.perm_dir_d, 3, PROP, &s.perm_dir, true};
  if (n == "inertial") return {&s.inertial_d, 1, PROP, &s.inertial, true};
  if (n == "k_s") return {&s.k_s_d, 1, PROP, &s.k_s, true};
  if (n == "rho_s") return {&s.rho_s_d, 1, PROP, &s.rho_s, true};
  if (n == "cp_s") return {&s.cp_s_d, 1, PROP, &s.cp_s, true};
  if (n == "disp_l") return {&s.disp_l_d, 1, PROP, &s.disp_l, true};
  if (n == "disp_t") return {&s.disp_t_d, 1, PROP, &s.disp_t, true};
  if (n == "bj_alpha") return {nullptr, 1, PROP, &s.bj_alpha, true};
  if (n == "h_sf") return {&s.h_sf_d, 1, PROP, &s.h_sf, true};
  if (n == "a_sf") return {&s.a_sf_d, 1, PROP, &s.a_sf, true};
  if (n == "q_src") return {&s.q_src_d, 1, PROP, &s.q_src, true};
  fail(MS_EINVAL, "unknown field name: " + n);
}

void to_device_order(const System &s, const double *h, int w, std::vector<double> &out) {
  out.resize((size_t)s.nc * w);
  for (int i = 0; i < s.nc; i++)
    for (int j = 0; j < w; j++) out[(size_t)s.o2n[i] * w + j] = h[(size_t)i * w + j];
}
void from_device_order(const System &s, const std::vector<double> &d, int w, double *h) {
  for (int i = 0; i < s.nc; i++)
    for (int j = 0; j < w; j++) h[(size_t)i * w + j] = d[(size_t)s.o2n[i] * w + j];
}
void download(const DBuf<double> &b, std::vector<double> &h, size_t n) {
  h.resize(n);
  b.download(h.data(), n);
  MS_HIP(hipStreamSynchronize(stream()));
}
System &sys(ms_system *s) {
  if (!s || !s->s) fail(MS_EINVAL, "null system handle");
  return *s->s;
}
}  // namespace

extern "C" {

int ms_init(int device) { return guarded([&] { ensure_runtime(device); }); }
const char *ms_last_error(void) { return last_error(); }
int ms_version(void) { return 100; }
int ms_set_graphs(int on) { return guarded([&] { set_graphs(on != 0); }); }
int ms_set_fused_momentum(int on) { return guarded([&] { set_momentum_fused(on != 0); }); }
int ms_set_schur_matfree(int on) { return guarded([&] { set_schur_matfree(on != 0); }); }
int ms_set_fused_p_update(int on) { return guarded([&] { set_p_fused(on != 0); }); }

int ms_dump_write(const char *path, int ncells, int ts_order, const char *source_mesh, const char *generated,
                  const double *u, const double *p, const double *T, const double *u_old, const double *u_old_old,
                  const double *T_old, const double *T_old_old) {
  return guarded([&] {
    const double *f[7] = {u, p, T, u_old, u_old_old, T_old, T_old_old};
    dump_write(path, ncells, ts_order, source_mesh, generated, f);
  });
}

int ms_dump_read(const char *path, int ncells, double *u, double *p, double *T, double *u_old, double *u_old_old,
                 double *T_old, double *T_old_old, int *ts_order, char *source_mesh, int source_len, int *ier) {
  return guarded([&] {
    if (!ier || !ts_order || !u || !p || !T || !u_old || !u_old_old || !T_old || !T_old_old || ncells <= 0)
      fail(MS_EINVAL, "ms_dump_read: null argument or empty mesh");
    double *const f[7] = {u, p, T, u_old, u_old_old, T_old, T_old_old};
    std::string src;
    *ier = dump_read(path, ncells, f, ts_order, src);
    if (source_mesh && source_len > 0) {
      const size_t n = std::min((size_t)source_len - 1, src.size());
      std::memcpy(source_mesh, src.data(), n);
      source_mesh[n] = '\0';
    }
  });
}

int ms_vtu_write(const char *path, int binary, int nnodes, const double *x, int nfaces, const long long *fptr,
                 const int *fnodes, const int *c0, const int *c1, int ncells, const double *xc, const double *p,
                 const double *u, const double *T, const double *T_s, const double *q_src) {
  return guarded([&] {
    vtu_write(path, binary != 0, nnodes, x, nfaces, fptr, fnodes, c0, c1, ncells, xc, p, u, T, T_s, q_src);
  });
}

int ms_system_create(ms_system **out, int nnodes, const double *x, int nfaces, const long long *fptr,
                     const int *fnodes, const int *c0, const int *c1, const int *zone, int ncells,
                     const ms_ctrl_t *ctrl, const ms_bcgroup_t *groups) {
  return guarded([&] {
    if (!out) fail(MS_EINVAL, "null out handle");
    ensure_runtime(-1);
    std::unique_ptr<ms_system> h(new ms_system());
    h->s.reset(create_system(nnodes, x, nfaces, fptr, fnodes, c0, c1, zone, ncells, ctrl, groups));
    *out = h.release();
  });
}

int ms_system_destroy(ms_system *s) {
  return guarded([&] {
    if (!s) return;
    if (s->s && s->s->nforks.load() > 0) fail(MS_EINVAL, "the system still has sweep cases (destroy the sweep first)");
    MS_HIP(hipStreamSynchronize(stream()));
    delete s;
  });
}

int ms_system_info(ms_system *h, long long *o) {
  return guarded([&] {
    System &s = sys(h);
    o[0] = s.nc;
    o[1] = s.nf;
    o[2] = s.pat.nnz;
    o[3] = s.ncolors;
    o[4] = s.pat.has_levels ? (long long)s.pat.lev_ptr_f.size() - 1 : 0;
    o[5] = (long long)s.device_bytes();
    o[6] = s.pat.nstore;
    o[7] = s.renumbered;
  });
}

int ms_coarse_info(ms_system *h, int *k, int *nchunks, int *agg) {
  return guarded([&] {
    System &s = sys(h);
    const Coarse &C = s.pat.coarse;
    if (k) *k = C.k;
    if (nchunks) *nchunks = C.nchunks;
    if (agg) {
      std::vector<int> dev(s.nc, 0);
      for (int c = 0; c < C.nchunks; c++)
        for (int r = C.chunk_ptr[c]; r < C.chunk_ptr[c + 1]; r++) dev[r] = C.chunk_agg[c];
      for (int i = 0; i < s.nc; i++) agg[i] = dev[s.o2n[i]];
    }
  });
}

int ms_periodic_match(int nnodes, const double *x, int nfaces, const long long *fptr, const int *fnodes,
                      const int *c0, const int *c1, const int *zone, int ncells, const ms_ctrl_t *ctrl,
                      int *partner) {
  return guarded([&] { periodic_match(nnodes, x, nfaces, fptr, fnodes, c0, c1, zone, ncells, ctrl, partner); });
}

int ms_coarse_aggregates(int n, const double *xc, int k, int *agg) {
  return guarded([&] {
    if (n <= 0 || !xc || !agg || k < 1 || k > n) fail(MS_EINVAL, "ms_coarse_aggregates: need 1 <= k <= n");
    std::vector<double> v(xc, xc + 3 * (size_t)n);
    std::vector<int> a = rcb_aggregates(n, v, k);
    std::copy(a.begin(), a.end(), agg);
  });
}

int ms_cell_ordering(int n, const int *row_ptr, const int *col_idx, const double *xc, int rcm, int colored,
                     int coarse, int *n2o, int *color, int *ncolors) {
  return guarded([&] {
    if (n <= 0 || !row_ptr || !col_idx || !n2o || !color || !ncolors) fail(MS_EINVAL, "ms_cell_ordering: bad arguments");
    if (coarse < 0 || coarse > n || (coarse > 0 && !xc)) fail(MS_EINVAL, "ms_cell_ordering: 0 <= coarse <= n, xc needed");
    std::vector<int> ptr(n + 1, 0), adj;
    for (int i = 0; i < n; i++) {
      for (int k = row_ptr[i] - 1; k < row_ptr[i + 1] - 1; k++) {
        const int j = col_idx[k] - 1;
        if (j < 0 || j >= n) fail(MS_EINVAL, "ms_cell_ordering: column index out of range");
        if (j != i) adj.push_back(j);
      }
      ptr[i + 1] = (int)adj.size();
    }
    std::vector<double> v;
    if (xc) v.assign(xc, xc + 3 * (size_t)n);
    std::vector<int> o, c, a;
    int nc = 0;
    cell_ordering(n, ptr, adj, v, rcm != 0, colored != 0, std::min(coarse, kMaxAgg), o, c, nc, a);
    std::copy(o.begin(), o.end(), n2o);
    std::copy(c.begin(), c.end(), color);
    *ncolors = nc;
  });
}

int ms_set_field(ms_system *h, const char *name, const double *host, long long n) {
  return guarded([&] {
    System &s = sys(h);
    FieldRef f = lookup(s, name);
    if (!host) fail(MS_EINVAL, "null host array");
    if (f.kind == PROP && s.nforks.load() > 0)
      fail(MS_EINVAL, std::string(name) + ": zone properties cannot change while sweep cases share this system");
    if (f.kind == FACE) {
      if (n != s.nf) fail(MS_EINVAL, "face field size mismatch");
      s.face_upload(*f.buf, std::vector<double>(host, host + n));
    } else {
      if (n != (long long)s.nc * f.width) fail(MS_EINVAL, "cell field size mismatch");
      std::vector<double> d;
      to_device_order(s, host, f.width, d);
      if (f.buf) f.buf->upload(d);
      if (f.kind == PROP) {
        f.host->assign(host, host + n);
        s.props_dirty = true;
      }
      std::string nm(name);
      if (nm == "T" || nm == "T_old" || nm == "T_old_old" || nm == "T_s" || nm == "T_s_old" || nm == "T_s_old_old")
        for (long long i = 0; i < n; i++)
          if (host[i] != 0.0) {
            s.T_nonzero = true;
            break;
          }
    }
    MS_HIP(hipStreamSynchronize(stream()));
  });
}

int ms_get_field(ms_system *h, const char *name, double *host, long long n) {
  return guarded([&] {
    System &s = sys(h);
    if (!host) fail(MS_EINVAL, "null host array");
    std::string nm(name ? name : "");
    // geometry (natural order, read-only)
    auto copy = [&](const std::vector<double> &v) {
      if ((size_t)n != v.size()) fail(MS_EINVAL, "geometry size mismatch");
      std::memcpy(host, v.data(), v.size() * sizeof(double));
    };
    if (nm == "xf") return copy(s.xf);
    if (nm == "sf") return copy(s.sf);
    if (nm == "area") return copy(s.area);
    if (nm == "lf") return copy(s.lf);
    if (nm == "vol") return copy(s.vol_nat);
    if (nm == "xc") return copy(s.xc_nat);
    FieldRef f = lookup(s, name);
    std::vector<double> d;
    if (f.kind == PROP) {
      if (n != (long long)s.nc * f.width) fail(MS_EINVAL, "cell field size mismatch");
      std::memcpy(host, f.host->data(), (size_t)n * sizeof(double));
    } else if (f.kind == FACE) {
      if (n != s.nf) fail(MS_EINVAL, "face field size mismatch");
      download(*f.buf, d, s.nf);
      for (int g = 0; g < s.nf; g++) host[g] = d[s.fo2d[g]];   // device face order -> face order
      // a periodic shadow face carries its joined face's flux, oriented out of its own cell
      for (int g = 0; g < s.nf; g++)
        if (s.shadow[g] >= 0) host[g] = -host[s.shadow[g]];
    } else {
      if (n != (long long)s.nc * f.width) fail(MS_EINVAL, "cell field size mismatch");
      download(*f.buf, d, (size_t)s.nc * f.width);
      from_device_order(s, d, f.width, host);
    }
  });
}

int ms_set_cell_types(ms_system *h, const int *cztype) {
  return guarded([&] {
    System &s = sys(h);
    if (s.nforks.load() > 0) fail(MS_EINVAL, "cell types cannot change while sweep cases share this system");
    if (!cztype) {
      s.cztype.clear();
    } else {
      for (int i = 0; i < s.nc; i++)
        if (cztype[i] < 1 || cztype[i] > 3) fail(MS_EINVAL, "cell type must be 1 (fluid), 2 (porous) or 3 (solid)");
      s.cztype.assign(cztype, cztype + s.nc);
    }
    s.props_dirty = true;
  });
}

int ms_set_ts_order(ms_system *h, int t) {
  return guarded([&] {
    if (t < 0 || t > 2) fail(MS_EINVAL, "ts_order must be 0, 1 or 2");
    sys(h).ts_order = t;
  });
}

int ms_set_interface(ms_system *h, const char *what, const double *v, long long n) {
  return guarded([&] {
    System &s = sys(h);
    if (!what || !v) fail(MS_EINVAL, "null argument");
    if (!s.has_iface) fail(MS_EINVAL, "ms_set_interface: the system has no coupling-interface faces");
    const std::string w(what);
    std::vector<double> *dst = nullptr;
    size_t width = 1;
    if (w == "vel") dst = &s.iface_vel, width = 3;
    else if (w == "p") dst = &s.iface_p;
    else if (w == "T") dst = &s.iface_T;
    else if (w == "Ts") dst = &s.iface_Ts;
    else if (w == "q") dst = &s.iface_q;
    else if (w == "slip_C") dst = &s.iface_slip_C;
    else if (w == "peer_vel") dst = &s.iface_peer_vel, width = 3;
    else if (w != "T_mode") fail(MS_EINVAL, "ms_set_interface: unknown item " + w);
    if (n != (long long)(width * s.nf)) fail(MS_EINVAL, "ms_set_interface: " + w + " needs one value per face");
    if (w == "T_mode") {
      for (int f = 0; f < s.nf; f++) {
        if (v[f] != 0.0 && v[f] != 1.0) fail(MS_EINVAL, "ms_set_interface: T_mode values are 0 or 1");
        s.iface_T_mode[f] = (int)v[f];
      }
    } else {
      dst->assign(v, v + n);
    }
    s.iface_dirty = true;
  });
}

int ms_set_interface_modes(ms_system *h, int p_dir, int slip, int zhang, int T_ltne) {
  return guarded([&] {
    System &s = sys(h);
    s.iface_p_dir = p_dir != 0;
    s.iface_slip = slip != 0;
    s.iface_zhang = zhang != 0;
    s.iface_T_ltne = T_ltne != 0;
    s.props_dirty = true;
  });
}

int ms_nan_scan(ms_system *h, int *field, int *cell, double *val, int *kind, int *ifin) {
  return guarded([&] {
    if (!field || !cell || !val || !kind || !ifin) fail(MS_EINVAL, "null output");
    System &s = sys(h);
    std::vector<double> d, u(3 * (size_t)s.nc), p(s.nc), T(s.nc), Ts(s.nc);
    download(s.u, d, 3 * (size_t)s.nc);
    from_device_order(s, d, 3, u.data());
    download(s.p, d, s.nc);
    from_device_order(s, d, 1, p.data());
    download(s.T, d, s.nc);
    from_device_order(s, d, 1, T.data());
    download(s.T_s, d, s.nc);
    from_device_order(s, d, 1, Ts.data());
    *field = -1, *cell = 0, *val = 0.0, *kind = 0, *ifin = 0;
    for (int c = 0; c < s.nc && !*ifin; c++) {
      const double v[6] = {u[3 * (size_t)c], u[3 * (size_t)c + 1], u[3 * (size_t)c + 2], p[c], T[c], Ts[c]};
      for (int k = 0; k < 6; k++)
        if (!std::isfinite(v[k])) {
          *field = k, *cell = c + 1, *val = v[k], *kind = std::isnan(v[k]) ? 1 : 2, *ifin = 1;
          break;
        }
    }
  });
}

int ms_get_ts_order(ms_system *h, int *t) {
  return guarded([&] {
    if (!t) fail(MS_EINVAL, "null output");
    *t = sys(h).ts_order;
  });
}

int ms_get_topology(ms_system *h, int *row_ptr, int *col_idx, int *fgrp) {
  return guarded([&] {
    System &s = sys(h);
    if (row_ptr) std::copy(s.ref_row_ptr.begin(), s.ref_row_ptr.end(), row_ptr);
    if (col_idx) std::copy(s.ref_col.begin(), s.ref_col.end(), col_idx);
    if (fgrp) std::copy(s.fgrp.begin(), s.fgrp.end(), fgrp);
  });
}

int ms_simple_run(ms_system *h, int nsteps, int *it_done, int *ier) {
  return guarded([&] {
    int r = run_simple(sys(h), nsteps, it_done);
    if (ier) *ier = r;
  });
}

int ms_piso_run(ms_system *h, int nsteps, int *ier) {
  return guarded([&] {
    int r = run_piso(sys(h), nsteps, false);
    if (ier) *ier = r;
  });
}

int ms_piso_continue(ms_system *h, int nsteps, int *ier) {
  return guarded([&] {
    int r = run_piso(sys(h), nsteps, true);
    if (ier) *ier = r;
  });
}

int ms_pimple_run(ms_system *h, int nsteps, int *ier) {
  return guarded([&] {
    int r = run_pimple(sys(h), nsteps, false);
    if (ier) *ier = r;
  });
}

int ms_pimple_continue(ms_system *h, int nsteps, int *ier) {
  return guarded([&] {
    int r = run_pimple(sys(h), nsteps, true);
    if (ier) *ier = r;
  });
}

int ms_get_history(ms_system *h, double *out, int maxrows, int *nrows) {
  return guarded([&] {
    System &s = sys(h);
    int rows = (int)(s.history.size() / 8);
    if (nrows) *nrows = rows;
    if (out) std::memcpy(out, s.history.data(), std::min(rows, maxrows) * 8 * sizeof(double));
  });
}

int ms_solver_stats(ms_system *h, long long *out8, int reset) {
  return guarded([&] {
    System &s = sys(h);
    if (out8) {
      out8[0] = s.st_ppe_solves;
      out8[1] = s.st_it32;
      out8[2] = s.st_it64;
      out8[3] = s.st_fallback;
      const bool st_full = s.kry && s.kry->fp32_stalled();
      const bool st_red = s.schur && s.schur->kry && s.schur->kry->fp32_stalled();
      out8[4] = st_full || st_red ? 1 : 0;
      out8[5] = stall_window();
      out8[6] = s.st_mom_solves;
      out8[7] = s.st_mom_its;
    }
    if (reset) s.st_ppe_solves = s.st_it32 = s.st_it64 = s.st_fallback = s.st_mom_solves = s.st_mom_its = 0;
  });
}

int ms_set_stall_window(int iters) {
  return guarded([&] {
    if (iters < 0) fail(MS_EINVAL, "ms_set_stall_window: iters >= 0 (0 = default 1024)");
    set_stall_window(iters);
  });
}

int ms_spmv_probe(int on, long long *launches, double *total_ms, long long *rows) {
  return guarded([&] {
    ensure_runtime(-1);
    if (launches && total_ms && rows) read_spmv_probe(launches, total_ms, rows);
    if (on >= 0) set_spmv_probe(on != 0);
  });
}

// ---- per-call level --------------------------------------------------------
static void matrix_out(System &s, double *val, double *rhs, double *ap) {
  std::vector<double> d;
  if (val) {
    download(s.val, d, s.pat.nstore);
    for (size_t k = 0; k < s.ref2dev.size(); k++) val[k] = d[s.ref2dev[k]];
  }
  if (rhs) {
    download(s.rhs, d, s.nc);
    from_device_order(s, d, 1, rhs);
  }
  if (ap) {
    download(s.ap, d, s.nc);
    from_device_order(s, d, 1, ap);
  }
}

int ms_compute_gradients(ms_system *h) { return guarded([&] { call_gradients(sys(h)); }); }
int ms_momentum_assembly(ms_system *h, int comp, double *val, double *rhs, double *ap) {
  return guarded([&] {
    if (comp < 1 || comp > 3) fail(MS_EINVAL, "comp must be 1..3");
    call_momentum(sys(h), comp - 1);
    matrix_out(sys(h), val, rhs, ap);
  });
}
int ms_temperature_assembly(ms_system *h, double *val, double *rhs, double *ap) {
  return guarded([&] {
    call_energy(sys(h));
    matrix_out(sys(h), val, rhs, ap);
  });
}
int ms_solid_temperature_assembly(ms_system *h, double *val, double *rhs, double *ap) {
  return guarded([&] {
    call_solid_energy(sys(h));
    matrix_out(sys(h), val, rhs, ap);
  });
}
int ms_flux_rhiechow(ms_system *h) { return guarded([&] { call_rhiechow(sys(h)); }); }
int ms_outflow_mass_rescale(ms_system *h) { return guarded([&] { call_outflow(sys(h)); }); }
int ms_ppe_assembly(ms_system *h, double *val, double *rhs) {
  return guarded([&] {
    call_ppe(sys(h));
    matrix_out(sys(h), val, rhs, nullptr);
  });
}
int ms_correct_fields(ms_system *h, const double *pp, double *dpp) {
  return guarded([&] {
    System &s = sys(h);
    std::vector<double> d;
    to_device_order(s, pp, 1, d);
    s.pp.upload(d);
    call_correct(s, s.pp.p);
    if (dpp) {
      download(s.dpp, d, 3 * (size_t)s.nc);
      from_device_order(s, d, 3, dpp);
    }
  });
}

// ---- solvers on a caller CSR ---------------------------------------------------
static void solve_csr(int n, const int *row_ptr, const int *col_idx, const double *val, const double *b, double *x,
                      double tol, int maxit, int solver, int precond, int precision, double inner_tol, int *it,
                      double *res, int *ierr) {
  if (n <= 0 || !row_ptr || !col_idx || !val || !b || !x) fail(MS_EINVAL, "bad solver arguments");
  ensure_runtime(-1);
  std::vector<int> rp(n + 1), ci(row_ptr[n] - 1);
  for (int i = 0; i <= n; i++) rp[i] = row_ptr[i] - 1;
  for (size_t k = 0; k < ci.size(); k++) {
    ci[k] = col_idx[k] - 1;
    if (ci[k] < 0 || ci[k] >= n) fail(MS_EINVAL, "column index out of range");
  }
  for (int i = 0; i < n; i++)
    for (int k = rp[i] + 1; k < rp[i + 1]; k++)
      if (ci[k] <= ci[k - 1]) fail(MS_EINVAL, "CSR columns must be strictly ascending per row");
  SellPattern pat;
  csr_to_sell(pat, n, rp, ci);
  if (precond == MS_PC_DILU) {
    // colour classes in natural order (rows renumbering is the caller's choice)
    fail(MS_EINVAL, "DILU needs a colour-blocked ordering; use a system with gpu_ppe_precond = dilu");
  }
  pat.upload();
  if (precond == MS_PC_NATURAL) pat.build_levels();
  std::vector<double> vs(pat.nstore, 0.0);
  for (int r = 0; r < n; r++)
    for (int k = 0; k < pat.rowlen[r]; k++) vs[pat.slot(r, k)] = val[rp[r] + k];
  DBuf<double> dv, db, dx;
  dv.upload(vs);
  db.upload(b, n);
  dx.upload(x, n);
  Krylov K(&pat);
  PcKind pc = precond == MS_PC_JACOBI ? PC_JACOBI : (solver == 0 ? PC_NATURAL_IC0 : PC_NATURAL_ILU0);
  SolveResult R;
  if (precision == MS_MIXED) R = K.mixed(dv.p, db.p, dx.p, tol, maxit, pc, solver, inner_tol);
  else if (solver == 0) R = K.pcg(dv.p, db.p, dx.p, tol, maxit, pc);
  else R = K.bicgstab(dv.p, db.p, dx.p, tol, maxit, pc);
  dx.download(x, n);
  MS_HIP(hipStreamSynchronize(stream()));
  if (it) *it = R.it;
  if (res) *res = R.res;
  if (ierr) *ierr = R.ierr;
}

int ms_cg_jacobi(int n, const int *rp, const int *ci, const double *v, const double *b, double *x, double tol,
                 int maxit, int *it, double *res, int *ierr) {
  return guarded([&] { solve_csr(n, rp, ci, v, b, x, tol, maxit, 0, MS_PC_JACOBI, MS_FP64, 0.0, it, res, ierr); });
}
int ms_cg_ic0(int n, const int *rp, const int *ci, const double *v, const double *b, double *x, double tol,
              int maxit, int *it, double *res, int *ierr) {
  return guarded([&] { solve_csr(n, rp, ci, v, b, x, tol, maxit, 0, MS_PC_NATURAL, MS_FP64, 0.0, it, res, ierr); });
}
int ms_bicgstab_ilu0(int n, const int *rp, const int *ci, const double *v, const double *b, double *x, double tol,
                     int maxit, int *it, double *res, int *ierr) {
  return guarded([&] { solve_csr(n, rp, ci, v, b, x, tol, maxit, 1, MS_PC_NATURAL, MS_FP64, 0.0, it, res, ierr); });
}
int ms_cg_jacobi_noalloc(int n, const int *rp, const int *ci, const double *v, const double *b, double *x,
                         double tol, int maxit, int *it, double *res, int *ierr) {
  return guarded([&] {
    if (!it || !res || !ierr) fail(MS_EINVAL, "ms_cg_jacobi_noalloc: null output");
    cg_jacobi_noalloc(n, rp, ci, v, b, x, tol, maxit, it, res, ierr);
  });
}
int ms_solve(int n, const int *rp, const int *ci, const double *v, const double *b, double *x, double tol, int maxit,
             int solver, int precond, int precision, double inner_tol, int *it, double *res, int *ierr) {
  return guarded([&] {
    if (solver < 0 || solver > 1 || precision < 0 || precision > 1) fail(MS_EINVAL, "bad solver selector");
    solve_csr(n, rp, ci, v, b, x, tol, maxit, solver, precond, precision, inner_tol, it, res, ierr);
  });
}

int ms_time_copy(long long bytes, int reps, double *avg_ms) {
  return guarded([&] {
    if (bytes < 16 || reps <= 0) fail(MS_EINVAL, "ms_time_copy: bytes >= 16, reps > 0");
    ensure_runtime(-1);
    const double t = time_copy(bytes, reps);
    if (avg_ms) *avg_ms = t;
  });
}

int ms_time_stream(long long bytes, int kind, int reps, double *avg_ms) {
  return guarded([&] {
    if (bytes < 16 || kind < 0 || kind > 2 || reps <= 0)
      fail(MS_EINVAL, "ms_time_stream: bytes >= 16, kind in {0,1,2}, reps > 0");
    ensure_runtime(-1);
    const double t = time_stream(bytes, kind, reps);
    if (avg_ms) *avg_ms = t;
  });
}

int ms_time_kernel(ms_system *h, int which, int reps, double *avg_ms, double *alg_bytes) {
  return guarded([&] {
    System &s = sys(h);
    if (which < 0 || which > 7 || reps <= 0) fail(MS_EINVAL, "ms_time_kernel: which in 0..7, reps > 0");
    // 4..7: the same four timings on the red-black reduced matrix S
    const bool red = which >= 4;
    if (red && (!s.schur || !s.schur->have_values))
      fail(MS_EINVAL, "ms_time_kernel: no reduced p' matrix yet (gpu_ppe_solver = schur and a solve)");
    const int w = which & 3;
    const bool fp32 = w == 0 || w == 2;
    Krylov &K = red ? *s.schur->kry : *s.kry;
    const SellPattern &P = red ? s.schur->pat : s.pat;
    double t = K.time_spmv(red ? s.schur->val.p : s.val.p, fp32, reps, w >= 2);
    const double sz = fp32 ? 4.0 : 8.0;
    // column index bytes of the format in use: 16-bit offsets plus the int32
    // column of every far entry, or int32 columns
    const double idx = (fp32 && P.has16) ? 2.0 + 4.0 * (double)P.nfar / (double)P.nnz : 4.0;
    if (avg_ms) *avg_ms = t;
    // algorithmic bytes of one SpMV: nnz (value + column index) + one int per row + x and y
    if (alg_bytes) *alg_bytes = (double)P.nnz * (sz + idx) + 4.0 * (P.n + 1) + 2.0 * P.n * sz;
    if (alg_bytes && red && fp32 && K.matfree_on()) {
      // the matrix-free product: A's nonzeros (FP32 weight + column index)
      // and one int per row of A, x read by each stage, the colour-0
      // intermediate (h, x_ref) written and read, q written
      const SellPattern &F = *s.P;
      const double n0 = s.schur->n0, n1 = s.schur->n1;
      *alg_bytes = (double)F.nnz * 4.0 + (F.has16 ? 2.0 * F.nnz + 4.0 * F.nfar : 4.0 * F.nnz) + 4.0 * (F.n + 1) +
                   2.0 * n1 * 4.0 +
                   2.0 * n0 * 8.0 + n1 * 4.0;
    }
  });
}

// ---- serving -------------------------------------------------------------
struct ms_server {
  ms::Server *s = nullptr;
};

int ms_server_create(ms_server **out, const ms_server_opts *o) {
  return guarded([&] {
    if (!out || !o) fail(MS_EINVAL, "ms_server_create: null argument");
    ServeOpts so;
    so.budget = o->hbm_budget;
    so.max_active = o->max_active > 0 ? o->max_active : 64;
    so.precond = o->precond;
    so.coarse = o->coarse;
    so.round_iters = o->round_iters > 0 ? o->round_iters : 16;
    so.delta = o->delta > 0.0 ? o->delta : 0.1;
    std::unique_ptr<ms_server> h(new ms_server());
    h->s = server_create(so);
    *out = h.release();
  });
}

int ms_server_destroy(ms_server *h) {
  return guarded([&] {
    if (!h) return;
    server_destroy(h->s);
    delete h;
  });
}

static ms::Server *srv(ms_server *h) {
  if (!h || !h->s) fail(MS_EINVAL, "null server handle");
  return h->s;
}

int ms_server_add_pattern(ms_server *h, int n, const int *rp, const int *ci, const double *xc, int *pattern) {
  return guarded([&] {
    int id = server_add_pattern(srv(h), n, rp, ci, xc);
    if (pattern) *pattern = id;
  });
}

int ms_server_submit(ms_server *h, int pattern, const double *val, const double *b, double tol, int maxit,
                     long long *ticket) {
  return guarded([&] {
    long long t = server_submit(srv(h), pattern, val, b, tol, maxit);
    if (ticket) *ticket = t;
  });
}

int ms_server_result(ms_server *h, long long ticket, double timeout_ms, double *x, int *it, double *res, int *ierr,
                     double *t_admit, double *t_done, int *ready) {
  return guarded([&] {
    ServeResult r;
    const bool ok = server_result(srv(h), ticket, timeout_ms, x, &r);
    if (ready) *ready = ok;
    if (!ok) return;
    if (it) *it = r.it;
    if (res) *res = r.res;
    if (ierr) *ierr = r.ierr;
    if (t_admit) *t_admit = r.t_admit;
    if (t_done) *t_done = r.t_done;
  });
}

int ms_server_stats(ms_server *h, long long *out8) {
  return guarded([&] {
    if (!out8) fail(MS_EINVAL, "null output");
    server_stats(srv(h), out8);
  });
}

// ---- parameter-sweep server ------------------------------------------------
struct ms_sweep {
  Sweep *s = nullptr;
};

int ms_sweep_create(ms_sweep **sw, ms_system *base, long long budget_bytes, int max_active) {
  return guarded([&] {
    if (!sw) fail(MS_EINVAL, "null output handle");
    *sw = nullptr;
    ensure_runtime(-1);
    auto h = std::make_unique<ms_sweep>();
    h->s = sweep_create(sys(base), budget_bytes, max_active);
    *sw = h.release();
  });
}

int ms_sweep_destroy(ms_sweep *sw) {
  return guarded([&] {
    if (!sw) return;
    sweep_destroy(sw->s);
    delete sw;
  });
}

int ms_sweep_submit(ms_sweep *sw, int group, const double *uvel, double speed, int nsteps, int keep_fields,
                    long long *id) {
  return guarded([&] {
    if (!sw || !sw->s) fail(MS_EINVAL, "null sweep handle");
    const long long i = sweep_submit(sw->s, group, uvel, speed, nsteps, keep_fields);
    if (id) *id = i;
  });
}

int ms_sweep_result(ms_sweep *sw, long long id, double wait_s, int *state, int *steps, int *ier, double *step_ms,
                    double *history, double *times3, long long *stats8, double *u, double *p, double *T,
                    double *T_s) {
  return guarded([&] {
    if (!sw || !sw->s) fail(MS_EINVAL, "null sweep handle");
    SweepCase c;
    const int st = sweep_result(sw->s, id, wait_s, &c);
    if (state) *state = st;
    if (st < 2) return;
    if (st == 3) fail(c.err ? c.err : MS_EINVAL, "sweep case " + std::to_string(id) + ": " + c.msg);
    if (steps) *steps = (int)c.step_ms.size();
    if (ier) *ier = c.ier;
    if (step_ms) std::copy(c.step_ms.begin(), c.step_ms.end(), step_ms);
    if (history) std::copy(c.history.begin(), c.history.end(), history);
    if (times3) {
      times3[0] = c.t_submit;
      times3[1] = c.t_admit;
      times3[2] = c.t_done;
    }
    if (stats8) std::copy(c.stats, c.stats + 8, stats8);
    if (c.keep) {
      if (u) std::copy(c.u.begin(), c.u.end(), u);
      if (p) std::copy(c.p.begin(), c.p.end(), p);
      if (T) std::copy(c.T.begin(), c.T.end(), T);
      if (T_s) std::copy(c.T_s.begin(), c.T_s.end(), T_s);
    }
  });
}

int ms_sweep_stats(ms_sweep *sw, long long *out10) {
  return guarded([&] {
    if (!sw || !sw->s || !out10) fail(MS_EINVAL, "null sweep handle or output");
    sweep_stats(sw->s, out10);
  });
}

// ---- wall-zone post-processing (mod_uns_forces.f90) -----------------------
static int wall_centre(System &s, double *xc, double *yc) {
  double sx = 0.0, sy = 0.0;
  for (int f : s.wall_faces) {
    sx = sx + s.xf[3 * (size_t)f];
    sy = sy + s.xf[3 * (size_t)f + 1];
  }
  const int n = (int)s.wall_faces.size();
  if (n > 0) {
    sx = sx / (double)n;
    sy = sy / (double)n;
  }
  *xc = sx;
  *yc = sy;
  return n;
}

int ms_compute_forces(ms_system *h, double *out14) {
  return guarded([&] {
    if (!out14) fail(MS_EINVAL, "null output");
    wall_forces(sys(h), out14);
  });
}

int ms_cp_distribution(ms_system *h, double *rows6, int maxrows, int *nrows) {
  return guarded([&] {
    System &s = sys(h);
    if (nrows) *nrows = 0;
    double u_inf, p_inf;
    free_stream(s, &u_inf, &p_inf);
    std::vector<double> pu;
    wall_samples(s, pu);
    if (u_inf <= 0.0) return;
    double xcyl, ycyl;
    const int n = wall_centre(s, &xcyl, &ycyl);
    if (nrows) *nrows = n;
    if (!rows6 || n == 0) return;
    if (maxrows < n) fail(MS_EINVAL, "ms_cp_distribution: rows6 holds fewer rows than wall faces");
    const double q = 0.5 * s.ctrl.rho * (u_inf * u_inf);
    const double pi = 3.141592653589793238462643;
    for (int k = 0; k < n; k++) {
      const double *xf = &s.xf[3 * (size_t)s.wall_faces[k]];
      const double th = std::atan2(xf[1] - ycyl, xf[0] - xcyl);
      const double sn = std::sin(th);
      double *r = &rows6[6 * (size_t)k];
      r[0] = th * 180.0 / pi;
      r[1] = (pu[4 * (size_t)k] - p_inf) / q;
      r[2] = 1.0 - 4.0 * (sn * sn);
      r[3] = xf[0];
      r[4] = xf[1];
      r[5] = xf[2];
    }
  });
}

int ms_separation(ms_system *h, double *out5) {
  return guarded([&] {
    if (!out5) fail(MS_EINVAL, "null output");
    System &s = sys(h);
    for (int k = 0; k < 5; k++) out5[k] = 0.0;
    std::vector<double> pu;
    wall_samples(s, pu);
    double xcyl, ycyl;
    const int n = wall_centre(s, &xcyl, &ycyl);
    if (n == 0) return;
    std::vector<double> th(n), ua(n);
    for (int k = 0; k < n; k++) {
      const int f = s.wall_faces[k];
      const double *sf = &s.sf[3 * (size_t)f], *xf = &s.xf[3 * (size_t)f];
      const double nx = sf[0] / norm2_ref(sf, 3), ny = sf[1] / norm2_ref(sf, 3);
      th[k] = std::atan2(xf[1] - ycyl, xf[0] - xcyl);
      ua[k] = pu[4 * (size_t)k + 1] * ny + pu[4 * (size_t)k + 2] * (-nx);
    }
    for (int i = 1; i < n; i++) {   // insertion sort by theta, stable (as the reference)
      const double t = th[i], u = ua[i];
      int k = i - 1;
      while (k >= 0 && th[k] > t) {
        th[k + 1] = th[k];
        ua[k + 1] = ua[k];
        k--;
      }
      th[k + 1] = t;
      ua[k + 1] = u;
    }
    double up = 0.0, tp = 0.0;
    bool ftop = false, fbot = false;
    for (int i = 0; i < n; i++) {
      if (std::fabs(ua[i]) < 1.0e-4) continue;
      if (ua[i] < 0.0 && up > 0.0) {
        const double ts = tp + (0.0 - up) * (th[i] - tp) / (ua[i] - up);
        if (th[i] < 0.0 && !fbot) {
          out5[2] = ts;
          fbot = true;
        } else if (th[i] > 0.0 && !ftop) {
          out5[0] = ts;
          ftop = true;
        }
      }
      up = ua[i];
      tp = th[i];
    }
    out5[1] = ftop ? 1.0 : 0.0;
    out5[3] = fbot ? 1.0 : 0.0;
    out5[4] = (wake_end(s, xcyl) - xcyl) / 1.0;
  });
}

}  // extern "C"
