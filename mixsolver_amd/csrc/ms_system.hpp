// ms_system.hpp -- one mesh + fields + matrices resident on the device.
#pragma once
#include <algorithm>
#include <atomic>
#include <cmath>
#include <memory>

#include "ms_common.hpp"
#include "ms_linalg.hpp"

namespace ms {

// Boundary group as the kernels see it (subset of ms_bcgroup_t).
struct DevBC {
  int btype, prfl_axis, has_lid, lid_dir;
  double uvel[3], pval, uspeed, prfl_mean, prfl_span, prfl_origin, lid_coord, lid_tol, lid_vel[3];
};

// Device views passed by value to the physics kernels.  Cell ids are in
// the device numbering (0-based); face order is the mesh's face order.
struct DMesh {
  int nc, nf, nbc, pin_cell, has_pdir, nonorth, blend_on;
  int pair_n0;                             // > 0: two-colour numbering, rows of colour 0 (cell_row)
  const int *c0, *c1, *fgrp;               // c1 < 0 on boundary faces
  const double *xf, *sf, *area, *lf, *vol, *xc;
  const int *cf_ptr, *cf_ent;              // cell -> faces ascending, ent = 2*face + side
  // the same per cell-face entry k (cf order), so a cell's face loads are
  // adjacent in memory: the other cell (-1 on boundary faces), lf, sf (3),
  // Dm, muf, tcor (3) of the entry's face
  const int *e_oth;
  const double *e_lf, *e_sf, *e_Dm, *e_muf, *e_tcor;
  const long long *cf_slot, *diag_slot;    // matrix slot of (cell, other) / (cell, cell)
  const double *dn, *dvs, *Dm, *muf, *tcor, *dw, *Dmb;   // momentum / PPE face constants
  const double *Dt, *kf, *Dtb;                           // energy face constants
  const signed char *tmode;                // boundary thermal mode 0 neumann0 1 dirichlet 2 flux
  const double *tval, *qval;
  const DevBC *bc;
  // periodic pairs: pidx[f] = k + 1 on a joined face of pair k (0 elsewhere);
  // its c1 cell sits at xc + psh[k] as seen from the face
  int nper;
  const signed char *pidx;
  double psh[4][3];
  // porous / solid zones (null unless the system has them): per-face flags
  // FF_KINK | FF_BJ | FF_CHT | FF_SOLID2, Beavers-Joseph coefficient, CHT
  // conductance, solid-phase face constants
  int has_face_flags, ltne, has_solid, has_disp, solid_phase;
  const signed char *fflag;
  const double *Cbj, *kch, *Ds, *ksf, *Dsb;
  const signed char *tmode_s;             // solid-phase boundary thermal mode
  const double *tval_s;                   // solid-phase Dirichlet values (null: tval)
  // coupling interface (BC_INTERFACE faces): per-face data in face order and
  // the switches; iface_ppe = the PPE treats interface faces as
  // pressure-Dirichlet (serial drivers under the interface pressure model)
  int iface_pdir, iface_ppe, iface_slip, iface_zhang;
  const double *iface_vel, *iface_p, *iface_slip_C, *iface_peer_vel;
};

enum { FF_KINK = 1, FF_BJ = 2, FF_CHT = 4, FF_SOLID2 = 8 };
enum { CZ_FLUID = 1, CZ_POROUS = 2, CZ_SOLID = 3 };

struct DFields {
  double *u, *p, *T, *T_s, *flux, *apc, *u_old, *u_old_old, *T_old, *T_old_old, *T_s_old, *T_s_old_old;
  double *gp, *gu, *gt, *gts, *psi;
  const double *porosity, *perm_dir, *inertial, *k_s, *rho_s, *cp_s;
  const double *disp_l, *disp_t, *h_sf, *a_sf, *q_src;
  const signed char *cztype;               // block type per cell (null: all fluid)
  double *kdisp;                           // (3,nc) dispersion conductivity work array
  const double *gp_prev;                   // gp before the current gradient pass (kink faces)
};

// The red-black reduced p' system of a two-colour system (ms_schur.hip):
// colour-1 rows of the Schur complement S = D1 - B^T D0^-1 B.
struct SchurSystem {
  int n0 = 0, n1 = 0;           // colour-0 rows [0,n0), colour-1 rows [n0,n0+n1)
  SellPattern pat;              // pattern of S, colour-1 local numbering
  const SellPattern *P = &pat;  // the pattern in use (a forked system's points to its base's)
  std::unique_ptr<Krylov> kry;  // PCG on S (Jacobi, coarse space of the colour-1 aggregates)
  DBuf<double> val, rhs, d, r64, part;
  // matrix-free S for the FP32 iterations: weights in A's layout and the
  // colour-0 intermediate (MatFree)
  DBuf<float> mf_w;
  DBuf<float2> mf_hx;
  DBuf<int> mf_agg;   // aggregate of every colour-1 row (the fused p update)
  MatFree mf;
  std::vector<double> hpart;
  bool have_values = false;     // val holds S of the last matrix
};

struct System {
  // sizes and control
  int nc = 0, nf = 0, nn = 0, nbc = 0;
  ms_ctrl_t ctrl{};
  std::vector<ms_bcgroup_t> groups;
  // numbering: dev = o2n[natural], natural = n2o[dev]
  std::vector<int> o2n, n2o;
  bool renumbered = false;
  int ncolors = 0;
  // host topology in device numbering
  std::vector<int> c0, c1, fgrp, cf_ptr, cf_ent;
  // periodic pairs: per face 0 or pair + 1 on joined faces; shadow[f] is the
  // joined face that replaces boundary face f (-1 if f is not a shadow)
  int nper = 0;
  double psh[4][3] = {};
  std::vector<signed char> pidx;
  std::vector<int> shadow;
  std::vector<long long> cf_slot, diag_slot;
  // Device face order: interior faces sorted by their lower device cell
  // (then the higher), boundary faces after them sorted by their cell, so
  // the faces of consecutive device cells sit together in memory and the
  // cell kernels' face loads coalesce.  fd2o[dev] = face, fo2d[face] = dev.
  // Host arrays (and the API) stay in face order; every device face array is
  // in device face order (face_upload), and the device cf_ent holds device
  // face ids in each cell's ascending face order (the reference's
  // accumulation order).
  std::vector<int> fd2o, fo2d;
  template <class T>
  void face_upload(DBuf<T> &d, const std::vector<T> &h, int width = 1) const {
    if (h.size() != (size_t)nf * width) fail(MS_EINVAL, "internal: face array size");
    std::vector<T> t(h.size());
    for (int f = 0; f < nf; f++)
      for (int w = 0; w < width; w++) t[(size_t)width * f + w] = h[(size_t)width * fd2o[f] + w];
    d.upload(t);
  }
  // natural-order geometry (face arrays: face order; cell arrays: natural order)
  std::vector<double> xf, sf, area, lf, vol_nat, xc_nat;
  // reference CSR (1-based, natural) and its slot map to the device matrix
  std::vector<int> ref_row_ptr, ref_col;
  std::vector<long long> ref2dev;
  SellPattern pat;
  // the matrix pattern in use: pat, or the base system's on a forked system
  // (fork_system), which shares its base's mesh, geometry, face constants,
  // zone properties and patterns and owns its fields, matrices and solvers
  const SellPattern *P = &pat;
  System *base = nullptr;               // set on a forked system
  std::atomic<int> nforks{0};           // live forks of this system
  std::unique_ptr<Krylov> kry;
  std::unique_ptr<SchurSystem> schur;   // gpu_ppe_solver = schur
  // per-cell properties (natural order, host copy)
  std::vector<double> porosity, perm_dir, inertial, k_s, rho_s, cp_s;
  std::vector<double> disp_l, disp_t, bj_alpha, h_sf, a_sf, q_src;
  std::vector<int> cztype;                 // natural order, CZ_* (empty: all fluid)
  bool props_dirty = true;
  // porous / solid flags derived from the properties (refresh_face_constants)
  bool has_face_flags = false, has_solid = false, has_disp = false, has_kink = false;
  bool has_bj = false;         // Beavers-Joseph faces
  bool solid_consts = false;   // keep solid-phase face constants (per-call assembly)
  bool bc_energy = false, has_qsrc = false;   // boundary / solid sources make T non-trivial
  bool T_nonzero = false;      // caller seeded a non-zero temperature field
  int ts_order = 0;
  double inlet_ramp = 1.0, pval_ramp = 1.0;
  double flux_ref = 1.0;
  double bnd_area = 0.0;   // boundary area (flux_ref)
  // device mesh data
  DBuf<int> d_c0, d_c1, d_fgrp, d_cf_ptr, d_cf_ent;
  DBuf<long long> d_cf_slot, d_diag_slot;
  DBuf<int> d_e_oth;
  DBuf<double> d_e_lf, d_e_sf, d_e_Dm, d_e_muf, d_e_tcor;
  // host face array (face order, width values per face) -> per cell-face entry
  template <class T>
  void entry_upload(DBuf<T> &d, const std::vector<T> &h, int width = 1) const {
    std::vector<T> t((size_t)width * cf_ent.size());
    for (size_t k = 0; k < cf_ent.size(); k++)
      for (int w = 0; w < width; w++) t[(size_t)width * k + w] = h[(size_t)width * (cf_ent[k] >> 1) + w];
    d.upload(t);
  }
  DBuf<double> d_xf, d_sf, d_area, d_lf, d_vol, d_xc;
  DBuf<double> d_dn, d_dvs, d_Dm, d_muf, d_tcor, d_dw, d_Dmb, d_Dt, d_kf, d_Dtb, d_tval, d_qval;
  DBuf<signed char> d_tmode, d_pidx, d_fflag, d_tmode_s, d_cztype;
  DBuf<double> d_Cbj, d_kch, d_Ds, d_ksf, d_Dsb;
  DBuf<DevBC> d_bc;
  // device fields (device numbering)
  DBuf<double> u, p, T, T_s, flux, apc, u_old, u_old_old, T_old, T_old_old, T_s_old, T_s_old_old;
  DBuf<double> gp, gu, gt, gts, psi;
  DBuf<double> porosity_d, perm_dir_d, inertial_d, k_s_d, rho_s_d, cp_s_d;
  DBuf<double> disp_l_d, disp_t_d, h_sf_d, a_sf_d, q_src_d, kdisp, gp_prev;
  // work
  DBuf<double> val, rhs, ap, pp, dpp, af, ucomp, part, diag8, tmpc;
  DBuf<double> mom_rhs3, mom_ap1;   // the three momentum components' interior-face sums (fused assembly)
  DBuf<unsigned int> ticket;
  std::vector<double> history;  // 8 per row
  // p' solve counters since creation: solves, iterations run in FP32 and in
  // FP64, solves that ran or finished in FP64 on a mixed-precision system
  long long st_ppe_solves = 0, st_it32 = 0, st_it64 = 0, st_fallback = 0;
  long long st_mom_solves = 0, st_mom_its = 0;   // momentum component solves and their iterations
  // transient drivers: steps taken since the last (re)started run, and the
  // ts_order that run found on entry (a continued run goes on from them)
  int steps_done = 0, ts_order_in = 0;
  int pin_cell = 0, has_pdir = 0, has_outflow = 0;
  // coupling interface: per-face data (face order, zero until set), the
  // switches of set_iface_p_dirichlet / _slip_mode / _zhang_vel /
  // _t_ltne_mode, and whether the current driver assembles the PPE the
  // serial way (interface faces pressure-Dirichlet under the pressure model)
  bool has_iface = false, iface_dirty = false, ppe_serial = true;
  int iface_p_dir = 0, iface_slip = 0, iface_zhang = 0, iface_T_ltne = 0;
  std::vector<double> iface_vel, iface_p, iface_T, iface_Ts, iface_q, iface_slip_C, iface_peer_vel;
  std::vector<int> iface_T_mode;
  DBuf<double> d_iface_vel, d_iface_p, d_iface_slip_C, d_iface_peer_vel, d_tval_s;
  // wall-zone post-processing: node z range (the span of the reference area)
  // and the wall faces in face order (built on first use)
  double zspan = 0.0;
  std::vector<int> wall_faces;
  DBuf<int> d_wall_faces;
  bool walls_ready = false;

  // the energy equations can have a non-zero solution (otherwise their
  // right-hand sides are zero and the solves return at once)
  bool energy_on() const { return T_nonzero || bc_energy || has_qsrc; }
  // min over cells of vol^(1/3) (host pow, as the reference's CFL estimate),
  // evaluated once: the volumes do not change
  double minv13_ = -1.0;
  double min_vol13() {
    if (minv13_ < 0.0) {
      double m = INFINITY;
      for (double v : vol_nat) m = std::min(m, std::pow(v, 1.0 / 3.0));
      minv13_ = m;
    }
    return minv13_;
  }
  bool solid_phase() const { return ctrl.thermal_model != 0 || has_solid; }
  DMesh dmesh() const;
  DFields dfields();
  void refresh_face_constants();
  size_t device_bytes() const;
  System() = default;
  System(const System &) = delete;
  System &operator=(const System &) = delete;
  ~System() {
    if (base) base->nforks--;
  }
};

System *create_system(int nnodes, const double *x, int nfaces, const long long *fptr, const int *fnodes,
                      const int *c0, const int *c1, const int *zone, int ncells, const ms_ctrl_t *ctrl,
                      const ms_bcgroup_t *groups);

// A case of a parameter sweep: a new system on the calling thread's stream
// that views base's mesh, geometry, face constants, zone properties and
// matrix patterns (read-only, no copy) and owns its fields (a copy of
// base's), matrices and solvers.  groups replaces base's boundary groups
// (same count and types; a sweep changes inlet speeds).  Base must outlive
// the fork and must not change its mesh-wide data while forks exist.  A fork
// supports the drivers and field reads; the host topology stays with base.
System *fork_system(System &base, const std::vector<ms_bcgroup_t> &groups);

// Host only: partner[f] = the face joined with boundary face f by the
// periodic pairs of ctrl (both directions), -1 elsewhere; same matching and
// errors as create_system.
void periodic_match(int nnodes, const double *x, int nfaces, const long long *fptr, const int *fnodes,
                    const int *c0, const int *c1, const int *zone, int ncells, const ms_ctrl_t *ctrl, int *partner);

// Wall-zone post-processing on the device fields (mod_uns_forces.f90):
// out14 = Fx Fy Fz CD CL u_inf A_ref p_inf, pressure parts, viscous parts;
// samples = (p, u_x, u_y, u_z) of the owner cell of every wall face, face
// order; wake_end = max(xcyl, max xc over cells with u_x < 0 behind xcyl).
void wall_forces(System &s, double *out14);
void wall_samples(System &s, std::vector<double> &pu);
double wake_end(System &s, double xcyl);
void free_stream(const System &s, double *u_inf, double *p_inf);

// VTU file of the cell fields (host only; vtk_write + vtu_write layouts)
void vtu_write(const char *path, bool binary, int nnodes, const double *x, int nfaces, const long long *fptr,
               const int *fnodes, const int *c0, const int *c1, int ncells, const double *xc, const double *p,
               const double *u, const double *T, const double *T_s, const double *q_src);

// field dump v1 (write_field_dump / read_field_dump); f = u p T u_old
// u_old_old T_old T_old_old; dump_read returns the reference's ier (0 ok, 1
// open, 2 header, 3 version, 4 ncells, 5 data)
void dump_write(const char *path, int ncells, int ts_order, const char *source_mesh, const char *generated,
                const double *const f[7]);
int dump_read(const char *path, int ncells, double *const f[7], int *ts_order, std::string &src);

// flang runtime NORM2 evaluation order (max-scaled accumulator), used for
// every geometric length so the device coefficients equal the reference's.
double norm2_ref(const double *v, int n);
std::vector<int> rcb_aggregates(int n, const std::vector<double> &xc, int k);
void build_schur(System &s, const std::vector<int> &agg_dev, int nagg);
// the reduced system's matrix-free S over the full pattern A (buffers and
// the Krylov hook)
void attach_matfree(SchurSystem &S, const SellPattern *A);
SolveResult schur_solve(System &s, const double *val, const double *b, double *x, double tol, int maxit, bool mixed,
                        double inner_tol, bool reuse);
void cell_ordering(int nc, const std::vector<int> &ptr, const std::vector<int> &adj, const std::vector<double> &xc,
                   bool rcm, bool colored, int nagg, std::vector<int> &n2o, std::vector<int> &color, int &ncolors,
                   std::vector<int> &agg);
std::vector<int> greedy_colors(int n, const std::vector<int> &ptr, const std::vector<int> &adj,
                               const std::vector<int> &visit, int &ncol);
void build_coarse(SellPattern &P, const std::vector<int> &agg_dev, int k);

// build a SellPattern from a 0-based CSR (columns sorted per row)
void csr_to_sell(SellPattern &pat, int n, const std::vector<int> &rp, const std::vector<int> &ci);

}  // namespace ms
