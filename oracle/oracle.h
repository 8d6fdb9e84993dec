/*
 * oracle.h -- CPU FP64 restatement of the reference's unstructured
 * SIMPLE/PISO pressure-velocity path (TEST INFRASTRUCTURE ONLY).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * this library, and only to check results and as the CPU baseline.  The product path
 * (libmixsolver_hip.so) never links or calls it.
 *
 * Every routine restates the arithmetic of the reference Fortran in the same
 * evaluation order (no FMA contraction, sequential sums in face order) so
 * that its results are bit-identical to the reference built with amdflang -O2.
 * Struct layouts are identical to ms_ctrl_t / ms_bcgroup_t in
 * include/mixsolver.h so the same ctypes objects feed both libraries.
 */
#ifndef MIXSOLVER_ORACLE_H
#define MIXSOLVER_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* boundary-condition type codes: mod_uns_control.f90:32-49 */
enum { OR_BC_NONE = 0, OR_BC_WALL = 1, OR_BC_SYMMETRY = 2, OR_BC_VINLET = 3,
       OR_BC_POUTLET = 4, OR_BC_FARFIELD = 5, OR_BC_SLIPWALL = 6,
       OR_BC_MASSINLET = 7, OR_BC_OUTFLOW = 8, OR_BC_INTERFACE = 9,
       OR_BC_VINLET_PARAB = 10 };

/* control (subset of ctrl_t, mod_uns_control.f90:148-280) */
typedef struct {
  double rho, mu, alpha_u, alpha_p;
  double outer_tol, lin_tol, conv_blend, nonorth_corr;
  double dt, cp, k_cond, tref, beta;
  double gravity[3], body_force[3];
  double init_T, t_pert;
  int outer_max, lin_max, ppe_precond, transient;  /* ppe_precond: 1 = ic0, else jacobi */
  int time_scheme, n_time_max, n_correct, n_out_every;
  int pimple, n_outer_iter, boussinesq, thermal_model;   /* thermal_model 0 lte, 1 ltne */
  int inlet_ramp, nbc;
  /* device-path extension keys (ignored by the oracle) */
  int gpu_ppe_precond, gpu_mom_precond, gpu_precision, gpu_ppe_solver;
  double gpu_inner_tol;
  int gpu_reorder, gpu_coarse;
  /* periodic pairs: the reference has none; or_create rejects nperiodic > 0 */
  int nperiodic;
  int periodic_zone[4][2];
  int periodic_auto[4];
  double periodic_shift[4][3];
  int porous_unified;   /* porous_model: 0 partitioned, 1 unified */
} or_ctrl_t;

/* one boundary group (bcgroup_t, mod_uns_bc.f90:80-110) */
typedef struct {
  int zone, btype;
  double uvel[3], pval, mdot, uspeed;
  double prfl_mean, prfl_span, prfl_origin;
  int prfl_axis, has_lid;
  int lid_dir, ttype;
  double lid_coord, lid_tol, lid_vel[3];
  double tval, qval;
  int ntbc_plane, tbc_pdir[4], tbc_pttype[4];
  double tbc_pcoord[4], tbc_tol, tbc_ptval[4], tbc_pqval[4];
} or_bcgroup_t;

typedef struct or_case or_case;

/* mesh: x[3*nn] node coordinates (node-major), face node CSR fptr[nf+1]
 * (0-based offsets) / fnodes (1-based), c0/c1 1-based (c1 = 0 boundary),
 * zone[nf]; groups[nbc] in control order.  Returns NULL on error and writes
 * an error code into *ier (10 unknown zone, 11 empty zone, 12 orphan face,
 * 30 non-positive volume, 50 unsupported feature). */
or_case *or_create(int nn, const double *x, int nf, const long long *fptr,
                   const int *fnodes, const int *c0, const int *c1,
                   const int *zone, int nc, const or_ctrl_t *ctrl,
                   const or_bcgroup_t *groups, int *ier);
void or_destroy(or_case *cs);

/* geometry / topology views (sizes in comments) */
const double *or_xf(or_case *);     /* 3*nf */
const double *or_sf(or_case *);     /* 3*nf */
const double *or_area(or_case *);   /* nf   */
const double *or_vol(or_case *);    /* nc   */
const double *or_xc(or_case *);     /* 3*nc */
const double *or_lf(or_case *);     /* nf   */
const int *or_cf_ptr(or_case *);    /* nc+1, 0-based */
const int *or_cf(or_case *);        /* 1-based face ids */
const int *or_row_ptr(or_case *);   /* nc+1, 1-based (Fortran) */
const int *or_col_idx(or_case *);   /* nnz, 1-based */
int or_nnz(or_case *);
const int *or_fgrp(or_case *);      /* nf, group id 1..nbc or 0 */

/* mutable fields (same layouts as fields_t: u(3,nc) etc.) */
double *or_field(or_case *, const char *name);
void or_set_ts_order(or_case *, int);
int or_get_ts_order(or_case *);

/* per-call routines */
void or_compute_gradients(or_case *);
void or_momentum_assembly(or_case *, int comp, double *val, double *rhs, double *ap);
void or_temperature_assembly(or_case *, double *val, double *rhs, double *ap);
void or_solid_temperature_assembly(or_case *, double *val, double *rhs, double *ap);
/* block type per cell (1 fluid, 2 porous, 3 solid; NULL = all fluid) */
void or_set_cztype(or_case *, const int *cztype);
void or_flux_rhiechow(or_case *);
void or_outflow_mass_rescale(or_case *);
void or_ppe_assembly(or_case *, double *val, double *rhs);
/* mpi = 1: ppe_assembly_mpi at one partition (mod_uns_simple_mpi.f90:882-968) */
void or_ppe_assembly_variant(or_case *, double *val, double *rhs, int mpi);
void or_correct_fields(or_case *, const double *pp, double *dpp);

/* drivers.  mpi = 1 selects the partitioned-path norms sqrt(gdot) and
 * ordering of mod_uns_simple_mpi.f90, mpi = 0 the serial norm2 variants. */
int or_simple_run(or_case *, int nsteps, int mpi, int *it_done);
int or_piso_run(or_case *, int nsteps, int mpi);
int or_pimple_run(or_case *, int nsteps, int mpi);
/* per-step diagnostics of the last driver call: n_steps x 6
 * (imbal, du_max, umax, cfl_max, lin_it_max, lin_res_mean) */
const double *or_history(or_case *, int *nrows);
/* wall time of the last PISO call, time inside its p' solves, and the
 * number of p' PCG iterations it performed (CPU-baseline extrapolation) */
void or_timing(or_case *, double *out3);

/* wall-zone post-processing (mod_uns_forces.f90) on the current fields:
 * compute_forces -> out14 = Fx Fy Fz CD CL u_inf A_ref p_inf, pressure parts
 * Fx_p Fy_p Fz_p, viscous parts Fx_v Fy_v Fz_v; write_cp_distribution ->
 * rows6 (theta deg, Cp, Cp_exact, x, y, z) per wall face, returns the row
 * count; compute_separation -> out5 = theta_top, found_top, theta_bot,
 * found_bot, wake length */
void or_compute_forces(or_case *, double *out14);
/* coupling interface data and switches (mod_uns_bc.f90:730-869) */
int or_set_interface(or_case *, const char *what, const double *values);
/* nan_scan: first non-finite value (field 0..5 = u_x u_y u_z p T T_s, cell
 * 1-based, kind 1 NaN / 2 Inf); returns ifin */
int or_nan_scan(or_case *, int *field, int *cell, double *val, int *kind);
void or_set_interface_modes(or_case *, int p_dir, int slip, int zhang, int T_ltne);
int or_cp_distribution(or_case *, double *rows6);
void or_separation(or_case *, double *out5);

/* linear solvers on a 1-based CSR (csr_t), mod_uns_linsolver(_mpi).f90 */
void or_cg_jacobi(int n, const int *rp, const int *ci, const double *v,
                  const double *b, double *x, double tol, int maxit,
                  int *it, double *res, int *ierr, int mpi);
void or_cg_ic0(int n, const int *rp, const int *ci, const double *v,
               const double *b, double *x, double tol, int maxit,
               int *it, double *res, int *ierr, int mpi);
void or_bicgstab_ilu0(int n, const int *rp, const int *ci, const double *v,
                      const double *b, double *x, double tol, int maxit,
                      int *it, double *res, int *ierr, int mpi);
int or_ic0_factor(int n, const int *rp, const int *ci, double *v);
int or_ilu0_factor(int n, const int *rp, const int *ci, double *v);
void or_csr_matvec(int n, const int *rp, const int *ci, const double *v,
                   const double *x, double *y);
double or_norm2(const double *v, int n);

#ifdef __cplusplus
}
#endif
#endif
