/*
 * mixsolver.h -- C ABI of libmixsolver_hip.so, the MI355X (gfx950) device
 * implementation of the unstructured SIMPLE/PISO pressure-velocity path.
 *
 * Plain pointers and sizes only.  Arrays use the reference's own layouts:
 * Fortran column-major fields (u(3,ncells) = xyz interleaved per cell),
 * 1-based owner/neighbour ids (c1 = 0 on boundary faces), 1-based CSR
 * (csr_t: row_ptr(n+1), col_idx(nnz) sorted per row, val(nnz)).
 * All host arrays are owned by the caller; device memory is owned by the
 * library and keyed by the opaque ms_system handle.
 *
 * Return value of every entry point: 0 on success, a negative MS_E* code on
 * an API/device error (message via ms_last_error()).  Solver/driver status
 * follows the reference's ierr/ier conventions in out-parameters.
 * Calls on one handle must not overlap; different handles are independent.
 */
#ifndef MIXSOLVER_H
#define MIXSOLVER_H
#ifdef __cplusplus
extern "C" {
#endif

#define MS_OK 0
#define MS_EINVAL (-1)     /* bad argument / size / unsupported feature */
#define MS_EDEVICE (-2)    /* HIP runtime error */
#define MS_ENOMEM (-3)     /* device allocation failed / over HBM budget */

/* BC type codes: mod_uns_control.f90:32-49 */
enum { MS_BC_NONE = 0, MS_BC_WALL = 1, MS_BC_SYMMETRY = 2, MS_BC_VINLET = 3,
       MS_BC_POUTLET = 4, MS_BC_FARFIELD = 5, MS_BC_SLIPWALL = 6,
       MS_BC_MASSINLET = 7, MS_BC_OUTFLOW = 8, MS_BC_INTERFACE = 9,
       MS_BC_VINLET_PARAB = 10 };

/* preconditioner / precision selectors for the device-path keys */
/* MS_PC_NATURAL: ILU(0) / IC(0) in the caller's cell order (the reference's
 * factors); MS_PC_MC: the same factorizations in the colour-blocked cell
 * numbering, so the rows of one colour are factored and swept together
 * (multicolour ILU(0) / IC(0)). */
enum { MS_PC_JACOBI = 0, MS_PC_NATURAL = 1, MS_PC_DILU = 2, MS_PC_MC = 3 };
enum { MS_FP64 = 0, MS_MIXED = 1 };

/* Control values: the fields of ctrl_t used on the path
 * (mod_uns_control.f90:148-280) plus device-path keys. */
typedef struct {
  double rho, mu, alpha_u, alpha_p;
  double outer_tol, lin_tol, conv_blend, nonorth_corr;
  double dt, cp, k_cond, tref, beta;
  double gravity[3], body_force[3];
  double init_T, t_pert;
  int outer_max, lin_max, ppe_precond, transient;   /* ppe_precond: 1 = 'ic0', else jacobi */
  int time_scheme, n_time_max, n_correct, n_out_every;
  int pimple, n_outer_iter, boussinesq, thermal_model;
  int inlet_ramp, nbc;
  int gpu_ppe_precond;   /* MS_PC_*: p' preconditioner on the device        */
  int gpu_mom_precond;   /* MS_PC_*: momentum / energy preconditioner        */
  int gpu_precision;     /* MS_FP64 or MS_MIXED (FP32 inner, FP64 refinement) */
  int gpu_ppe_solver;    /* 0 = PCG (reference), 1 = BiCGSTAB (IC(0)-type
                            preconditioners become ILU(0)), 2 = PCG on the
                            red-black Schur complement of a two-colour cell
                            graph (colour 0 eliminated; Jacobi on S, the
                            coarse space of gpu_coarse on colour 1)        */
  double gpu_inner_tol;  /* relative tolerance of each FP32 inner solve      */
  int gpu_reorder;       /* 0 = natural order, 1 = RCM                       */
  int gpu_coarse;        /* p' PCG coarse space: number of aggregates, 0 = off (max 1024) */
  /* Periodic face pairs (no reference counterpart; the reference's BC list
   * has no periodic type).  Pair i joins every boundary face of zone
   * periodic_zone[i][0] to the face of zone periodic_zone[i][1] whose
   * centroid translated by periodic_shift[i] coincides with it; the two
   * become one interior face.  periodic_auto[i] != 0 takes the shift from the
   * difference of the two zones' mean face centroids.  Periodic zones have
   * no bc group. */
  int nperiodic;                    /* 0..4 */
  int periodic_zone[4][2];
  int periodic_auto[4];
  double periodic_shift[4][3];
  /* porous_model (mod_uns_control.f90:488-499): 0 = partitioned (kink face
   * pressure and Beavers-Joseph slip at fluid/porous faces), 1 = unified
   * (harmonic face viscosity, no interface treatment) */
  int porous_unified;
} ms_ctrl_t;

/* One boundary group (bcgroup_t, mod_uns_bc.f90:80-110), control order. */
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
} ms_bcgroup_t;

typedef struct ms_system ms_system;

/* ---- library ---------------------------------------------------------- */
int ms_init(int device);                 /* select device, create streams   */
const char *ms_last_error(void);
int ms_version(void);
/* Replay the mixed-precision PCG's 16-iteration chunks, and the request
 * server's rounds while its iterating set is unchanged, from captured HIP
 * graphs (1, default; the environment variable MIXSOLVER_GRAPHS=0 sets 0 at
 * start-up) or launch their kernels one by one (0).  Results are identical
 * either way.  No reference counterpart (library option). */
int ms_set_graphs(int on);
/* The drivers assemble the three momentum components' interior faces in one
 * pass (1, default; MIXSOLVER_MOM_FUSED=0 sets 0 at start-up) where that is
 * exact: no Beavers-Joseph faces and, with the convection limiter on, no
 * symmetry or slip-wall group; else (0) one pass per component as
 * momentum_assembly (mod_uns_simple.f90:928-1338).  Results are identical
 * either way.  No reference counterpart (library option). */
int ms_set_fused_momentum(int on);
/* The FP32 iterations of the red-black reduced p' solve (gpu_ppe_solver =
 * schur) apply S = D1 - B^T D0^-1 B matrix-free, in two launches over the
 * full matrix's rows (1, default; MIXSOLVER_SCHUR_MF=0 sets 0 at start-up),
 * or the formed S (0).  Both reach the same FP64 residual; the FP32
 * rounding, and so the iteration count, differ.  No reference counterpart
 * (library option). */
int ms_set_schur_matfree(int on);
/* With the matrix-free S, the reduced PCG's p update (p = z + nu + beta p)
 * is applied inside the next product's two launches (1; MIXSOLVER_P_FUSED=1
 * at start-up) or by a launch of its own (0, default: measured faster).
 * Results are identical either way.  No reference counterpart (library
 * option). */
int ms_set_fused_p_update(int on);

/* vtk_write (mod_uns_output.f90:81-240, tetrahedral fan) through vtu_write
 * (mod_vtu_writer.f90): binary = 0 inline ASCII (Float64, ES24.16E3),
 * 1 appended raw binary (Float32).  Mesh as for ms_system_create; xc (3,
 * ncells) cell centroids; p, u (3,ncells), T cell fields; T_s written as
 * temperature_solid when not NULL, q_src as heat_source_solid when not NULL
 * (the reference writes them when LTNE exchange is active or a solid zone
 * exists, and when a solid zone exists).  Host only; byte-identical to the
 * reference's file for the same inputs. */
/* write_field_dump (mod_uns_restart.f90:77-152): field dump v1, ASCII,
 * ES24.16, the header the reference writes ('# generated:' from `generated`
 * when not NULL/empty, else the local time).  u, u_old, u_old_old (3,ncells);
 * p, T, T_old, T_old_old (ncells).  Host only. */
int ms_dump_write(const char *path, int ncells, int ts_order, const char *source_mesh, const char *generated,
                  const double *u, const double *p, const double *T, const double *u_old, const double *u_old_old,
                  const double *T_old, const double *T_old_old);
/* read_field_dump (mod_uns_restart.f90:166-334, serial): the arrays as for
 * ms_dump_write, *ts_order, the source_mesh header (truncated to
 * source_len - 1 characters) and *ier with the reference's codes: 0 ok,
 * 1 cannot open, 2 header ended early or malformed, 3 version mismatch,
 * 4 ncells mismatch, 5 data.  Host only. */
int ms_dump_read(const char *path, int ncells, double *u, double *p, double *T, double *u_old, double *u_old_old,
                 double *T_old, double *T_old_old, int *ts_order, char *source_mesh, int source_len, int *ier);

int ms_vtu_write(const char *path, int binary, int nnodes, const double *x, int nfaces, const long long *fptr,
                 const int *fnodes, const int *c0, const int *c1, int ncells, const double *xc, const double *p,
                 const double *u, const double *T, const double *T_s, const double *q_src);

/* ---- system = one mesh + fields + matrices resident on the device ------
 * Replaces the setup half of uns_run_mpi / uns_solver_init
 * (mod_uns_driver_mpi.f90:114-228, mod_uns_driver.f90:60-239):
 * build_connectivity, compute_geometry, build_bc, init_fields,
 * build_cell_csr -- from the arrays read_cas fills in mesh_t
 * (mod_uns_mesh.f90:17-54).  fptr: face-node offsets (0-based, nfaces+1),
 * fnodes: 1-based node ids, x: (3,nnodes). */
int ms_system_create(ms_system **out, int nnodes, const double *x, int nfaces,
                     const long long *fptr, const int *fnodes, const int *c0,
                     const int *c1, const int *zone, int ncells,
                     const ms_ctrl_t *ctrl, const ms_bcgroup_t *groups);
int ms_system_destroy(ms_system *s);
int ms_system_info(ms_system *s, long long *out8);  /* ncells nfaces nnz ncolors nlevels bytes_device
                                                       nstore renumbered */
/* p' coarse space of a system: k aggregates, nchunks row chunks; agg (may be
 * NULL) receives the aggregate of every cell in the caller's numbering. */
int ms_coarse_info(ms_system *s, int *k, int *nchunks, int *agg);

/* Aggregates of the p' coarse space (gpu_coarse > 0): recursive coordinate
 * bisection of the cell centres xc (3,n) into k groups of equal size (+-1),
 * splitting each group along the longest side of its bounding box.  agg[i]
 * in 0..k-1.  Host only (no device needed); no reference counterpart. */
int ms_coarse_aggregates(int n, const double *xc, int k, int *agg);

/* The device numbering ms_system_create uses (no reference counterpart;
 * the reference keeps the mesh's order): n2o[new] = old cell (0-based) of a
 * cell graph given as a 1-based CSR pattern (row_ptr, col_idx; the diagonal
 * may be present).  rcm != 0: reverse Cuthill-McKee visiting order (from a
 * minimum-degree cell of each component, ties by index), else natural;
 * colored != 0: greedy colouring in the visiting order, colour classes
 * ascending (color[old] = class, *ncolors classes); coarse = k > 0: k
 * recursive-bisection aggregates of xc (3,n), each contiguous inside each
 * class.  Host only (no device needed), deterministic. */
int ms_cell_ordering(int n, const int *row_ptr, const int *col_idx, const double *xc, int rcm, int colored,
                     int coarse, int *n2o, int *color, int *ncolors);

/* Periodic face matching of ctrl's periodic pairs on a mesh given as for
 * ms_system_create: partner[f] (nfaces) = the face joined with boundary face
 * f (0-based, both directions), -1 elsewhere.  Host only (no device
 * needed); the same matching and errors as ms_system_create.  No reference
 * counterpart (the reference has no periodic boundary). */
int ms_periodic_match(int nnodes, const double *x, int nfaces, const long long *fptr, const int *fnodes,
                      const int *c0, const int *c1, const int *zone, int ncells, const ms_ctrl_t *ctrl,
                      int *partner);

/* fields_t components by name: "u"(3,nc) "p" "T" "T_s" "flux"(nf) "apc"
 * "u_old" "u_old_old" "T_old" "T_old_old" "T_s_old" "T_s_old_old" "gp"(3,nc)
 * "gu"(3,3,nc) "gt" "gts"; per-cell porous-medium coefficients
 * (mod_uns_fields.f90:50-71, filled by setup_porous_fields :418-504)
 * "porosity" "perm_dir"(3,nc) "inertial" "disp_l" "disp_t" "bj_alpha" "k_s"
 * "rho_s" "cp_s" "h_sf" "a_sf" "q_src";
 * geometry (read-only): "xf" "sf" "area" "vol" "xc" "lf". */
int ms_set_field(ms_system *s, const char *name, const double *host, long long n);
int ms_get_field(ms_system *s, const char *name, double *host, long long n);
int ms_set_ts_order(ms_system *s, int ts_order);
int ms_get_ts_order(ms_system *s, int *ts_order);   /* fld%ts_order of the system */
/* nan_scan (mod_uns_simple.f90:2375-2426) on the current fields: the first
 * non-finite value in cell order (caller's numbering), u_x u_y u_z p T T_s
 * within a cell: *field 0..5, *cell 1-based, *val, *kind 1 NaN / 2 Inf,
 * *ifin 1 when found.  The drivers return 60 (NAN_IER) when a step leaves a
 * non-finite value; this names it for the report. */
int ms_nan_scan(ms_system *s, int *field, int *cell, double *val, int *kind, int *ifin);

/* ---- coupling interface (mode = coupled callers) ------------------------
 * Boundary faces of face zones whose condition or user name contains
 * "interface" and that have no bc group form a BC_INTERFACE group (build_bc,
 * mod_uns_bc.f90:282-323): pass it as a group with btype MS_BC_INTERFACE and
 * the zone id.  Per-face data, natural face order, all faces (values of
 * other faces are ignored), replacing set_interface_vel / _p / _T_ltne /
 * _q / _T_mode / _slip (mod_uns_bc.f90:730-821): what = "vel" (3,nfaces),
 * "p", "T", "Ts", "q", "slip_C", "peer_vel" (3,nfaces), "T_mode" (0
 * Dirichlet iface T, 1 Neumann with q; set_interface_q sets 1 where it
 * applies).  n = number of values. */
int ms_set_interface(ms_system *s, const char *what, const double *values, long long n);
/* set_iface_p_dirichlet / set_iface_slip_mode / set_iface_zhang_vel /
 * set_iface_t_ltne_mode (mod_uns_bc.f90:823-857); per system here (module
 * state in the reference). */
int ms_set_interface_modes(ms_system *s, int p_dir, int slip, int zhang, int T_ltne);
/* Block type of every cell (m%cztype after resolve_cell_zones,
 * mod_uns_control.f90:575-733): 1 fluid, 2 porous, 3 solid; NULL = all
 * fluid.  Porous coefficients come through ms_set_field. */
int ms_set_cell_types(ms_system *s, const int *cztype);
int ms_get_topology(ms_system *s, int *row_ptr, int *col_idx, int *fgrp); /* 1-based CSR */

/* ---- loop level: keep everything resident across iterations ----------- */
/* simple_run (mod_uns_simple.f90:88) / simple_run_mpi
 * (mod_uns_simple_mpi.f90:57): nsteps caps outer_max (0 = outer_max).
 * ier: 0 ok, 20-23 breakdowns, 60 non-finite (reference codes). */
int ms_simple_run(ms_system *s, int nsteps, int *it_done, int *ier);
/* piso_run_mpi at one partition (mod_uns_simple_mpi.f90:298); nsteps
 * overrides n_time_max when > 0.  ier: 30-36 / 60 as the reference. */
int ms_piso_run(ms_system *s, int nsteps, int *ier);
/* pimple_run_mpi at one partition (mod_uns_simple_mpi.f90:580): n_outer_iter
 * outer iterations per step; ier: 50-56 / 60 as the reference. */
int ms_pimple_run(ms_system *s, int nsteps, int *ier);
/* The next nsteps steps of the last ms_piso_run / ms_pimple_run on this
 * system, as if that call had been given more steps: no re-initialisation
 * (apc, gradients and fluxes stay as the last step left them) and the step
 * count goes on (the first-order start applies to the run's first step
 * only).  The same loop as piso_run_mpi / pimple_run_mpi, cut where a
 * caller syncs fields (snapshots every n_out_every steps). */
int ms_piso_continue(ms_system *s, int nsteps, int *ier);
int ms_pimple_continue(ms_system *s, int nsteps, int *ier);
/* per-iteration/step diagnostics of the last driver call, 8 doubles per row:
 * imbal du_max dT_max|umax cfl lin_it_max lin_res_mean ppe_it_total
 * ppe_res_max (the largest final p' residual over the step's correctors) */
int ms_get_history(ms_system *s, double *out, int maxrows, int *nrows);
/* Solve counters of the system since creation or the last reset (no
 * reference counterpart): out8 = p' solves, p' iterations run in FP32, p'
 * iterations run in FP64, p' solves that ran or finished in FP64 on a
 * mixed-precision system, 1 when the system's FP32 p' iteration has stalled
 * (every later p' solve is FP64), the stall window in force, momentum
 * component solves, their iterations (FP32 inner iterations on a
 * mixed-precision system).  reset != 0 zeroes the counters after reading
 * them (the stall switch stays). */
int ms_solver_stats(ms_system *s, long long *out8, int reset);
/* Stall guard of the mixed-precision PCG (no reference counterpart): when the
 * FP32 residual has not halved over `iters` iterations (checked every 16
 * iterations; 0 = the default 1024) the FP32 iteration stops, the FP64 PCG
 * with the same preconditioner finishes from the current x with the
 * remaining budget, and the system solves p' in FP64 from then on.  An
 * exhausted mixed PCG reports it = maxit + 1 and ierr = 1, as the FP64 CG.
 * Library-wide setting. */
int ms_set_stall_window(int iters);
/* In-loop timing of the mixed-precision PCG's FP32 SpMV (measurement, no
 * reference counterpart).  The counters (launches timed, their summed
 * duration in ms, rows of the last matrix timed) are read into the three
 * pointers when they are all non-null, then on = 1 turns the probe on and
 * zeroes them, on = 0 turns it off, on = -1 leaves it as it is.  With the
 * probe on, the last SpMV of every 16-iteration chunk is launched with its
 * own start / stop events (the kernel's dispatch timestamps) behind the
 * chunk's captured graph, which then holds the other 15 iterations.
 * Library-wide setting. */
int ms_spmv_probe(int on, long long *launches, double *total_ms, long long *rows);

/* ---- wall-zone post-processing on the resident fields ------------------
 * compute_forces (mod_uns_forces.f90:43-185): out14 = Fx Fy Fz CD CL u_inf
 * A_ref p_inf, then the pressure parts Fx_p Fy_p Fz_p and the viscous parts
 * Fx_v Fy_v Fz_v (the reference prints these).  Pressure force sum
 * (p_c0 - p_inf) S_f and wall shear mu u_t / dn along n x z over the wall
 * faces, reduced on the device; u_inf / p_inf of the first far-field group;
 * A_ref = (x extent of the wall faces) x (z extent of the nodes). */
int ms_compute_forces(ms_system *s, double *out14);
/* write_cp_distribution (mod_uns_forces.f90:195-254) as rows of 6 doubles,
 * one per wall face in face order: theta (deg, about the mean of the wall
 * face centroids), Cp, 1 - 4 sin^2(theta), x, y, z.  *nrows = number of wall
 * faces (0 without a far-field group); rows6 may be null to query it. */
int ms_cp_distribution(ms_system *s, double *rows6, int maxrows, int *nrows);
/* compute_separation (mod_uns_forces.f90:275-407): out5 = top separation
 * angle (rad, from +x), found (1/0), bottom separation angle, found, wake
 * length (farthest reverse-flow cell centroid behind the wall centre). */
int ms_separation(ms_system *s, double *out5);

/* ---- per-call level (copy in/out through ms_set/get_field) -------------
 * val/rhs/ap are host arrays in the reference CSR order (nnz / ncells). */
int ms_compute_gradients(ms_system *s);                       /* mod_uns_fields.f90:164 */
int ms_momentum_assembly(ms_system *s, int comp, double *val, double *rhs, double *ap); /* mod_uns_simple.f90:928 */
int ms_temperature_assembly(ms_system *s, double *val, double *rhs, double *ap); /* :1748 */
int ms_solid_temperature_assembly(ms_system *s, double *val, double *rhs, double *ap); /* :2109 */
int ms_flux_rhiechow(ms_system *s);                           /* :1413 */
int ms_outflow_mass_rescale(ms_system *s);                    /* :1536 */
int ms_ppe_assembly(ms_system *s, double *val, double *rhs);  /* :1552 / mod_uns_simple_mpi.f90:882 */
int ms_correct_fields(ms_system *s, const double *pp, double *dpp); /* :1649 */

/* ---- linear solvers on a caller CSR (csr_t) ----------------------------
 * Same argument meaning as cg_jacobi / cg_ic0 / bicgstab_ilu0
 * (mod_uns_linsolver.f90:230, :405, :148) and their _mpi twins at one
 * partition (mod_uns_linsolver_mpi.f90:235, :306, :155): x is the initial
 * guess on entry, it/res/ierr as the reference (ierr 0 converged, 1 maxit,
 * 2 breakdown, 11/12 factorization failure). */
int ms_cg_jacobi(int n, const int *row_ptr, const int *col_idx, const double *val,
                 const double *b, double *x, double tol, int maxit,
                 int *it, double *res, int *ierr);
int ms_cg_ic0(int n, const int *row_ptr, const int *col_idx, const double *val,
              const double *b, double *x, double tol, int maxit,
              int *it, double *res, int *ierr);
int ms_bicgstab_ilu0(int n, const int *row_ptr, const int *col_idx, const double *val,
                     const double *b, double *x, double tol, int maxit,
                     int *it, double *res, int *ierr);
/* cg_jacobi that allocates nothing.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as ms_cg_jacobi (cg_jacobi_mpi recurrences,
 * it = maxit+1 / ierr = 1 on exhaustion).  The matrix, vectors, reduction
 * partials and solver scalars are staged in one fixed device array per GPU
 * (about 20 MB, held until the process ends): n <= 131072 and
 * nnz <= 1048576, larger systems return MS_EINVAL.  ms_init must have run on
 * the current device.  Calls share that GPU's array: calls that overlap on
 * one GPU race. */
int ms_cg_jacobi_noalloc(int n, const int *row_ptr, const int *col_idx, const double *val,
                         const double *b, double *x, double tol, int maxit,
                         int *it, double *res, int *ierr);
/* Extended solver entry (no reference counterpart: the device path's
 * solver options).  solver 0 PCG / 1 BiCGSTAB; precond MS_PC_*; precision
 * MS_FP64 / MS_MIXED; inner_tol for MS_MIXED.  res is the FP64 true
 * relative residual ||b-Ax||/||b|| for MS_MIXED. */
int ms_solve(int n, const int *row_ptr, const int *col_idx, const double *val,
             const double *b, double *x, double tol, int maxit, int solver,
             int precond, int precision, double inner_tol,
             int *it, double *res, int *ierr);

/* ---- serving (no reference counterpart) ---------------------------------
 * Continuous-batching server for SPD pressure-correction solves A x = b,
 * x0 = 0, FP32 inner iterations with FP64 reliable updates to tol (the FP64
 * true relative residual).  Patterns (one per mesh) are registered once and
 * stay resident; a request brings values in the pattern's CSR order and a
 * right-hand side, copied into pinned host memory before ms_server_submit
 * returns.  All admitted requests are packed into every launch; a result
 * does not depend on which other requests share the device (bit for bit).
 * The server's worker thread runs the iterations on the library stream and
 * the set-up of admitted requests (uploads, FP32 copy, preconditioner, coarse
 * inverse) on a second stream, so set-ups overlap the rounds.  Device and
 * pinned buffers of finished requests are kept per pattern and reused; the
 * pool is released when the budget needs room for another pattern. */
typedef struct ms_server ms_server;
typedef struct {
  long long hbm_budget;  /* bytes admitted requests may hold; 0 = half of free memory */
  int max_active;        /* requests in flight (default 64 when 0) */
  int precond;           /* MS_PC_JACOBI or MS_PC_DILU */
  int coarse;            /* max aggregates per pattern (0..256; needs xc): min(coarse, n/4096) are used */
  int round_iters;       /* iterations between scheduling decisions (default 16 when 0) */
  double delta;          /* reliable-update trigger (default 0.1 when 0) */
} ms_server_opts;
int ms_server_create(ms_server **out, const ms_server_opts *opts);
int ms_server_destroy(ms_server *s);
/* 1-based CSR pattern (csr_t layout, sorted columns, diagonal present);
 * xc (3,n) cell centres or NULL.  Returns the pattern id. */
int ms_server_add_pattern(ms_server *s, int n, const int *row_ptr, const int *col_idx,
                          const double *xc, int *pattern);
int ms_server_submit(ms_server *s, int pattern, const double *val, const double *b,
                     double tol, int maxit, long long *ticket);
/* timeout_ms < 0 waits, 0 polls.  *ready = 1 when x, it, res (FP64 true
 * relative residual), ierr (0 converged, 1 maxit, 2 breakdown, 3 FP32 stall:
 * the residual did not halve over the stall window of ms_set_stall_window,
 * and x is the iterate at that point, for the caller to finish in FP64,
 * e.g. with ms_solve from x as the initial guess) and the times
 * from submission to admission and to completion (seconds) were written; the
 * ticket is then released. */
int ms_server_result(ms_server *s, long long ticket, double timeout_ms, double *x, int *it,
                     double *res, int *ierr, double *t_admit, double *t_done, int *ready);
/* submitted completed active pending rounds launches peak_budget_bytes patterns */
int ms_server_stats(ms_server *s, long long *out8);

/* ---- parameter-sweep server (no reference counterpart) --------------------
 * Serves whole transient cases of one mesh that differ in an inlet speed
 * (a parameter sweep).  Every case is a fork of `base`: it reads the base's
 * mesh, geometry, face constants, zone properties and matrix patterns in
 * place and owns its fields (copied from the base's when it starts),
 * matrices and solver state.  A case runs `nsteps` steps of the partitioned
 * transient PISO driver at one partition (piso_run_mpi,
 * mod_uns_simple_mpi.f90:298-578; the first step as ms_piso_run, the others
 * as ms_piso_continue) with the base's control.  Cases are admitted by
 * device-memory budget (budget_bytes; 0 = 90% of the free device memory at
 * creation): the first case runs alone until its first step has allocated
 * what a case allocates; then cases start while the footprints of the cases
 * in flight plus one more fit the budget, at most max_active (1..64) at
 * once, each on a host thread and HIP stream of its own.  A case's results
 * do not depend on the mix (bit for bit).  While a sweep exists the base's
 * zone properties and cell types cannot change and the base cannot be
 * destroyed; do not run the base itself while cases are being admitted. */
typedef struct ms_sweep ms_sweep;
int ms_sweep_create(ms_sweep **sw, ms_system *base, long long budget_bytes, int max_active);
/* waits for the cases in flight; queued cases are dropped */
int ms_sweep_destroy(ms_sweep *sw);
/* group: 1-based boundary group whose speed the case sets (0: none):
 * velocity inlet / far field: uvel (3); parabolic inlet: mean speed `speed`;
 * mass inlet: speed.  keep_fields: keep the final u (3,ncells), p, T, T_s
 * for ms_sweep_result.  *id receives the case id. */
int ms_sweep_submit(ms_sweep *sw, int group, const double *uvel, double speed, int nsteps, int keep_fields,
                    long long *id);
/* Waits up to wait_s seconds for case id to end.  *state: 0 queued,
 * 1 running, 2 done, 3 failed (the call then returns the case's error code
 * and ms_last_error its message).  When done: *steps the steps run, *ier
 * the driver's ier, step_ms (nsteps) the wall time of each step from its
 * launch to its completion, history (8 per step) the rows of
 * ms_get_history, times3 = seconds since the sweep started at submission,
 * admission and completion, stats8 as ms_solver_stats; u, p, T, T_s
 * (natural order) when the case kept its fields.  Any output may be NULL. */
int ms_sweep_result(ms_sweep *sw, long long id, double wait_s, int *state, int *steps, int *ier, double *step_ms,
                    double *history, double *times3, long long *stats8, double *u, double *p, double *T,
                    double *T_s);
/* submitted done failed queued active peak_active footprint_bytes_per_case
 * budget_bytes base_device_bytes steps_run */
int ms_sweep_stats(ms_sweep *sw, long long *out10);

/* ---- measurement (no reference counterpart) -----------------------------
 * Average duration of the dominant solver kernel (PCG SpMV with its fused
 * p.q partials) on the system's current matrix, timed with HIP events on the
 * library stream over reps launches.  which: 0 = FP32 difference-form kernel
 * of the mixed path, 1 = FP64 kernel, back to back (the matrix stays in the
 * last-level cache when it fits); 2 / 3 = the same kernels, each launch timed
 * alone after a 512 MiB read that evicts the last-level cache (cold).  alg_bytes: nnz*(s+i) + 4(n+1) + 2ns with
 * s the value size and i the column-index size of the format in use (2 for the
 * FP32 kernel's 16-bit slice offsets, else 4). */
/* which 4..7: the same four timings on the red-black reduced p' matrix S of a
 * system with gpu_ppe_solver = 2, after its first p' solve.  With the
 * matrix-free S on (ms_set_schur_matfree), which 4 and 6 time its two
 * launches, and alg_bytes = nnz_A (4 + i) + 4 (n + 1) + 8 n1 + 16 n0 + 4 n1
 * over the full matrix A (n0 / n1 rows of colour 0 / 1). */
int ms_time_kernel(ms_system *s, int which, int reps, double *avg_ms, double *alg_bytes);
/* Average duration of a plain device-to-device copy kernel (k_copy4) moving
 * `bytes` (read once, written once): the calibration of traffic counters and
 * the achievable-bandwidth reference beside the published peak. */
int ms_time_copy(long long bytes, int reps, double *avg_ms);
/* Average duration of a streaming kernel over `bytes`: kind 0 = the copy
 * kernel above, 1 = the same copy with non-temporal stores, 2 = a read-only
 * pass.  The achievable-HBM references beside the published peak (buffers
 * larger than the 256 MB last-level cache).  No reference counterpart. */
int ms_time_stream(long long bytes, int kind, int reps, double *avg_ms);

#ifdef __cplusplus
}
#endif
#endif
