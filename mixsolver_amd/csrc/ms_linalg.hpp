// ms_linalg.hpp -- device Krylov solvers on SELL-64 matrices.
#pragma once
#include "ms_common.hpp"

namespace ms {

struct KS {              // device-resident solver scalars (one per solve)
  double bnrm, rz, pap, alpha, beta, res, rho, rho_old, omega, snrm, tt, rr;
  double rmax, delta;    // reliable updates: largest sloppy |r| since the last update, trigger ratio
  double sigma, scale;   // FP32 iteration runs on r/sigma; scale = sigma_old/sigma_new at an update
  double cnu;            // coarse part of r.z: (Z^T r) . E^-1 (Z^T r)
  double tol;
  double rzd, rz_next;   // split coarse finisher: Jacobi part of r.z, r.z waiting to become rz
  int rzp;               // rz_next pending
  int it, ierr, done, maxit;  // done: 0 running, 1 finished, 2 paused for a reliable update
  int coarse;            // coarse correction active
  int ru;                // reliable update requested for this iteration
  int nru;               // reliable updates performed
  int pup;               // a p update is pending for the next product (fused p update)
};

struct SolveResult {
  int it = 0, ierr = 0;
  double res = 0.0;
  int outer = 0;          // refinement cycles / reliable updates (mixed precision)
  int stalled = 0;        // mixed precision: the solve ran or finished in FP64 after an FP32 stall
  int it32 = 0, it64 = 0; // iterations run in FP32 and in FP64 (it32 + it64 = it, bar the exhaustion count)
};

// Stall guard of the mixed-precision PCG: the FP32 iteration gives way to
// FP64 when its residual has not halved over this many iterations (default
// 1024; checked at the end of each 16-iteration chunk).
void set_stall_window(int iters);
int stall_window();
// in-loop timing of the mixed PCG's FP32 SpMV (the last launch of every
// 16-iteration chunk; measurement only)
void set_spmv_probe(bool on);
void read_spmv_probe(long long *launches, double *ms, long long *rows);

// average duration (ms) of a device-to-device copy of `bytes` (read + write)
double time_copy(long long bytes, int reps);
double time_stream(long long bytes, int kind, int reps);

// Matrix-free product with the reduced matrix S = D1 - B^T D0^-1 B of a
// two-colour system (rows [0,n0) colour 0, [n0,n0+n1) colour 1 of the full
// matrix A), for the FP32 inner iterations of the reduced solve: two
// launches over A's rows instead of one over the formed S.  w holds, in A's
// SELL layout, -a_kj / a_kk off the diagonal and (row sum) / a_kk on it for
// the colour-0 rows, -a_ij and the row sum for the colour-1 rows; hx holds
// (h_k, x_ref_k) per colour-0 row between the two launches.
struct MatFree {
  int n0 = 0, n1 = 0;
  const SellPattern *A = nullptr;
  float *w = nullptr;
  float2 *hx = nullptr;
  const int *agg1 = nullptr;   // aggregate of every colour-1 row (coarse space), for the fused p update
};
// the pending p update p = r/d + nu[agg] + beta p, applied inside the next
// matrix-free product instead of its own launch (MIXSOLVER_P_FUSED=1;
// default off: the neighbours' r, d, aggregate and p gathered by the product
// cost more than the launch they save, 15.87 against 17.68 PISO steps/s at
// 2M cells)
bool p_fused_enabled();
void set_p_fused(bool on);
// arguments of the fused p update (beta from the coarse finisher's parts)
struct PUp {
  const float *r, *dg;
  const int *agg1;
  const double *nu, *cnu;
  int G;
};
// the reduced solve's FP32 iterations apply S matrix-free (default on;
// MIXSOLVER_SCHUR_MF=0 or ms_set_schur_matfree(0) forms S for them too)
bool schur_matfree_enabled();
void set_schur_matfree(bool on);

// Preconditioner kinds used internally.
enum PcKind { PC_JACOBI = 0, PC_NATURAL_IC0 = 1, PC_NATURAL_ILU0 = 2, PC_DILU = 3 };

class Krylov {
 public:
  explicit Krylov(const SellPattern *pat);
  ~Krylov();
  // FP64 solvers with the reference recurrences (mod_uns_linsolver_mpi.f90):
  // x is the initial guess on entry.  Returns reference it/res/ierr.
  // coarse: add the aggregate coarse correction of pat->coarse to the
  // preconditioner (PCG only).
  // reuse: the matrix values are those of the previous solve with the same
  // preconditioner (e.g. the second PISO corrector): its set-up (FP32 copy,
  // preconditioner, coarse inverse) is kept.
  SolveResult pcg(const double *val, const double *b, double *x, double tol, int maxit, PcKind pc,
                  bool coarse = false, bool reuse = false);
  // coarse: add the aggregate coarse correction Z E^-1 Z^T to the
  // preconditioner (the p' matrix is symmetric, so E is)
  SolveResult bicgstab(const double *val, const double *b, double *x, double tol, int maxit, PcKind pc,
                       bool coarse = false);
  // Mixed precision.  PCG: FP32 matrix, preconditioner and Krylov vectors
  // with FP32 accumulation, FP64 solution and reliable updates of the true
  // FP64 residual.  BiCGSTAB: FP32 inner solves inside FP64 iterative
  // refinement.  res is the FP64 true relative residual.
  SolveResult mixed(const double *val, const double *b, double *x, double tol, int maxit, PcKind pc,
                    int solver, double inner_tol, bool coarse = false, bool reuse = false);
  void spmv(const double *val, const double *x, double *y);
  // average duration (ms) of the PCG SpMV+dot kernel on matrix val, measured
  // with HIP events on the library stream over `reps` launches; fp32 selects
  // the FP32 difference-form kernel of the mixed path.
  double time_spmv(const double *val, bool fp32, int reps, bool cold = false);
  long long launches = 0;
  bool fp32_stalled() const { return fp32_stalled_; }
  // the FP32 SpMV of the mixed PCG applies this operator instead of val32
  void set_matfree(const MatFree *m) { mf_ = m; }
  bool matfree_on() const { return mf_ && schur_matfree_enabled(); }

  const SellPattern *pat;
  int n;

 private:
  int setup_pc64(PcKind pc, const double *val);
  void setup_pc32(PcKind pc, const double *val);
  template <class T> void apply_pc(PcKind pc, const T *val, const T *fac, const T *dg, const T *in, T *out,
                                   const int *done);
  template <class T> KS read_state(KS *d);
  SolveResult bicg32(PcKind pc, int maxit, int window, bool coarse);
  template <class T> void coarse_apply(const T *in, T *out, const int *done, KS *st);
  void ensure32();
  void coarse_setup(const double *val);   // E = Z^T A Z and its inverse

  DBuf<double> r_, r0_, p_, q_, z_, s_, t_, ph_, sh_, fac_, dg_, part_, prz_;   // prz_: DILU sweep partials
  DBuf<float> val32_, r32_, r032_, p32_, q32_, z32_, s32_, t32_, ph32_, sh32_, x32_, fac32_, dg32_, part32_, prz32_;
  DBuf<KS> st_, st32_;
  DBuf<double> cnu_split_;   // c.nu parts of k_finish_split
  DBuf<float> pq32_;         // the SpMV's p.q partials on the split path
  DBuf<int> flag_;
  CoarseMat cm_;             // E, E^-1 and nu of the coarse space for the current matrix
  DBuf<double> co_part_;     // chunk sums of a BiCGSTAB coarse correction
  const MatFree *mf_ = nullptr;
  // q = op p with the fused p.q partials; start / stop events (may be null)
  // bracket the product's launches
  void spmv32(const double *val64, const float *v32, float *p, float *q, KS *st, float *part, hipEvent_t e0,
              hipEvent_t e1, const PUp *pu = nullptr);
  void *hpin_ = nullptr;
  hipEvent_t state_ev_[2] = {nullptr, nullptr};   // the two pipelined chunk states (mixed PCG)
  void *state_slot(int slot) const { return (char *)hpin_ + 1024 * (slot + 1); }
  int nblk_;
  bool have32_ = false;
  struct SetupKey {
    const double *val = nullptr;
    int pc = -1, fp32 = -1, coarse = -1;
    bool operator==(const SetupKey &o) const {
      return val == o.val && pc == o.pc && fp32 == o.fp32 && coarse == o.coarse;
    }
  } last_;   // the set-up the preconditioner buffers currently hold
  // captured 16-iteration chunk of the mixed PCG: valid while the kernels'
  // arguments (buffers, preconditioner, coarse mode) are those of the key
  struct ChunkKey {
    const void *v32 = nullptr, *fac = nullptr, *dg = nullptr, *nu = nullptr;
    int pc = -1, coarse = -1, fused = -1, len = 0, mf = 0, pf = 0;
    bool operator==(const ChunkKey &o) const {
      return v32 == o.v32 && fac == o.fac && dg == o.dg && nu == o.nu && pc == o.pc && coarse == o.coarse &&
             fused == o.fused && len == o.len && mf == o.mf && pf == o.pf;
    }
  } chunk_key_;
  hipGraphExec_t chunk_exec_ = nullptr;
  bool fp32_stalled_ = false;   // the mixed PCG stalled on this system: FP64 PCG from then on
};

}  // namespace ms
