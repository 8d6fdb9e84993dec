// ms_linalg.hpp -- device Krylov solvers on SELL-64 matrices.
#pragma once
#include "ms_common.hpp"

namespace ms {

template <class T>
struct KS {              // device-resident solver scalars (one per solve)
  T bnrm, rz, pap, alpha, beta, res, rho, rho_old, omega, snrm, tt, rr;
  int it, ierr, done, maxit;
  T tol;
  unsigned int ticket;
  int pad;
};

struct SolveResult {
  int it = 0, ierr = 0;
  double res = 0.0;
  int outer = 0;          // refinement cycles (mixed precision)
};

// Preconditioner kinds used internally.
enum PcKind { PC_JACOBI = 0, PC_NATURAL_IC0 = 1, PC_NATURAL_ILU0 = 2, PC_DILU = 3 };

class Krylov {
 public:
  explicit Krylov(SellPattern *pat);
  ~Krylov();
  // FP64 solvers with the reference recurrences (mod_uns_linsolver_mpi.f90):
  // x is the initial guess on entry.  Returns reference it/res/ierr.
  SolveResult pcg(const double *val, const double *b, double *x, double tol, int maxit, PcKind pc);
  SolveResult bicgstab(const double *val, const double *b, double *x, double tol, int maxit, PcKind pc);
  // FP32 inner Krylov inside FP64 iterative refinement; res = FP64 true
  // relative residual.  solver 0 = PCG, 1 = BiCGSTAB.
  SolveResult mixed(const double *val, const double *b, double *x, double tol, int maxit, PcKind pc,
                    int solver, double inner_tol);
  // y = A x in FP64 (for tests / residual checks)
  void spmv(const double *val, const double *x, double *y);
  double true_residual(const double *val, const double *b, const double *x);  // ||b-Ax||/||b||
  long long launches = 0;  // kernel launches issued (diagnostics)

  SellPattern *pat;
  int n;

 private:
  int setup_pc64(PcKind pc, const double *val);           // returns factorization ierr
  void apply_pc64(PcKind pc, const double *val, const double *in, double *out);
  void setup_pc32(PcKind pc, const double *val);
  void apply_pc32(PcKind pc, const float *in, float *out);
  template <class T> KS<T> read_state(KS<T> *d);
  SolveResult inner32(int solver, PcKind pc, double inner_tol, int maxit);

  DBuf<double> r_, r0_, p_, q_, z_, s_, t_, ph_, sh_, fac_, dinv_, part_;
  DBuf<float> val32_, b32_, x32_, r32_, r032_, p32_, q32_, z32_, s32_, t32_, ph32_, sh32_, fac32_, dinv32_, part32_;
  DBuf<KS<double>> st_;
  DBuf<KS<float>> st32_;
  DBuf<int> flag_;
  void *hpin_ = nullptr;   // pinned host staging for state reads
  int nblk_;
  bool have32_ = false;
};

}  // namespace ms
