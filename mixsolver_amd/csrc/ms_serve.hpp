// ms_serve.hpp -- continuous-batching request server for pressure-correction
// solves (host interface of ms_serve.hip).
#pragma once
#include <memory>

#include "ms_system.hpp"

namespace ms {

struct ServeOpts {
  long long budget = 0;      // bytes of device memory admitted requests may hold
  int max_active = 64;       // requests in flight, packed into every launch
  int precond = MS_PC_DILU;  // MS_PC_JACOBI or MS_PC_DILU
  int coarse = 0;            // aggregates per pattern (needs cell centres), 0 = off
  int round_iters = 16;      // iterations per scheduling round
  double delta = 0.1;        // reliable-update trigger of the FP32 iteration
};

struct ServeResult {
  int it = 0, ierr = 0;
  double res = 0.0;
  double t_admit = 0.0, t_done = 0.0;   // seconds after submission
};

class Server;
Server *server_create(const ServeOpts &o);
void server_destroy(Server *s);
int server_add_pattern(Server *s, int n, const int *row_ptr, const int *col_idx, const double *xc);
long long server_submit(Server *s, int pattern, const double *val, const double *b, double tol, int maxit);
// true when the result is ready (x in the caller's numbering)
bool server_result(Server *s, long long ticket, double timeout_ms, double *x, ServeResult *r);
void server_stats(Server *s, long long *out8);

}  // namespace ms
