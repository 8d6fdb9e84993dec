// ms_sweep.hpp -- parameter-sweep server: whole transient cases forked from
// one resident system (host interface of ms_sweep.cpp).
#pragma once
#include <string>
#include <vector>

#include "ms_system.hpp"

namespace ms {

// One case of a sweep: the base system with other boundary groups (inlet
// speeds), run for nsteps PISO steps.  Results are in the caller's natural
// cell order.
struct SweepCase {
  long long id = 0;
  std::vector<ms_bcgroup_t> groups;
  int nsteps = 0, keep = 0;
  int state = 0;                 // 0 queued, 1 running, 2 done, 3 failed
  int ier = 0;                   // the driver's ier of the last step
  int err = 0;                   // MS_E* of a failed case
  std::string msg;
  double t_submit = 0.0, t_admit = 0.0, t_done = 0.0;   // seconds since the sweep started
  std::vector<double> step_ms;   // wall time of each step (launch to completion)
  std::vector<double> history;   // 8 per step, as ms_get_history
  long long stats[8] = {};       // as ms_solver_stats
  std::vector<double> u, p, T, T_s;   // final fields (keep != 0)
};

class Sweep;
Sweep *sweep_create(System &base, long long budget, int max_active);
void sweep_destroy(Sweep *s);
// group: 1-based boundary group whose speed the case sets (0: none):
// velocity inlet / far field: uvel; parabolic inlet: mean speed `speed`;
// mass inlet: speed
long long sweep_submit(Sweep *s, int group, const double *uvel, double speed, int nsteps, int keep);
// state of the case after waiting up to wait_s seconds for it to end; out
// receives a copy of the case
int sweep_result(Sweep *s, long long id, double wait_s, SweepCase *out);
void sweep_stats(Sweep *s, long long *out10);

}  // namespace ms
