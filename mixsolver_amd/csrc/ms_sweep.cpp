// ms_sweep.cpp -- parameter-sweep server: many transient cases on one mesh.
//
// The cases of a sweep (the same mesh, zones and control, different inlet
// speeds) are forks of one resident base system: each views the base's
// mesh, geometry, face constants, zone properties and matrix patterns and
// owns only its fields, matrices and solver state, so the HBM a case holds
// is what differs between cases.  Admission is by device-memory budget: the
// first case runs alone until its first step has allocated everything a
// case allocates (its footprint, measured by the device bytes its worker
// thread holds); later cases are admitted while the cases in flight plus one
// more footprint fit the budget, and a case that finishes frees its share
// for the next.  Every admitted case runs on a worker thread with a HIP
// stream of its own (default priority), so the kernels of concurrent cases
// overlap on the device and each case's host-side waits (solver state reads
// between chunks, step diagnostics) hide behind the other cases' work.  A
// case's arithmetic does not depend on the mix: its kernels read only its
// own buffers and the shared read-only data, and every reduction has a fixed
// order, so a case's fields are the same bit for bit alone and in any mix.
//
// Per case and step the server records the wall time of the step (launch to
// completion: the per-step latency a case sees while it shares the device)
// and the driver's history row.  The per-case physics is the partitioned
// transient driver at one partition (piso_run_mpi, mod_uns_simple_mpi.f90:298-578).
#include "ms_sweep.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

namespace ms {

int run_piso(System &s, int nsteps, bool resume);

namespace {
using Clock = std::chrono::steady_clock;
double secs(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double>(b - a).count(); }
}  // namespace

class Sweep {
 public:
  Sweep(System &base, long long budget, int max_active) : base_(base), budget_(budget) {
    if (base.base) fail(MS_EINVAL, "ms_sweep_create: the base is itself a sweep case");
    if (!base.ctrl.transient) fail(MS_EINVAL, "ms_sweep_create: the base system's control is not transient");
    if (max_active < 1 || max_active > 64) fail(MS_EINVAL, "ms_sweep_create: max_active in 1..64");
    dev_ = runtime_device();
    // everything a fork reads from the base is final before any worker runs
    base.refresh_face_constants();
    (void)base.min_vol13();
    MS_HIP(hipStreamSynchronize(stream()));
    if (budget_ <= 0) {
      size_t fr = 0, tot = 0;
      MS_HIP(hipMemGetInfo(&fr, &tot));
      budget_ = (long long)(0.9 * (double)fr);
    }
    shared_ = (long long)base.device_bytes();
    base.nforks++;   // the base stays as it is while the sweep exists
    t0_ = Clock::now();
    for (int i = 0; i < max_active; i++) workers_.emplace_back([this] { worker(); });
  }

  ~Sweep() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto &t : workers_) t.join();
    base_.nforks--;
  }

  long long submit(int group, const double *uvel, double speed, int nsteps, int keep) {
    if (nsteps < 1) fail(MS_EINVAL, "ms_sweep_submit: nsteps >= 1");
    std::vector<ms_bcgroup_t> g = base_.groups;
    if (group < 0 || group > (int)g.size()) fail(MS_EINVAL, "ms_sweep_submit: group out of range");
    if (group > 0) {
      ms_bcgroup_t &b = g[group - 1];
      switch (b.btype) {
        case MS_BC_VINLET:
        case MS_BC_FARFIELD:
          if (!uvel) fail(MS_EINVAL, "ms_sweep_submit: uvel needed for a velocity inlet / far-field group");
          for (int d = 0; d < 3; d++) b.uvel[d] = uvel[d];
          break;
        case MS_BC_VINLET_PARAB:
          b.prfl_mean = speed;
          break;
        case MS_BC_MASSINLET:
          b.uspeed = speed;
          break;
        default:
          fail(MS_EINVAL, "ms_sweep_submit: group " + std::to_string(group) + " is not an inlet or far-field group");
      }
    }
    auto c = std::make_unique<SweepCase>();
    c->groups = std::move(g);
    c->nsteps = nsteps;
    c->keep = keep;
    std::lock_guard<std::mutex> lk(mu_);
    if (stop_) fail(MS_EINVAL, "ms_sweep_submit: the sweep is shutting down");
    c->id = next_id_++;
    c->t_submit = secs(t0_, Clock::now());
    SweepCase *raw = c.get();
    cases_[raw->id] = std::move(c);
    queue_.push_back(raw);
    submitted_++;
    cv_.notify_all();
    return raw->id;
  }

  int result(long long id, double wait_s, SweepCase *out) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = cases_.find(id);
    if (it == cases_.end()) fail(MS_EINVAL, "ms_sweep_result: unknown case id");
    SweepCase *c = it->second.get();
    if (wait_s > 0.0)
      done_cv_.wait_for(lk, std::chrono::duration<double>(wait_s), [&] { return c->state >= 2; });
    if (out && c->state >= 2) *out = *c;
    else if (out) {
      out->state = c->state;
      out->id = c->id;
    }
    return c->state;
  }

  void stats(long long *o) {
    std::lock_guard<std::mutex> lk(mu_);
    o[0] = submitted_;
    o[1] = done_;
    o[2] = failed_;
    o[3] = (long long)queue_.size();
    o[4] = active_;
    o[5] = peak_;
    o[6] = footprint_;
    o[7] = budget_;
    o[8] = shared_;
    o[9] = steps_;
  }

 private:
  // a case may start: the first runs alone until its footprint is known;
  // then while the reserved footprints plus one more fit the budget (a case
  // alone always starts)
  bool admissible() const {
    if (queue_.empty() || probing_) return false;
    if (active_ == 0) return true;
    if (footprint_ <= 0) return false;
    return reserved_ + reserve_of(footprint_) <= budget_;
  }
  static long long reserve_of(long long fp) { return fp + fp / 32; }   // allocation granularity margin

  void worker() {
    if (hipSetDevice(dev_ < 0 ? 0 : dev_) != hipSuccess) return;
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return;
    set_thread_stream(st);
    for (;;) {
      SweepCase *c = nullptr;
      long long held = 0;   // bytes this case has reserved
      bool probe = false;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return stop_ || admissible(); });
        if (stop_) break;
        c = queue_.front();
        queue_.pop_front();
        active_++;
        peak_ = std::max(peak_, active_);
        if (footprint_ > 0) {
          held = reserve_of(footprint_);
          reserved_ += held;
        } else {
          probe = true;
          probing_ = true;
        }
        c->state = 1;
        c->t_admit = secs(t0_, Clock::now());
      }
      const long long b0 = thread_device_bytes();
      long long fp = 0;
      int state = 2, steps = 0;
      try {
        std::unique_ptr<System> f;
        {
          std::lock_guard<std::recursive_mutex> g(device_alloc_mutex());
          f.reset(fork_system(base_, c->groups));
        }
        for (int k = 0; k < c->nsteps; k++) {
          const auto a = Clock::now();
          const int ier = run_piso(*f, 1, k > 0);
          c->step_ms.push_back(1e3 * secs(a, Clock::now()));
          c->history.insert(c->history.end(), f->history.begin(), f->history.end());
          steps++;
          if (k == 0) {
            fp = thread_device_bytes() - b0;
            if (probe) {
              std::lock_guard<std::mutex> lk(mu_);
              footprint_ = std::max(footprint_, fp);
              held = reserve_of(footprint_);
              reserved_ += held;
              probing_ = false;
              probe = false;
              cv_.notify_all();
            }
          }
          c->ier = ier;
          if (ier) break;
        }
        c->stats[0] = f->st_ppe_solves;
        c->stats[1] = f->st_it32;
        c->stats[2] = f->st_it64;
        c->stats[3] = f->st_fallback;
        c->stats[4] = (f->kry && f->kry->fp32_stalled()) || (f->schur && f->schur->kry->fp32_stalled()) ? 1 : 0;
        c->stats[5] = stall_window();
        c->stats[6] = f->st_mom_solves;
        c->stats[7] = f->st_mom_its;
        fp = std::max(fp, thread_device_bytes() - b0);
        if (c->keep) keep_fields(*f, *c);
        std::lock_guard<std::recursive_mutex> g(device_alloc_mutex());
        f.reset();
      } catch (const Error &e) {
        state = 3;
        c->err = e.code;
        c->msg = e.msg;
      } catch (const std::bad_alloc &) {
        state = 3;
        c->err = MS_ENOMEM;
        c->msg = "host out of memory";
      } catch (...) {
        state = 3;
        c->err = MS_EINVAL;
        c->msg = "unknown error";
      }
      {
        std::lock_guard<std::mutex> lk(mu_);
        if (probe) probing_ = false;   // failed before its footprint was known
        if (fp > footprint_ && state == 2) footprint_ = fp;
        reserved_ -= held;
        active_--;
        steps_ += steps;
        (state == 2 ? done_ : failed_)++;
        c->state = state;
        c->t_done = secs(t0_, Clock::now());
      }
      cv_.notify_all();
      done_cv_.notify_all();
    }
    set_thread_stream(nullptr);
    (void)hipStreamDestroy(st);
  }

  // final fields in the natural cell order
  void keep_fields(System &f, SweepCase &c) {
    const int nc = f.nc;
    std::vector<double> d;
    auto get = [&](const DBuf<double> &b, int w, std::vector<double> &out) {
      d.resize((size_t)nc * w);
      b.download(d.data(), d.size());
      MS_HIP(hipStreamSynchronize(stream()));
      out.resize(d.size());
      for (int i = 0; i < nc; i++)
        for (int j = 0; j < w; j++) out[(size_t)i * w + j] = d[(size_t)base_.o2n[i] * w + j];
    };
    get(f.u, 3, c.u);
    get(f.p, 1, c.p);
    get(f.T, 1, c.T);
    get(f.T_s, 1, c.T_s);
  }

  System &base_;
  long long budget_ = 0, shared_ = 0;
  int dev_ = -1;
  Clock::time_point t0_;
  std::vector<std::thread> workers_;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  std::deque<SweepCase *> queue_;
  std::map<long long, std::unique_ptr<SweepCase>> cases_;
  long long next_id_ = 1, submitted_ = 0, done_ = 0, failed_ = 0, steps_ = 0;
  long long footprint_ = 0, reserved_ = 0;
  long long active_ = 0, peak_ = 0;
  bool probing_ = false, stop_ = false;
};

Sweep *sweep_create(System &base, long long budget, int max_active) { return new Sweep(base, budget, max_active); }
void sweep_destroy(Sweep *s) { delete s; }
long long sweep_submit(Sweep *s, int group, const double *uvel, double speed, int nsteps, int keep) {
  return s->submit(group, uvel, speed, nsteps, keep);
}
int sweep_result(Sweep *s, long long id, double wait_s, SweepCase *out) { return s->result(id, wait_s, out); }
void sweep_stats(Sweep *s, long long *out10) { s->stats(out10); }

}  // namespace ms
