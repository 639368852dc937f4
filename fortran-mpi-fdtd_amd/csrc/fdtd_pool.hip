// fdtd_pool.hip -- serving layer: many independent simulations on one GPU.
//
// Jobs are set-up handles plus a step count.  A job waits until its device
// footprint fits the HBM budget and a slot is free; admitted jobs each run on
// their own HIP stream and advance in chunks, up to two queued per job, so
// the chunks of co-resident jobs overlap on the device.
// After every chunk a check on the job's stream flags NaN or Inf in its E
// and H; the flag comes back with the chunk's event, and a job whose fields
// are no longer finite ends FAILED (FDTD_ERR_NONFINITE) without touching
// the others.  A finished job keeps its memory (and its budget share) until
// the caller releases it after reading its results.  No counterpart in the
// reference.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <chrono>
#include <deque>
#include <vector>

#include "fdtd_internal.h"

namespace fdtd {

namespace {

double now_ms() {
  using clk = std::chrono::steady_clock;
  return std::chrono::duration<double, std::milli>(clk::now().time_since_epoch()).count();
}

struct Job {
  fdtd_sim* sim;
  int nsteps = 0, left = 0;
  int state = FDTD_JOB_WAITING;
  size_t bytes = 0;
  bool counted = false;
  double t_submit = 0, t_start = 0, t_done = 0;
  // chunks queued and not yet seen complete: events ev[head], ev[head^1]
  hipEvent_t ev[2] = {nullptr, nullptr};
  int head = 0, inflight = 0;
  bool cancel = false;
  int err = 0;
  // finite-field sentinel: device flag (on the handle, set by the check
  // after each chunk) and its pinned host copies, one per queued chunk
  int* d_flag = nullptr;
  int* h_flag = nullptr;
};

// chunks kept queued per running job: the next one is on the device while
// the host polls, so a busy host loop does not leave the GPU idle
constexpr int kAhead = 2;

}  // namespace
}  // namespace fdtd

using namespace fdtd;

struct fdtd_pool {
  size_t budget = 0, used = 0;
  int max_active = 1, chunk = 64;
  std::vector<Job> jobs;      // by ticket
  std::deque<int> waiting;    // FIFO of tickets
  std::vector<int> running;   // tickets
};

namespace {

// Poll a running job: retire completed chunks in order, then keep kAhead
// chunks queued while steps remain.
int advance(fdtd_pool* p, Job& j) {
  Sim* s = j.sim;
  while (j.inflight > 0) {
    const hipError_t q = hipEventQuery(j.ev[j.head]);
    if (q == hipErrorNotReady) break;
    if (q != hipSuccess) {
      j.err = hip_check(q, "job stream");
      j.state = FDTD_JOB_FAILED;
      j.t_done = now_ms();
      return 0;
    }
    if (j.h_flag[j.head]) {  // this chunk left NaN or Inf in the fields
      j.err = set_error(FDTD_ERR_NONFINITE, "served job: fields are no longer finite");
      j.state = FDTD_JOB_FAILED;
      j.t_done = now_ms();
      return 0;
    }
    j.head ^= 1;
    --j.inflight;
  }
  if (j.cancel || j.left == 0) {
    if (j.inflight == 0) {
      j.state = j.cancel ? FDTD_JOB_CANCELLED : FDTD_JOB_DONE;
      j.t_done = now_ms();
    }
    return 0;
  }
  while (j.inflight < kAhead && j.left > 0) {
    const int k = std::min(j.left, p->chunk);
    int rc = fdtd_step(j.sim, k);
    if (rc) {
      j.err = rc;
      j.state = FDTD_JOB_FAILED;
      j.t_done = now_ms();
      return 0;
    }
    j.left -= k;
    const int slot = (j.head + j.inflight) & 1;
    // the finite-field check of this chunk, its flag copied back before the
    // chunk's event
    rc = launch_nonfinite(s, j.d_flag);
    if (rc == 0 && hipMemcpyAsync(&j.h_flag[slot], j.d_flag, sizeof(int), hipMemcpyDeviceToHost, s->stream) !=
                       hipSuccess)
      rc = set_error(FDTD_ERR_HIP, "sentinel copy");
    if (rc) {
      j.err = rc;
      j.state = FDTD_JOB_FAILED;
      j.t_done = now_ms();
      return 0;
    }
    if (!j.ev[slot] && hipEventCreateWithFlags(&j.ev[slot], hipEventDisableTiming) != hipSuccess) {
      j.err = set_error(FDTD_ERR_HIP, "event create");
      j.state = FDTD_JOB_FAILED;
      return 0;
    }
    if (hipEventRecord(j.ev[slot], s->stream) != hipSuccess) {
      j.err = set_error(FDTD_ERR_HIP, "event record");
      j.state = FDTD_JOB_FAILED;
      return 0;
    }
    ++j.inflight;
  }
  return 0;
}

}  // namespace

extern "C" {

int fdtd_estimate_bytes(fdtd_sim* sim, size_t* bytes) {
  if (!sim || !bytes) return set_error(FDTD_ERR_ARG, "null argument");
  Sim* s = (Sim*)sim;
  *bytes = s->committed ? s->dev_bytes : estimate_bytes(s);
  return 0;
}

int fdtd_pool_create(fdtd_pool** out, size_t hbm_budget, int max_active, int chunk_steps) {
  if (!out || hbm_budget == 0 || max_active < 1 || chunk_steps < 1) return set_error(FDTD_ERR_ARG, "bad pool");
  fdtd_pool* p = new fdtd_pool();
  p->budget = hbm_budget;
  p->max_active = max_active;
  // whole DFT blocks: a chunk then ends where a run stepped at once folds
  // its spectra, so a served job's results equal its run alone bit for bit
  const int64_t whole = ((int64_t)chunk_steps + kDftBlock - 1) / kDftBlock * kDftBlock;
  p->chunk = (int)std::min<int64_t>(whole, (int64_t)INT_MAX / kDftBlock * kDftBlock);
  *out = p;
  return 0;
}

int fdtd_pool_destroy(fdtd_pool* p) {
  if (!p) return 0;
  // events of jobs never released (their handles must still be alive)
  for (auto& j : p->jobs) {
    for (int e = 0; e < 2; ++e)
      if (j.ev[e]) {
        (void)hipEventSynchronize(j.ev[e]);
        (void)hipEventDestroy(j.ev[e]);
      }
    if (j.h_flag) (void)hipHostFree(j.h_flag);
  }
  delete p;
  return 0;
}

int fdtd_pool_submit(fdtd_pool* p, fdtd_sim* sim, int nsteps, int* ticket) {
  if (!p || !sim || !ticket || nsteps < 0) return set_error(FDTD_ERR_ARG, "bad argument");
  Sim* s = (Sim*)sim;
  Job j;
  j.sim = sim;
  j.nsteps = j.left = nsteps;
  j.bytes = s->committed ? s->dev_bytes : estimate_bytes(s);
  if (j.bytes > p->budget) return set_error(FDTD_ERR_NOMEM, "job larger than the pool's HBM budget");
  j.t_submit = now_ms();
  p->jobs.push_back(j);
  *ticket = (int)p->jobs.size() - 1;
  p->waiting.push_back(*ticket);
  return 0;
}

int fdtd_pool_progress(fdtd_pool* p, int* n_running, int* n_waiting) {
  if (!p) return set_error(FDTD_ERR_ARG, "null pool");
  // advance running jobs
  std::vector<int> still;
  for (int t : p->running) {
    Job& j = p->jobs[t];
    advance(p, j);
    if (j.state == FDTD_JOB_RUNNING) still.push_back(t);
  }
  p->running.swap(still);
  // admit waiting jobs in arrival order while memory and slots allow
  while (!p->waiting.empty() && (int)p->running.size() < p->max_active) {
    Job& j = p->jobs[p->waiting.front()];
    if (p->used + j.bytes > p->budget) break;
    p->waiting.pop_front();
    int rc = fdtd_commit(j.sim);
    j.t_start = now_ms();
    if (rc) {
      j.err = rc;
      j.state = FDTD_JOB_FAILED;
      j.t_done = j.t_start;
      continue;
    }
    Sim* s = (Sim*)j.sim;
    j.d_flag = (int*)dev_alloc(s, sizeof(int));  // zero-filled, on the handle
    if (!j.d_flag || hipHostMalloc((void**)&j.h_flag, 2 * sizeof(int), hipHostMallocDefault) != hipSuccess) {
      j.h_flag = nullptr;
      j.err = set_error(FDTD_ERR_NOMEM, "finite-field sentinel");
      j.state = FDTD_JOB_FAILED;
      j.t_done = j.t_start;
      continue;
    }
    j.h_flag[0] = j.h_flag[1] = 0;
    j.bytes = s->dev_bytes;
    p->used += j.bytes;
    j.counted = true;
    j.state = FDTD_JOB_RUNNING;
    advance(p, j);
    if (j.state == FDTD_JOB_RUNNING) p->running.push_back((int)(&j - p->jobs.data()));
  }
  if (n_running) *n_running = (int)p->running.size();
  if (n_waiting) *n_waiting = (int)p->waiting.size();
  return 0;
}

int fdtd_pool_status(fdtd_pool* p, int ticket, int* state, double* wait_ms, double* run_ms) {
  if (!p || ticket < 0 || ticket >= (int)p->jobs.size()) return set_error(FDTD_ERR_ARG, "bad ticket");
  const Job& j = p->jobs[ticket];
  if (state) *state = j.state;
  const double t = now_ms();
  if (wait_ms) *wait_ms = (j.state == FDTD_JOB_WAITING ? t : j.t_start) - j.t_submit;
  if (run_ms)
    *run_ms = j.state == FDTD_JOB_WAITING ? 0.0 : ((j.state == FDTD_JOB_RUNNING ? t : j.t_done) - j.t_start);
  if (j.state == FDTD_JOB_FAILED) return j.err ? j.err : FDTD_ERR_STATE;
  return 0;
}

int fdtd_pool_cancel(fdtd_pool* p, int ticket) {
  if (!p || ticket < 0 || ticket >= (int)p->jobs.size()) return set_error(FDTD_ERR_ARG, "bad ticket");
  Job& j = p->jobs[ticket];
  if (j.state == FDTD_JOB_WAITING) {
    p->waiting.erase(std::remove(p->waiting.begin(), p->waiting.end(), ticket), p->waiting.end());
    j.state = FDTD_JOB_CANCELLED;
    j.t_start = j.t_done = now_ms();
  } else if (j.state == FDTD_JOB_RUNNING) {
    j.cancel = true;
  }
  return 0;
}

int fdtd_pool_release(fdtd_pool* p, int ticket) {
  if (!p || ticket < 0 || ticket >= (int)p->jobs.size()) return set_error(FDTD_ERR_ARG, "bad ticket");
  Job& j = p->jobs[ticket];
  if (j.state == FDTD_JOB_RUNNING) return set_error(FDTD_ERR_STATE, "job still running");
  if (j.state == FDTD_JOB_WAITING) {
    p->waiting.erase(std::remove(p->waiting.begin(), p->waiting.end(), ticket), p->waiting.end());
    j.state = FDTD_JOB_FAILED;
    j.err = FDTD_ERR_STATE;
  }
  if (j.counted) {
    p->used -= j.bytes;
    j.counted = false;
  }
  // the chunk events go now, while the job's stream still exists: the
  // caller may destroy the handle next, and an event must not outlive the
  // stream it was recorded on
  for (int e = 0; e < 2; ++e)
    if (j.ev[e]) {
      (void)hipEventSynchronize(j.ev[e]);
      (void)hipEventDestroy(j.ev[e]);
      j.ev[e] = nullptr;
    }
  if (j.h_flag) {  // its copies have landed: the events above are complete
    (void)hipHostFree(j.h_flag);
    j.h_flag = nullptr;
  }
  j.inflight = 0;
  return 0;
}

int fdtd_pool_usage(fdtd_pool* p, size_t* used, size_t* budget) {
  if (!p) return set_error(FDTD_ERR_ARG, "null pool");
  if (used) *used = p->used;
  if (budget) *budget = p->budget;
  return 0;
}

}  // extern "C"
