// libtbec time stepping: real-time RK3, renormalised imaginary-time RK3 and the
// reference's imaginary Euler, each step = 3 (or 1) fused stage launches.  Runs of
// steps are captured once into a hipGraph and replayed (the graph holds only kernel
// nodes; device state is updated by kernels, never by memset nodes).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <string>

#include "gpe_stage.hpp"
#include "tbec_internal.hpp"

using namespace tbec;

#include <chrono>
#include <cstdlib>

double tbec::progress_interval() {
  const char *e = std::getenv("TBEC_PROGRESS");
  if (!e) return -1.0;
  const double v = std::atof(e);
  return v > 0 ? v : 10.0;
}

#define HIPCHK2(x)                                                                     \
  do {                                                                                 \
    hipError_t e_ = (x);                                                               \
    if (e_ != hipSuccess)                                                              \
      return set_error(TBEC_ERR_HIP, std::string(#x " failed: ") + hipGetErrorString(e_)); \
  } while (0)

// Graph replays are asynchronous: a replay can still be in flight when a caller drops
// the graphs (a physics change mid-run, tuning, destroy), so the stream is drained
// before any executable graph is destroyed.
void tbec_drop_graphs(tbec_handle *h) {
  if (h->graphs.empty()) return;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (auto &kv : h->graphs) (void)hipGraphExecDestroy(kv.second);
  h->graphs.clear();
}

// Enqueue `n` steps of `mode` directly (no graph).
static int enqueue_steps(tbec_handle *h, long long n, int mode) {
  for (long long s = 0; s < n; ++s) {
    int rc = 0;
    if (mode == TBEC_REAL_RK3) {
      rc = enqueue_stage(h, ST_RK1, false, false);
      if (!rc) rc = enqueue_stage(h, ST_RK2, false, false);
      if (!rc) rc = enqueue_stage(h, ST_RK3, false, false);
    } else if (mode == TBEC_IMAG_RK3) {
      rc = enqueue_stage(h, ST_RK1, true, false);
      if (!rc) rc = enqueue_stage(h, ST_RK2, true, false);
      if (!rc) rc = enqueue_stage(h, ST_RK3, true, false);
    } else {
      rc = enqueue_stage(h, ST_EULER, false, false);
    }
    if (rc) return rc;
  }
  return 0;
}

// Replay (capturing on first use) a graph of `n` steps starting from the current
// parity.  n is even for the 3-stage modes so the parity is restored.
static int replay_chunk(tbec_handle *h, long long n, int mode) {
  const auto key = std::make_tuple(mode, n, h->cur);
  auto it = h->graphs.find(key);
  if (it == h->graphs.end()) {
    const int cur0 = h->cur;
    // no allocation may happen inside the capture: size the reduction buffer first
    int rc0 = tbec_reserve_partials(h, fill_args_count(h));
    if (rc0) return rc0;
    HIPCHK2(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_steps(h, n, mode);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(h->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      h->cur = cur0;
      return rc;
    }
    if (e != hipSuccess) return set_error(TBEC_ERR_HIP, std::string("graph capture failed: ") +
                                                            hipGetErrorString(e));
    hipGraphExec_t ex = nullptr;
    e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return set_error(TBEC_ERR_HIP, std::string("graph instantiate: ") +
                                                            hipGetErrorString(e));
    h->cur = cur0;  // capture recorded the parity flips; the replay performs them
    it = h->graphs.emplace(key, ex).first;
  }
  HIPCHK2(hipGraphLaunch(it->second, h->stream));
  const int flips = mode == TBEC_IMAG_EULER_REF ? (int)(n & 1) : (int)((3 * n) & 1);
  h->cur ^= flips;
  return 0;
}

static const long long kChunk = 16;

static int run_steps(tbec_handle *h, long long nsteps, int mode, long long chunk) {
  if (chunk > 1 && (chunk & 1)) ++chunk;
  long long done = 0;
  if (chunk > 1) {
    while (nsteps - done >= chunk) {
      int rc = replay_chunk(h, chunk, mode);
      if (rc) return rc;
      done += chunk;
    }
  }
  return enqueue_steps(h, nsteps - done, mode);
}

extern "C" {

int tbec_step(tbec_handle *h, int64_t nsteps, int mode, double *t_inout) {
  if (!h) return set_error(TBEC_ERR_ARG, "null handle");
  if (nsteps < 0) return set_error(TBEC_ERR_ARG, "nsteps must be >= 0");
  if (mode != TBEC_REAL_RK3 && mode != TBEC_IMAG_RK3 && mode != TBEC_IMAG_EULER_REF)
    return set_error(TBEC_ERR_ARG, "unknown step mode");
  if (t_inout) {
    int rc = tbec_set_time(h, *t_inout);
    if (rc) return rc;
  }
  if (mode == TBEC_IMAG_RK3) {
    if (h->p.dt_re != 0.0 || !(h->p.dt_im < 0.0))
      return set_error(TBEC_ERR_ARG, "TBEC_IMAG_RK3 needs dt = (0, -dtau) with dtau > 0");
    // a step-limited run: no convergence stop, flags cleared; scale, c0 and the step
    // counter continue from the previous imaginary-time call
    h->tol = 0.0;
    int rc = prepare_physics(h, h->t_now);   // imaginary time does not advance t
    if (rc) return rc;
    const double zero = 0.0;
    const int zflags[4] = {0, 0, 0, 0};
    HIPCHK2(hipMemcpyAsync(h->dstate + ST_TOL, &zero, sizeof(double), hipMemcpyHostToDevice,
                           h->stream));
    HIPCHK2(hipMemcpyAsync(h->dflag, zflags, sizeof zflags, hipMemcpyHostToDevice, h->stream));
    HIPCHK2(hipStreamSynchronize(h->stream));
    h->scale_pending = true;
    rc = run_steps(h, nsteps, mode, nsteps >= kChunk ? kChunk : 0);
    if (rc) return rc;
    if (t_inout) *t_inout = h->t_now;
    // the finalisers flag a non-finite norm or energy (and every later stage skips)
    HIPCHK2(hipMemcpyAsync(h->hflag, h->dflag, sizeof(int) * 4, hipMemcpyDeviceToHost,
                           h->stream));
    HIPCHK2(hipStreamSynchronize(h->stream));
    if (h->hflag[0] == 2) {
      h->scale_pending = false;
      return set_error(TBEC_ERR_NONFINITE, "non-finite norm or energy in imaginary time");
    }
    return tbec_materialize_scale(h);
  }
  int rc = tbec_materialize_scale(h);
  if (rc) return rc;
  if (mode == TBEC_IMAG_EULER_REF) {
    rc = prepare_physics(h, h->t_now);
    if (!rc) rc = run_steps(h, nsteps, mode, nsteps >= kChunk ? kChunk : 0);
    if (!rc && t_inout) *t_inout = h->t_now;
    return rc;
  }
  // Real time: each step sees the potential, frame term and switch-offs of the time it
  // starts at; runs of steps whose physics agrees share one table and replay graphs.
  double t = h->t_now;
  const bool moves = physics_time_dependent(h);
  long long done = 0;
  while (done < nsteps) {
    rc = prepare_physics(h, t);
    if (rc) return rc;
    long long n = 1;
    double tn = advance_time(h, t);
    if (!moves) {
      for (; done + n < nsteps; ++n) tn = advance_time(h, tn);
    } else {
      while (done + n < nsteps && same_step_physics(h, t, tn)) {
        ++n;
        tn = advance_time(h, tn);
      }
    }
    rc = run_steps(h, n, mode, n >= kChunk ? kChunk : 0);
    if (rc) return rc;
    done += n;
    t = tn;
  }
  h->t_now = t;
  if (t_inout) *t_inout = t;
  return 0;
}

int tbec_imag_solve(tbec_handle *h, int64_t max_steps, double tol, double norm,
                    int check_every, tbec_solve_result *out) {
  if (!h || !out) return set_error(TBEC_ERR_ARG, "null argument");
  if (h->p.dt_re != 0.0 || !(h->p.dt_im < 0.0))
    return set_error(TBEC_ERR_ARG, "imaginary time needs dt = (0, -dtau) with dtau > 0");
  if (!(norm > 0) || max_steps < 1 || check_every < 1)
    return set_error(TBEC_ERR_ARG, "norm, max_steps and check_every must be positive");
  int rc = tbec_materialize_scale(h);
  if (!rc) rc = prepare_physics(h, h->t_now);   // imaginary time does not advance t
  if (rc) return rc;
  h->norm_target = norm;
  h->tol = tol;
  rc = tbec_reset_state(h);  // scale 1, c0_eff = c0, step 0, flag 0
  if (rc) return rc;
  h->scale_pending = true;
  long long chunk = check_every;
  if (chunk & 1) ++chunk;
  long long done = 0;
  int flag = 0;
  const double every = progress_interval();
  auto last = std::chrono::steady_clock::now();
  while (done < max_steps) {
    long long n = std::min<long long>(chunk, max_steps - done);
    if (n == chunk && chunk > 1) rc = replay_chunk(h, n, TBEC_IMAG_RK3);
    else rc = enqueue_steps(h, n, TBEC_IMAG_RK3);
    if (rc) return rc;
    done += n;
    HIPCHK2(hipMemcpyAsync(h->hflag, h->dflag, sizeof(int) * 4, hipMemcpyDeviceToHost,
                           h->stream));
    HIPCHK2(hipStreamSynchronize(h->stream));
    flag = h->hflag[0];
    if (flag) break;
    if (every > 0) {
      const auto now = std::chrono::steady_clock::now();
      if (std::chrono::duration<double>(now - last).count() >= every) {
        std::fprintf(stderr, "[tbec] imaginary-time solve: %lld steps\n", done);
        last = now;
      }
    }
  }
  HIPCHK2(hipMemcpyAsync(h->hstate, h->dstate, sizeof(double) * ST_COUNT,
                         hipMemcpyDeviceToHost, h->stream));
  HIPCHK2(hipStreamSynchronize(h->stream));
  const double *s = h->hstate;
  out->steps = (int64_t)s[ST_STEP];
  out->converged = flag == 1;
  out->energy = s[ST_E];
  out->mu = s[ST_MU];
  out->residual = s[ST_RES];
  out->norm = s[ST_N];
  out->rel_change = s[ST_REL];
  if (flag == 1) {
    // the converged psi_n is the input of the skipped step: buffer parity of step n
    const long long n_conv = h->hflag[1];
    const long long skipped = done - n_conv;  // steps whose stages were skipped
    if ((3 * skipped) & 1) h->cur ^= 1;
  }
  if (flag == 2) return set_error(TBEC_ERR_NONFINITE, "non-finite norm or energy in imaginary time");
  return tbec_materialize_scale(h);
}

int tbec_time_stages(tbec_handle *h, int nstages, double *ms_per_stage) {
  if (!h || !ms_per_stage || nstages < 1) return set_error(TBEC_ERR_ARG, "bad argument");
  int rc = tbec_materialize_scale(h);
  if (!rc) rc = prepare_physics(h, h->t_now);
  if (rc) return rc;
  // a handle with dt = (0, -dtau) times the imaginary-time stages (reductions and
  // finalisers included) from a fresh solver state with no convergence stop
  const bool imag = h->p.dt_re == 0.0 && h->p.dt_im < 0.0;
  if (imag) {
    h->tol = 0.0;
    rc = tbec_reset_state(h);
    if (rc) return rc;
    h->scale_pending = true;
  }
  hipEvent_t e0, e1;
  HIPCHK2(hipEventCreate(&e0));
  HIPCHK2(hipEventCreate(&e1));
  HIPCHK2(hipEventRecord(e0, h->stream));
  for (int i = 0; i < nstages; ++i) {
    rc = enqueue_stage(h, ST_RK1 + (i % 3), imag, false);
    if (rc) return rc;
  }
  HIPCHK2(hipEventRecord(e1, h->stream));
  HIPCHK2(hipEventSynchronize(e1));
  float ms = 0;
  HIPCHK2(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms_per_stage = ms / nstages;
  return imag ? tbec_materialize_scale(h) : 0;
}

}  // extern "C"
