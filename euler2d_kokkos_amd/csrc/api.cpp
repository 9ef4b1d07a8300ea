// api.cpp -- C ABI (include/euler2d_hip.h) over the HIP kernels.
//
// The handle plays the role of the reference's HydroRun<device_t>
// (src/HydroRun.h:44-134): it owns U, U2 (and, for the pass schedule, Q,
// Fluxes_x, Fluxes_y) in device memory as four SoA planes each, one HIP stream,
// and the device-resident loop control block.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include <cmath>

#include "config.hpp"
#include "euler2d_hip.h"
#include "kernels.hpp"
#include "params.hpp"
#include "status.hpp"
#include "trace.hpp"

using e2d::Grid;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

int hip_fail(hipError_t e, const char *what) {
  return fail(E2D_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr)                               \
  do {                                              \
    hipError_t e_ = (expr);                         \
    if (e_ != hipSuccess) return hip_fail(e_, #expr); \
  } while (0)

// Events that time one call on a stream, destroyed on every return path.
template <int N>
struct Events {
  hipEvent_t e[N] = {};
  hipError_t create() {
    for (auto &x : e) {
      const hipError_t rc = hipEventCreate(&x);
      if (rc != hipSuccess) return rc;
    }
    return hipSuccess;
  }
  ~Events() {
    for (auto x : e)
      if (x) (void)hipEventDestroy(x);
  }
};

// steps launched between two host checks of the loop state
constexpr int kChunk = 256;

}  // namespace

int e2d::report(int code, const std::string &msg) { return fail(code, msg); }

struct e2d_config {
  e2d::ConfigMap map;
};

struct e2d_run {
  e2d_params p{};
  e2d::KParams kp{};
  int precision = 0;
  int schedule = E2D_SCHEDULE_FUSED;
  int seg_rows = 16;
  int device = 0;
  hipStream_t stream = nullptr;
  Grid U{}, U2{}, Q{}, Fx{}, Fy{}, Sx{}, Sy{};
  e2d::StepCtl *ctl = nullptr;
  e2d::StepState *host_s = nullptr;  // pinned copy of ctl->s, written by the step kernels
  hipEvent_t ev0 = nullptr, ev1 = nullptr;  // the run loop's timing events
  int phase = 0;  // phase of the next run-loop launch in the control rings
  // captured runs of kLoopGraphSteps[g] run-loop launches (for seg_rows = loop_graph_tile[g])
  hipGraphExec_t loop_graph[2] = {nullptr, nullptr};
  int loop_graph_tile[2] = {0, 0};
  unsigned long long *bits = nullptr;
  unsigned int *count = nullptr;
  double *partials = nullptr;
  double *sums4 = nullptr;
  double timers[5] = {0, 0, 0, 0, 0};
  std::vector<void *> allocs;
};

namespace {

int alloc_dev(e2d_run *r, void **ptr, size_t bytes) {
  hipError_t e = hipMalloc(ptr, bytes);
  if (e != hipSuccess) {
    *ptr = nullptr;
    return fail(E2D_ERR_ALLOC, std::string("hipMalloc(") + std::to_string(bytes) +
                                   " bytes): " + hipGetErrorString(e));
  }
  r->allocs.push_back(*ptr);
  return E2D_OK;
}

// Rows start on 128-byte and planes on 512-byte boundaries.
int alloc_grid(e2d_run *r, Grid &g, int precision) {
  const int sb = e2d::grid_scalar_bytes(precision);
  const int row_align = 128 / sb, plane_align = 512 / sb;
  const int pitch = (r->p.isize + row_align - 1) / row_align * row_align;
  long long plane = (long long)pitch * r->p.jsize;
  plane = (plane + plane_align - 1) / plane_align * plane_align;
  void *ptr = nullptr;
  int rc = alloc_dev(r, &ptr, (size_t)plane * e2d::grid_planes(precision) * sb);
  if (rc) return rc;
  g.p = ptr;
  g.pitch = pitch;
  g.plane = plane;
  return E2D_OK;
}

Grid &state(e2d_run *r, int which) { return which == E2D_U ? r->U : r->U2; }

int check_run(e2d_run *r) {
  if (!r) return fail(E2D_ERR_ARG, "null e2d_run handle");
  hipError_t e = hipSetDevice(r->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  return E2D_OK;
}

int ensure_pass_arrays(e2d_run *r) {
  int rc = E2D_OK;
  if (!r->Q.p) {
    rc = alloc_grid(r, r->Q, E2D_PRECISION_FP64);
    if (!rc) rc = alloc_grid(r, r->Fx, E2D_PRECISION_FP64);
    if (!rc) rc = alloc_grid(r, r->Fy, E2D_PRECISION_FP64);
    if (!rc) {
      const size_t bytes = (size_t)r->Fx.plane * 4 * sizeof(double);
      HIP_TRY(hipMemsetAsync(r->Fx.p, 0, bytes, r->stream));
      HIP_TRY(hipMemsetAsync(r->Fy.p, 0, bytes, r->stream));
    }
  }
  if (!rc && r->schedule == E2D_SCHEDULE_DIRECTIONAL && !r->Sx.p) {
    // Slopes_x / Slopes_y of implementation version 1 (HydroRun.h:168-172)
    rc = alloc_grid(r, r->Sx, E2D_PRECISION_FP64);
    if (!rc) rc = alloc_grid(r, r->Sy, E2D_PRECISION_FP64);
    if (!rc) {
      const size_t bytes = (size_t)r->Sx.plane * 4 * sizeof(double);
      HIP_TRY(hipMemsetAsync(r->Sx.p, 0, bytes, r->stream));
      HIP_TRY(hipMemsetAsync(r->Sy.p, 0, bytes, r->stream));
    }
  }
  return rc;
}

// one godunov step with an explicit dt; in/out by the parity of nStep
int do_step(e2d_run *r, int nStep, double dt, bool timed) {
  Grid in = (nStep % 2 == 0) ? r->U : r->U2;
  Grid out = (nStep % 2 == 0) ? r->U2 : r->U;
  const double dtdx = dt / r->p.dx;
  const double dtdy = dt / r->p.dy;
  Events<5> evs;
  hipEvent_t *ev = evs.e;
  if (timed) HIP_TRY(evs.create());
  if (r->schedule != E2D_SCHEDULE_FUSED) {
    if (r->precision != E2D_PRECISION_FP64)
      return fail(E2D_ERR_UNSUPPORTED, "the pass schedules exist for FP64 only");
    int rc = ensure_pass_arrays(r);
    if (rc) return rc;
    // HydroRun::godunov_unsplit_impl, HydroRun.h:279-364
    if (timed) HIP_TRY(hipEventRecord(ev[0], r->stream));
    {
      e2d::Region reg("make_boundaries");
      HIP_TRY(e2d::launch_boundaries(r->kp, in, r->stream));
    }
    HIP_TRY(e2d::launch_copy(r->kp, in, out, r->stream));
    if (timed) HIP_TRY(hipEventRecord(ev[1], r->stream));
    {
      e2d::Region reg("compute_primitives");
      HIP_TRY(e2d::launch_prim(r->kp, in, r->Q, r->stream));
    }
    if (timed) HIP_TRY(hipEventRecord(ev[2], r->stream));
    if (r->schedule == E2D_SCHEDULE_PASSES) {
      e2d::Region impl("hydro_impl0");
      {
        e2d::Region reg("compute_fluxes");
        HIP_TRY(e2d::launch_fluxes(r->kp, r->Q, r->Fx, r->Fy, dtdx, dtdy, r->stream));
      }
      if (timed) HIP_TRY(hipEventRecord(ev[3], r->stream));
      e2d::Region reg("update_hydro");
      HIP_TRY(e2d::launch_update(r->kp, out, r->Fx, r->Fy, r->stream));
    } else if (r->schedule == E2D_SCHEDULE_DIRECTIONAL) {
      e2d::Region impl("hydro_impl1");
      HIP_TRY(e2d::launch_slopes(r->kp, r->Q, r->Sx, r->Sy, r->stream));
      HIP_TRY(e2d::launch_trace_fluxes(r->kp, 0, r->Q, r->Sx, r->Sy, r->Fx, dtdx, dtdy,
                                       r->stream));
      HIP_TRY(e2d::launch_update_dir(r->kp, 0, out, r->Fx, r->stream));
      HIP_TRY(e2d::launch_trace_fluxes(r->kp, 1, r->Q, r->Sx, r->Sy, r->Fy, dtdx, dtdy,
                                       r->stream));
      HIP_TRY(e2d::launch_update_dir(r->kp, 1, out, r->Fy, r->stream));
    } else {
      e2d::Region impl("hydro_impl2");
      HIP_TRY(e2d::launch_fluxes_atomic(r->kp, r->Q, out, dtdx, dtdy, r->stream));
    }
    if (timed) HIP_TRY(hipEventRecord(ev[4], r->stream));
  } else {
    if (timed) HIP_TRY(hipEventRecord(ev[0], r->stream));
    // implementation version 0 as one fused launch (boundaries to update)
    e2d::Region impl("hydro_impl0");
    HIP_TRY(e2d::launch_step_fused(r->kp, in, out, dt, nullptr, 0, r->seg_rows, r->stream));
    if (timed) HIP_TRY(hipEventRecord(ev[4], r->stream));
  }
  if (timed) {
    HIP_TRY(hipEventSynchronize(ev[4]));
    float ms = 0.f;
    if (r->schedule != E2D_SCHEDULE_FUSED) {
      HIP_TRY(hipEventElapsedTime(&ms, ev[0], ev[1]));
      r->timers[0] += ms * 1e-3;
      HIP_TRY(hipEventElapsedTime(&ms, ev[1], ev[2]));
      r->timers[2] += ms * 1e-3;
      if (r->schedule == E2D_SCHEDULE_PASSES) {  // flux / update timers of version 0 only
        HIP_TRY(hipEventElapsedTime(&ms, ev[2], ev[3]));
        r->timers[3] += ms * 1e-3;
        HIP_TRY(hipEventElapsedTime(&ms, ev[3], ev[4]));
        r->timers[4] += ms * 1e-3;
      }
      HIP_TRY(hipEventElapsedTime(&ms, ev[1], ev[4]));
      r->timers[1] += ms * 1e-3;
    } else {
      HIP_TRY(hipEventElapsedTime(&ms, ev[0], ev[4]));
      r->timers[1] += ms * 1e-3;
    }
  }
  return E2D_OK;
}

// Host-side loop state -> control block; dt of the pending step computed on
// the device from the current state.
// Phase 0 of the control rings: the state before the loop's first step and
// the CFL rate of its input; every other entry cleared.
// One launch: at nStep == 0 it also fills the ghost cells of U and U2
// (main.cpp:90-91).
int loop_begin(e2d_run *r, double t, int nStep) {
  r->phase = 0;
  const bool even = nStep % 2 == 0;
  HIP_TRY(e2d::launch_loop_begin(r->kp, even ? r->U : r->U2, even ? r->U2 : r->U, nStep == 0,
                                 r->ctl, t, nStep, r->partials, r->stream));
  return E2D_OK;
}

// Run-loop launches per captured graph: multiples of 6, so that a replay that
// starts at ring phase 0 on an even step ends there (phase ring 3, array
// parity 2), the long one below the launches between two host checks of an
// open-ended loop.  Between two graph launches the GPU idles about 8.6 us,
// between two single launches it does not: the long graph carries long loops
// with few such gaps, and budgets below the short one run as single launches
// (a 20-step budget ran 1.5% faster than as three 6-launch graphs and two
// single launches).
constexpr int kLoopGraphSteps[2] = {240, 30};
static_assert(kLoopGraphSteps[0] <= kChunk, "the long loop graph must fit in one chunk");

int loop_graph(e2d_run *r, int g) {
  if (r->loop_graph[g] && r->loop_graph_tile[g] == r->seg_rows) return E2D_OK;
  if (r->loop_graph[g]) (void)hipGraphExecDestroy(r->loop_graph[g]);
  r->loop_graph[g] = nullptr;
  hipGraph_t graph = nullptr;
  HIP_TRY(hipStreamBeginCapture(r->stream, hipStreamCaptureModeThreadLocal));
  for (int k = 0; k < kLoopGraphSteps[g]; ++k) {
    const Grid &in = (k % 2 == 0) ? r->U : r->U2;
    const Grid &out = (k % 2 == 0) ? r->U2 : r->U;
    hipError_t e = e2d::launch_step_fused(r->kp, in, out, 0.0, r->ctl, k % 3, r->seg_rows,
                                          r->stream);
    if (e != hipSuccess) {
      (void)hipStreamEndCapture(r->stream, &graph);
      if (graph) (void)hipGraphDestroy(graph);
      return hip_fail(e, "launch_step_fused (capture)");
    }
  }
  HIP_TRY(hipStreamEndCapture(r->stream, &graph));
  hipError_t e = hipGraphInstantiate(&r->loop_graph[g], graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) {
    r->loop_graph[g] = nullptr;
    return hip_fail(e, "hipGraphInstantiate");
  }
  r->loop_graph_tile[g] = r->seg_rows;
  return E2D_OK;
}

// Captures, ahead of a timed region, the graphs a loop of `count` launches
// will replay (capture and instantiation are host work of about a millisecond
// per 30 launches).
int loop_graphs_for(e2d_run *r, int count) {
  for (int g = 0; g < 2; ++g)
    if (count >= kLoopGraphSteps[g]) {
      int rc = loop_graph(r, g);
      if (rc) return rc;
    }
  return E2D_OK;
}

// Queues `count` run-loop launches from step nStep0 on: captured runs of 240
// or 30 launches where the ring phase is 0 and the step even, single launches
// elsewhere.
int loop_steps(e2d_run *r, int nStep0, int count) {
  int k = 0;
  while (k < count) {
    const int n = nStep0 + k;
    if (r->phase == 0 && n % 2 == 0 && count - k >= kLoopGraphSteps[1]) {
      const int g = count - k >= kLoopGraphSteps[0] ? 0 : 1;
      int rc = loop_graph(r, g);
      if (rc) return rc;
      HIP_TRY(hipGraphLaunch(r->loop_graph[g], r->stream));
      k += kLoopGraphSteps[g];
      continue;
    }
    Grid in = (n % 2 == 0) ? r->U : r->U2;
    Grid out = (n % 2 == 0) ? r->U2 : r->U;
    HIP_TRY(e2d::launch_step_fused(r->kp, in, out, 0.0, r->ctl, r->phase, r->seg_rows,
                                   r->stream));
    r->phase = r->phase == 2 ? 0 : r->phase + 1;
    ++k;
  }
  return E2D_OK;
}

// Loop state after the launched steps, with the while condition of main.cpp:100.
struct LoopView {
  double t = 0.0;
  int nstep = 0;
  bool active = false;
};

LoopView loop_view(const e2d_params &p, const e2d::StepState &s) {
  LoopView v;
  v.t = s.t;
  v.nstep = s.nstep;
  v.active = s.t < p.tEnd && s.nstep < p.nStepmax;
  return v;
}

// The kernels also store each ring entry in pinned host memory, so reading
// the loop state needs only the synchronisation.
int loop_read(e2d_run *r, LoopView &v) {
  HIP_TRY(hipStreamSynchronize(r->stream));
  const volatile e2d::StepState &hs = r->host_s[r->phase];
  e2d::StepState s{};
  s.t = hs.t;
  s.nstep = hs.nstep;
  v = loop_view(r->p, s);
  return E2D_OK;
}

int run_loop_impl(e2d_run *r, int max_steps, double *t, int *nStep, double *elapsed_ms,
                  int *executed) {
  if (!t || !nStep) return fail(E2D_ERR_ARG, "null t / nStep");
  e2d::Region reg("main_loop");
  if (r->schedule != E2D_SCHEDULE_FUSED) {
    // reference structure: host dt per step, exactly main.cpp:100-145
    const int n0 = *nStep;
    if (*nStep == 0) {
      HIP_TRY(e2d::launch_boundaries(r->kp, r->U, r->stream));
      HIP_TRY(e2d::launch_boundaries(r->kp, r->U2, r->stream));
    }
    Events<2> ev;
    HIP_TRY(ev.create());
    HIP_TRY(hipEventRecord(ev.e[0], r->stream));
    int done = 0;
    while (*t < r->p.tEnd && *nStep < r->p.nStepmax && (max_steps <= 0 || done < max_steps)) {
      double dt;
      int rc = e2d_compute_dt(r, *nStep % 2, &dt);
      if (rc) return rc;
      if (*t + dt > r->p.tEnd) dt = r->p.tEnd - *t;
      rc = do_step(r, *nStep, dt, false);
      if (rc) return rc;
      *nStep += 1;
      *t += dt;
      done++;
    }
    HIP_TRY(hipEventRecord(ev.e[1], r->stream));
    HIP_TRY(hipEventSynchronize(ev.e[1]));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ev.e[0], ev.e[1]));
    r->timers[1] += ms * 1e-3;
    if (elapsed_ms) *elapsed_ms = ms;
    if (executed) *executed = *nStep - n0;
    return E2D_OK;
  }

  const int n0 = *nStep;
  // graphs first: their capture is host work that would otherwise fall inside
  // the timed region
  int rc = loop_graphs_for(r, max_steps > 0 ? std::min(max_steps, r->p.nStepmax - n0) : kChunk);
  if (rc) return rc;
  rc = loop_begin(r, *t, *nStep);
  if (rc) return rc;
  hipEvent_t e0 = r->ev0, e1 = r->ev1;
  HIP_TRY(hipEventRecord(e0, r->stream));
  LoopView c;
  int launched = 0;
  for (;;) {
    // with an explicit step budget every launch is queued at once (steps past
    // the end of the loop return immediately); otherwise check every kChunk
    int todo = max_steps > 0 ? max_steps - launched : kChunk;
    const long long left = (long long)r->p.nStepmax - (n0 + launched);
    if (left < todo) todo = (int)std::max(0LL, left);
    if (todo <= 0) break;
    rc = loop_steps(r, n0 + launched, todo);
    if (rc) return rc;
    launched += todo;
    // an explicit budget is queued whole: the end event follows the last
    // launch directly, with no host round trip in between
    if (max_steps > 0 && launched >= max_steps) break;
    rc = loop_read(r, c);
    if (rc) return rc;
    if (!c.active) break;
  }
  HIP_TRY(hipEventRecord(e1, r->stream));
  rc = loop_read(r, c);
  if (rc) return rc;
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
  r->timers[1] += ms * 1e-3;
  *t = c.t;
  *nStep = c.nstep;
  if (elapsed_ms) *elapsed_ms = ms;
  if (executed) *executed = c.nstep - n0;
  return E2D_OK;
}

// Device state -> host FP64 planes [v][j][i] (isize x jsize, ghosts included);
// the split representation is combined as (double)hi + (double)lo.
int download_planes(e2d_run *r, const Grid &g, std::vector<double> &out) {
  const int isz = r->p.isize, jsz = r->p.jsize;
  const size_t cells = (size_t)isz * jsz;
  out.assign(4 * cells, 0.0);
  if (r->precision == E2D_PRECISION_FP64) {
    const double *src = static_cast<const double *>(g.p);
    for (int v = 0; v < 4; v++)
      HIP_TRY(hipMemcpy2DAsync(out.data() + v * cells, (size_t)isz * sizeof(double),
                               src + v * g.plane, (size_t)g.pitch * sizeof(double),
                               (size_t)isz * sizeof(double), (size_t)jsz, hipMemcpyDeviceToHost,
                               r->stream));
    HIP_TRY(hipStreamSynchronize(r->stream));
    return E2D_OK;
  }
  // split: (hi, lo) float pairs; FP32: floats
  const int per = r->precision == E2D_PRECISION_FP32_SPLIT ? 2 : 1;
  const size_t eb = per * sizeof(float);
  std::vector<float> tmp((size_t)per * 4 * cells);
  const char *src = static_cast<const char *>(g.p);
  for (int v = 0; v < 4; v++)
    HIP_TRY(hipMemcpy2DAsync(tmp.data() + (size_t)per * v * cells, (size_t)isz * eb,
                             src + (size_t)v * g.plane * eb, (size_t)g.pitch * eb, (size_t)isz * eb,
                             (size_t)jsz, hipMemcpyDeviceToHost, r->stream));
  HIP_TRY(hipStreamSynchronize(r->stream));
  for (size_t k = 0; k < 4 * cells; k++)
    out[k] = per == 2 ? (double)tmp[2 * k] + (double)tmp[2 * k + 1] : (double)tmp[k];
  return E2D_OK;
}

// Host FP64 planes -> device state (split: hi = (float)x, lo = (float)(x - hi)).
int upload_planes(e2d_run *r, const Grid &g, const std::vector<double> &in) {
  const int isz = r->p.isize, jsz = r->p.jsize;
  const size_t cells = (size_t)isz * jsz;
  if (r->precision == E2D_PRECISION_FP64) {
    double *dst = static_cast<double *>(g.p);
    for (int v = 0; v < 4; v++)
      HIP_TRY(hipMemcpy2DAsync(dst + v * g.plane, (size_t)g.pitch * sizeof(double),
                               in.data() + v * cells, (size_t)isz * sizeof(double),
                               (size_t)isz * sizeof(double), (size_t)jsz, hipMemcpyHostToDevice,
                               r->stream));
    HIP_TRY(hipStreamSynchronize(r->stream));
    return E2D_OK;
  }
  const int per = r->precision == E2D_PRECISION_FP32_SPLIT ? 2 : 1;
  const size_t eb = per * sizeof(float);
  std::vector<float> tmp((size_t)per * 4 * cells);
  for (size_t k = 0; k < 4 * cells; k++) {
    const double x = in[k];
    const float hi = (float)x;
    if (per == 2) {
      tmp[2 * k] = hi;
      tmp[2 * k + 1] = (float)(x - (double)hi);
    } else {
      tmp[k] = hi;
    }
  }
  char *dst = static_cast<char *>(g.p);
  for (int v = 0; v < 4; v++)
    HIP_TRY(hipMemcpy2DAsync(dst + (size_t)v * g.plane * eb, (size_t)g.pitch * eb,
                             tmp.data() + (size_t)per * v * cells, (size_t)isz * eb,
                             (size_t)isz * eb, (size_t)jsz, hipMemcpyHostToDevice, r->stream));
  HIP_TRY(hipStreamSynchronize(r->stream));
  return E2D_OK;
}

void planes_to_layout(const e2d_run *r, const std::vector<double> &pl, double *host, int layout) {
  const int isz = r->p.isize, jsz = r->p.jsize, nx = r->p.nx, ny = r->p.ny;
  const size_t cells = (size_t)isz * jsz;
  for (int v = 0; v < 4; v++)
    for (int j = 0; j < jsz; j++)
      for (int i = 0; i < isz; i++) {
        const double x = pl[v * cells + (size_t)j * isz + i];
        if (layout == E2D_LAYOUT_RIGHT)
          host[((size_t)i * jsz + j) * 4 + v] = x;
        else if (layout == E2D_LAYOUT_LEFT)
          host[v * cells + (size_t)j * isz + i] = x;
        else if (i >= 2 && i < nx + 2 && j >= 2 && j < ny + 2)
          host[(size_t)v * nx * ny + (size_t)(j - 2) * nx + (i - 2)] = x;
      }
}

void layout_to_planes(const e2d_run *r, const double *host, int layout, std::vector<double> &pl) {
  const int isz = r->p.isize, jsz = r->p.jsize, nx = r->p.nx, ny = r->p.ny;
  const size_t cells = (size_t)isz * jsz;
  for (int v = 0; v < 4; v++)
    for (int j = 0; j < jsz; j++)
      for (int i = 0; i < isz; i++) {
        double &x = pl[v * cells + (size_t)j * isz + i];
        if (layout == E2D_LAYOUT_RIGHT)
          x = host[((size_t)i * jsz + j) * 4 + v];
        else if (layout == E2D_LAYOUT_LEFT)
          x = host[v * cells + (size_t)j * isz + i];
        else if (i >= 2 && i < nx + 2 && j >= 2 && j < ny + 2)
          x = host[(size_t)v * nx * ny + (size_t)(j - 2) * nx + (i - 2)];
      }
}

template <class F>
int guarded(F &&f) {
  try {
    return f();
  } catch (const std::bad_alloc &) {
    return fail(E2D_ERR_ALLOC, "host allocation failed");
  } catch (...) {
    return fail(E2D_ERR_ARG, "unexpected C++ exception");
  }
}

}  // namespace

extern "C" {

const char *e2d_last_error(void) { return g_last_error.c_str(); }

int e2d_version(void) { return 1; }

int e2d_params_size(void) { return (int)sizeof(e2d_params); }

int e2d_device_info(char *name, int name_len, int *num_cu) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return fail(E2D_ERR_HIP, "no HIP device visible");
  hipDeviceProp_t prop;
  e = hipGetDeviceProperties(&prop, 0);
  if (e != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
  if (name && name_len > 0) {
    // the marketing name can be empty (a device seen without its product
    // table); the architecture name is always there
    if (prop.name[0])
      std::snprintf(name, (size_t)name_len, "%s (%s)", prop.name, prop.gcnArchName);
    else
      std::snprintf(name, (size_t)name_len, "%s", prop.gcnArchName);
  }
  if (num_cu) *num_cu = prop.multiProcessorCount;
  return E2D_OK;
}

// ------------------------------------------------------------ ConfigMap
int e2d_config_open(const char *path, e2d_config **out, int *parse_error) {
  return guarded([&]() -> int {
    if (!path || !out) return fail(E2D_ERR_ARG, "null path / out");
    auto *c = new e2d_config{e2d::ConfigMap(path)};
    *out = c;
    if (parse_error) *parse_error = c->map.ParseError();
    if (c->map.ParseError() == -1) return fail(E2D_ERR_IO, std::string("cannot open ") + path);
    return E2D_OK;
  });
}

int e2d_config_close(e2d_config *cfg) {
  delete cfg;
  return E2D_OK;
}

int e2d_config_get_string(const e2d_config *cfg, const char *section, const char *name,
                          const char *default_value, char *out, int out_len) {
  if (!cfg || !section || !name || !default_value) return fail(E2D_ERR_ARG, "null argument");
  return guarded([&]() -> int {
    const std::string v = cfg->map.getString(section, name, default_value);
    if (out && out_len > 0) {
      std::strncpy(out, v.c_str(), (size_t)out_len);
      out[out_len - 1] = '\0';
    }
    return (int)v.size();
  });
}

int e2d_config_get_integer(const e2d_config *cfg, const char *section, const char *name,
                           long default_value, long *out) {
  if (!cfg || !section || !name || !out) return fail(E2D_ERR_ARG, "null argument");
  return guarded([&]() -> int {
    *out = cfg->map.getInteger(section, name, default_value);
    return E2D_OK;
  });
}

int e2d_config_get_float(const e2d_config *cfg, const char *section, const char *name,
                         float default_value, float *out) {
  if (!cfg || !section || !name || !out) return fail(E2D_ERR_ARG, "null argument");
  return guarded([&]() -> int {
    *out = cfg->map.getFloat(section, name, default_value);
    return E2D_OK;
  });
}

int e2d_config_get_bool(const e2d_config *cfg, const char *section, const char *name,
                        int default_value, int *out) {
  if (!cfg || !section || !name || !out) return fail(E2D_ERR_ARG, "null argument");
  return guarded([&]() -> int {
    *out = cfg->map.getBool(section, name, default_value != 0) ? 1 : 0;
    return E2D_OK;
  });
}

// ---------------------------------------------------------- HydroParams
int e2d_params_setup(const e2d_config *cfg, e2d_params *out) {
  if (!cfg || !out) return fail(E2D_ERR_ARG, "null argument");
  return guarded([&]() -> int {
    e2d::params_setup(cfg->map, *out);
    return E2D_OK;
  });
}

int e2d_params_from_ini(const char *path, e2d_params *out, int *parse_error) {
  if (!path || !out) return fail(E2D_ERR_ARG, "null argument");
  return guarded([&]() -> int {
    e2d::ConfigMap map(path);
    if (parse_error) *parse_error = map.ParseError();
    e2d::params_setup(map, *out);  // defaults when the file is missing, like the reference
    if (map.ParseError() == -1) return fail(E2D_ERR_IO, std::string("cannot open ") + path);
    return E2D_OK;
  });
}

int e2d_params_init(e2d_params *p) {
  if (!p) return fail(E2D_ERR_ARG, "null params");
  e2d::params_init(*p);
  return E2D_OK;
}

int e2d_params_print(const e2d_params *p) {
  if (!p) return fail(E2D_ERR_ARG, "null params");
  e2d::params_print(*p);
  return E2D_OK;
}

int e2d_params_validate(const e2d_params *p) {
  if (!p) return fail(E2D_ERR_ARG, "null params");
  const char *why = e2d::params_check(*p);
  if (why) return fail(E2D_ERR_UNSUPPORTED, why);
  return E2D_OK;
}

// --------------------------------------------------------------- HydroRun
int e2d_create(const e2d_params *p, int precision, e2d_run **out) {
  if (!p || !out) return fail(E2D_ERR_ARG, "null argument");
  *out = nullptr;
  int rc = e2d_params_validate(p);
  if (rc) return rc;
  if (precision != E2D_PRECISION_FP64 && precision != E2D_PRECISION_FP32_SPLIT &&
      precision != E2D_PRECISION_FP32)
    return fail(E2D_ERR_ARG, "precision must be one of E2D_PRECISION_FP64 / _FP32_SPLIT / _FP32");
  return guarded([&]() -> int {
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) return fail(E2D_ERR_HIP, "no HIP device visible");
    auto *r = new e2d_run();
    // every failure past this point releases what the handle holds
    const int rc1 = [&]() -> int {
    r->p = *p;
    r->kp = e2d::make_kparams(*p, precision);
    r->precision = precision;
    // implementationVersion selects the schedule, as in godunov_unsplit_impl
    // (HydroRun.h:312-357); version 0 runs fused, with the same bits
    if (precision == E2D_PRECISION_FP64 && p->implementationVersion == 1)
      r->schedule = E2D_SCHEDULE_DIRECTIONAL;
    else if (precision == E2D_PRECISION_FP64 && p->implementationVersion == 2)
      r->schedule = E2D_SCHEDULE_ATOMIC;
    HIP_TRY(hipGetDevice(&r->device));
    int ncu = 0;
    HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, r->device));
    r->kp.simds = 4 * ncu;
    HIP_TRY(hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking));
    void *ptr = nullptr;
    int rc2 = alloc_grid(r, r->U, precision);
    if (!rc2) rc2 = alloc_grid(r, r->U2, precision);
    if (!rc2) rc2 = alloc_dev(r, &ptr, sizeof(e2d::StepCtl));
    if (!rc2) r->ctl = static_cast<e2d::StepCtl *>(ptr);
    if (!rc2) rc2 = alloc_dev(r, &ptr, 64);
    if (!rc2) r->bits = static_cast<unsigned long long *>(ptr);
    if (!rc2) rc2 = alloc_dev(r, &ptr, 64);
    if (!rc2) r->count = static_cast<unsigned int *>(ptr);
    if (!rc2) rc2 = alloc_dev(r, &ptr, (size_t)e2d::sums_partials_needed(r->kp) * sizeof(double));
    if (!rc2) r->partials = static_cast<double *>(ptr);
    if (!rc2) rc2 = alloc_dev(r, &ptr, 4 * sizeof(double));
    if (!rc2) r->sums4 = static_cast<double *>(ptr);
    if (rc2) return rc2;
    HIP_TRY(hipEventCreate(&r->ev0));
    HIP_TRY(hipEventCreate(&r->ev1));
    HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&r->host_s), 3 * sizeof(e2d::StepState),
                          hipHostMallocDefault));
    std::memset(r->host_s, 0, 3 * sizeof(e2d::StepState));
    // 32-bit row offsets in the step kernels need planes of at most 4 GiB;
    // E2D_WIDE_ADDRESSING=1 selects the 64-bit form for any size
    const char *wide = std::getenv("E2D_WIDE_ADDRESSING");
    r->kp.wide_addr = (wide && *wide && std::strcmp(wide, "0") != 0) ||
                      (unsigned long long)r->U.plane * e2d::grid_scalar_bytes(precision) >
                          (1ull << 32);
    // defined contents everywhere (ghosts included) before the first init
    const size_t bytes =
      (size_t)r->U.plane * e2d::grid_planes(precision) * e2d::grid_scalar_bytes(precision);
    HIP_TRY(hipMemsetAsync(r->U.p, 0, bytes, r->stream));
    HIP_TRY(hipMemsetAsync(r->U2.p, 0, bytes, r->stream));
    HIP_TRY(hipMemsetAsync(r->ctl, 0, sizeof(e2d::StepCtl), r->stream));
    HIP_TRY(hipMemcpyAsync(&r->ctl->host_s, &r->host_s, sizeof(r->host_s), hipMemcpyHostToDevice,
                           r->stream));
    HIP_TRY(hipStreamSynchronize(r->stream));
    return E2D_OK;
    }();
    if (rc1) {
      const std::string msg = g_last_error;
      e2d_destroy(r);
      return fail(rc1, msg);
    }
    *out = r;
    return E2D_OK;
  });
}

int e2d_destroy(e2d_run *run) {
  if (!run) return fail(E2D_ERR_ARG, "null e2d_run handle");
  (void)hipSetDevice(run->device);
  if (run->stream) (void)hipStreamSynchronize(run->stream);
  for (hipGraphExec_t g : run->loop_graph)
    if (g) (void)hipGraphExecDestroy(g);
  if (run->ev0) (void)hipEventDestroy(run->ev0);
  if (run->ev1) (void)hipEventDestroy(run->ev1);
  if (run->host_s) (void)hipHostFree(run->host_s);
  for (void *ptr : run->allocs) (void)hipFree(ptr);
  if (run->stream) (void)hipStreamDestroy(run->stream);
  delete run;
  return E2D_OK;
}

int e2d_init_problem(e2d_run *r) {
  int rc = check_run(r);
  if (rc) return rc;
  const bool sedov = r->p.problemType == 1 && r->p.blast_total_energy_inside > 0;
  if (sedov) HIP_TRY(hipMemsetAsync(r->count, 0, sizeof(unsigned int), r->stream));
  HIP_TRY(e2d::launch_init(r->kp, r->p, r->U, sedov ? r->count : nullptr, r->stream));
  if (sedov) {
    // volume inside the disk, accumulated in a fixed sequential order
    unsigned int n = 0;
    HIP_TRY(hipMemcpyAsync(&n, r->count, sizeof(n), hipMemcpyDeviceToHost, r->stream));
    HIP_TRY(hipStreamSynchronize(r->stream));
    double volume = 0.0;
    for (unsigned int k = 0; k < n; ++k) volume += r->p.dx * r->p.dy;
    HIP_TRY(e2d::launch_blast_energy(r->kp, r->p, r->U, r->p.blast_total_energy_inside / volume,
                                     r->stream));
  }
  HIP_TRY(e2d::launch_copy(r->kp, r->U, r->U2, r->stream));
  HIP_TRY(hipStreamSynchronize(r->stream));
  return E2D_OK;
}

int e2d_copy_array(e2d_run *r, int src, int dst) {
  int rc = check_run(r);
  if (rc) return rc;
  if ((src != E2D_U && src != E2D_U2) || (dst != E2D_U && dst != E2D_U2))
    return fail(E2D_ERR_ARG, "src/dst must be E2D_U or E2D_U2");
  if (src == dst) return E2D_OK;
  HIP_TRY(e2d::launch_copy(r->kp, state(r, src), state(r, dst), r->stream));
  return E2D_OK;
}

int e2d_compute_dt(e2d_run *r, int useU, double *dt) {
  int rc = check_run(r);
  if (rc) return rc;
  if (!dt) return fail(E2D_ERR_ARG, "null dt");
  e2d::Region reg("compute_dt");
  HIP_TRY(hipMemsetAsync(r->bits, 0, sizeof(unsigned long long), r->stream));
  HIP_TRY(e2d::launch_dt_reduce(r->kp, useU == 0 ? r->U : r->U2, r->bits, r->stream));
  unsigned long long bits = 0;
  HIP_TRY(hipMemcpyAsync(&bits, r->bits, sizeof(bits), hipMemcpyDeviceToHost, r->stream));
  HIP_TRY(hipStreamSynchronize(r->stream));
  double invdt;
  std::memcpy(&invdt, &bits, sizeof(invdt));
  *dt = r->p.cfl / invdt;  // HydroRun.h:246
  return E2D_OK;
}

int e2d_make_boundaries(e2d_run *r, int which) {
  int rc = check_run(r);
  if (rc) return rc;
  if (which != E2D_U && which != E2D_U2) return fail(E2D_ERR_ARG, "which must be E2D_U or E2D_U2");
  e2d::Region reg("make_boundaries");
  HIP_TRY(e2d::launch_boundaries(r->kp, state(r, which), r->stream));
  return E2D_OK;
}

int e2d_godunov_unsplit(e2d_run *r, int nStep, double dt) {
  int rc = check_run(r);
  if (rc) return rc;
  if (nStep < 0) return fail(E2D_ERR_ARG, "nStep must be >= 0");
  return do_step(r, nStep, dt, true);
}

int e2d_run_loop(e2d_run *r, int max_steps, double *t, int *nStep) {
  int rc = check_run(r);
  if (rc) return rc;
  return run_loop_impl(r, max_steps, t, nStep, nullptr, nullptr);
}

int e2d_time_loop(e2d_run *r, int nsteps, double *elapsed_ms, int *executed, double *t,
                  int *nStep) {
  int rc = check_run(r);
  if (rc) return rc;
  if (nsteps <= 0) return fail(E2D_ERR_ARG, "nsteps must be > 0");
  return run_loop_impl(r, nsteps, t, nStep, elapsed_ms, executed);
}

int e2d_copy_state(e2d_run *r, int which, double *host, int layout) {
  int rc = check_run(r);
  if (rc) return rc;
  if (!host || (which != E2D_U && which != E2D_U2)) return fail(E2D_ERR_ARG, "bad argument");
  if (layout != E2D_LAYOUT_INTERIOR && layout != E2D_LAYOUT_LEFT && layout != E2D_LAYOUT_RIGHT)
    return fail(E2D_ERR_ARG, "unknown layout");
  const Grid &g = state(r, which);
  return guarded([&]() -> int {
    std::vector<double> planes;
    rc = download_planes(r, g, planes);
    if (rc) return rc;
    planes_to_layout(r, planes, host, layout);
    return E2D_OK;
  });
}

int e2d_upload_state(e2d_run *r, int which, const double *host, int layout) {
  int rc = check_run(r);
  if (rc) return rc;
  if (!host || (which != E2D_U && which != E2D_U2)) return fail(E2D_ERR_ARG, "bad argument");
  if (layout != E2D_LAYOUT_INTERIOR && layout != E2D_LAYOUT_LEFT && layout != E2D_LAYOUT_RIGHT)
    return fail(E2D_ERR_ARG, "unknown layout");
  const Grid &g = state(r, which);
  return guarded([&]() -> int {
    std::vector<double> planes;
    if (layout == E2D_LAYOUT_INTERIOR) {
      // keep the ghost cells the device already holds
      rc = download_planes(r, g, planes);
      if (rc) return rc;
    } else {
      planes.assign((size_t)4 * r->p.isize * r->p.jsize, 0.0);
    }
    layout_to_planes(r, host, layout, planes);
    return upload_planes(r, g, planes);
  });
}

int e2d_conserved_sums(e2d_run *r, int which, double sums[4]) {
  int rc = check_run(r);
  if (rc) return rc;
  if (!sums || (which != E2D_U && which != E2D_U2)) return fail(E2D_ERR_ARG, "bad argument");
  HIP_TRY(e2d::launch_sums(r->kp, state(r, which), r->partials, r->sums4, r->stream));
  HIP_TRY(hipMemcpyAsync(sums, r->sums4, 4 * sizeof(double), hipMemcpyDeviceToHost, r->stream));
  HIP_TRY(hipStreamSynchronize(r->stream));
  return E2D_OK;
}

int e2d_kinetic_cells(e2d_run *r, int which, long long *count) {
  int rc = check_run(r);
  if (rc) return rc;
  if (!count || (which != E2D_U && which != E2D_U2)) return fail(E2D_ERR_ARG, "bad argument");
  unsigned long long *dev = r->bits + 1;  // bits[0] is compute_dt's
  HIP_TRY(hipMemsetAsync(dev, 0, sizeof(unsigned long long), r->stream));
  HIP_TRY(e2d::launch_kinetic_count(r->kp, state(r, which), dev, r->stream));
  unsigned long long n = 0;
  HIP_TRY(hipMemcpyAsync(&n, dev, sizeof(n), hipMemcpyDeviceToHost, r->stream));
  HIP_TRY(hipStreamSynchronize(r->stream));
  *count = (long long)n;
  return E2D_OK;
}

int e2d_save_vtk(e2d_run *r, int which, const char *output_dir, const char *output_prefix,
                 int step) {
  int rc = check_run(r);
  if (rc) return rc;
  if (which != E2D_U && which != E2D_U2) return fail(E2D_ERR_ARG, "which must be E2D_U or E2D_U2");
  return guarded([&]() -> int {
    std::vector<double> host((size_t)4 * r->p.nx * r->p.ny);
    int rc2 = e2d_copy_state(r, which, host.data(), E2D_LAYOUT_INTERIOR);
    if (rc2) return rc2;
    return e2d_write_vti(&r->p, host.data(), r->precision == E2D_PRECISION_FP32 ? 1 : 0,
                         output_dir, output_prefix, step);
  });
}

int e2d_radial_profile(e2d_run *r, int which, int nbins, double *distances, double *density,
                       int64_t *counts) {
  int rc = check_run(r);
  if (rc) return rc;
  if (which != E2D_U && which != E2D_U2) return fail(E2D_ERR_ARG, "which must be E2D_U or E2D_U2");
  if (nbins < 1 || !distances || !density) return fail(E2D_ERR_ARG, "bad argument");
  const e2d_params &p = r->p;
  // ComputeRadialProfileFunctor::apply, ComputeRadialProfileFunctor.h:83-128
  e2d::RadialGeom geo{};
  geo.xmin = p.xmin;
  geo.ymin = p.ymin;
  geo.cx = (p.xmin + p.xmax) / 2;
  geo.cy = (p.ymin + p.ymax) / 2;
  const double Dx = (p.xmax - p.xmin) / 2;
  const double Dy = (p.ymax - p.ymin) / 2;
  geo.max_radial_distance = std::sqrt(Dx * Dx + Dy * Dy);
  geo.nbins = nbins;
  return guarded([&]() -> int {
    void *dev = nullptr;
    const size_t bytes = (size_t)nbins * (sizeof(unsigned long long) + sizeof(double));
    HIP_TRY(hipMalloc(&dev, bytes));
    auto *cnt = static_cast<unsigned long long *>(dev);
    auto *sum = reinterpret_cast<double *>(cnt + nbins);
    std::vector<unsigned long long> hc(nbins);
    std::vector<double> hs(nbins);
    hipError_t e = e2d::launch_radial_profile(r->kp, state(r, which), geo, cnt, sum, r->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(hc.data(), cnt, sizeof(unsigned long long) * nbins,
                         hipMemcpyDeviceToHost, r->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(hs.data(), sum, sizeof(double) * nbins, hipMemcpyDeviceToHost,
                         r->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
    (void)hipFree(dev);
    if (e != hipSuccess) return hip_fail(e, "radial profile");
    const double dr = geo.max_radial_distance / nbins;
    for (int i = 0; i < nbins; ++i) {
      distances[i] = (i + 0.5) * dr;
      density[i] = hs[i] / (int)hc[i];  // empty bin: 0/0, as the reference
      if (counts) counts[i] = (int64_t)hc[i];
    }
    return E2D_OK;
  });
}

int e2d_timers(e2d_run *r, double out[5]) {
  if (!r || !out) return fail(E2D_ERR_ARG, "null argument");
  for (int k = 0; k < 5; k++) out[k] = r->timers[k];
  return E2D_OK;
}

int e2d_set_schedule(e2d_run *r, int schedule) {
  if (!r) return fail(E2D_ERR_ARG, "null e2d_run handle");
  if (schedule < E2D_SCHEDULE_FUSED || schedule > E2D_SCHEDULE_ATOMIC)
    return fail(E2D_ERR_ARG, "unknown schedule");
  r->schedule = schedule;
  return E2D_OK;
}

int e2d_get_schedule(e2d_run *r, int *schedule) {
  if (!r || !schedule) return fail(E2D_ERR_ARG, "null argument");
  *schedule = r->schedule;
  return E2D_OK;
}

int e2d_set_tile(e2d_run *r, int seg_rows) {
  if (!r) return fail(E2D_ERR_ARG, "null e2d_run handle");
  if (seg_rows < 1 || seg_rows > 4096) return fail(E2D_ERR_ARG, "seg_rows must be in [1, 4096]");
  r->seg_rows = seg_rows;
  return E2D_OK;
}

int e2d_get_tile(e2d_run *r, int *seg_rows) {
  if (!r || !seg_rows) return fail(E2D_ERR_ARG, "null argument");
  *seg_rows = r->seg_rows;
  return E2D_OK;
}

int e2d_autotune(e2d_run *r, const int *candidates, int ncand, int steps, int *best_seg_rows,
                 double *best_ms_per_step) {
  int rc = check_run(r);
  if (rc) return rc;
  if (steps <= 0) return fail(E2D_ERR_ARG, "bad argument");
  std::vector<int> auto_cand;
  if (!candidates || ncand <= 0) {
    // A wavefront alone keeps its SIMD's VALU about as busy as two sharing
    // it, so the step time follows the rows marched per SIMD: the smallest
    // tiles at which the step's wavefronts fit k per SIMD (k = 1..4) are the
    // natural candidates, next to a fixed set.
    auto_cand = {8, 12, 14, 16, 20, 24, 32, 48, 64};
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, r->device) == hipSuccess) {
      const int simds = 4 * prop.multiProcessorCount;
      for (int k = 1; k <= 4; ++k) {
        int tile = 1;
        while (tile < 4096 && e2d::fused_map(r->kp, tile).nwaves > k * simds) ++tile;
        if (std::find(auto_cand.begin(), auto_cand.end(), tile) == auto_cand.end())
          auto_cand.push_back(tile);
      }
    }
    candidates = auto_cand.data();
    ncand = (int)auto_cand.size();
  }
  for (int c = 0; c < ncand; ++c)
    if (candidates[c] < 1 || candidates[c] > 4096) return fail(E2D_ERR_ARG, "bad candidate");
  // snapshot of both arrays; the tuning runs advance a throw-away copy of the
  // loop.  Every exit below releases the snapshots and restores the schedule.
  Grid s1{}, s2{};
  rc = alloc_grid(r, s1, r->precision);
  if (!rc) rc = alloc_grid(r, s2, r->precision);
  hipError_t e = hipSuccess;
  if (!rc) {
    e = e2d::launch_copy(r->kp, r->U, s1, r->stream);
    if (e == hipSuccess) e = e2d::launch_copy(r->kp, r->U2, s2, r->stream);
    if (e != hipSuccess) rc = hip_fail(e, "launch_copy(snapshot)");
  }
  const int schedule = r->schedule;
  r->schedule = E2D_SCHEDULE_FUSED;
  const double timers1 = r->timers[1];
  const int tile0 = r->seg_rows;
  int best = tile0;
  double best_ms = 1e300;
  // two interleaved passes over the candidates, the faster of each
  // candidate's two timings counts: a clock ramp or a transient during one
  // short timing does not decide the choice
  std::vector<double> cand_ms((size_t)ncand, 1e300);
  for (int pass = 0; pass < 2 && !rc; ++pass) {
    for (int c = 0; c < ncand && !rc; ++c) {
      r->seg_rows = candidates[c];
      double t = 0.0, ms = 0.0;
      int n = 0, ex = 0;
      rc = run_loop_impl(r, 2, &t, &n, nullptr, nullptr);  // warm-up
      if (!rc) {
        t = 0.0;
        n = 0;
        rc = run_loop_impl(r, steps, &t, &n, &ms, &ex);
      }
      if (!rc && ex > 0) cand_ms[(size_t)c] = std::min(cand_ms[(size_t)c], ms / ex);
      e = e2d::launch_copy(r->kp, s1, r->U, r->stream);
      if (e == hipSuccess) e = e2d::launch_copy(r->kp, s2, r->U2, r->stream);
      if (e != hipSuccess && !rc) rc = hip_fail(e, "launch_copy(restore)");
    }
  }
  for (int c = 0; c < ncand; ++c)
    if (cand_ms[(size_t)c] < best_ms) {
      best_ms = cand_ms[(size_t)c];
      best = candidates[c];
    }
  e = hipStreamSynchronize(r->stream);
  if (e != hipSuccess && !rc) rc = hip_fail(e, "hipStreamSynchronize");
  // release the snapshot arrays
  for (void *ptr : {s1.p, s2.p}) {
    if (!ptr) continue;
    for (auto it = r->allocs.begin(); it != r->allocs.end(); ++it)
      if (*it == ptr) {
        r->allocs.erase(it);
        break;
      }
    (void)hipFree(ptr);
  }
  r->timers[1] = timers1;
  r->schedule = schedule;
  if (rc) {
    r->seg_rows = tile0;
    return rc;
  }
  r->seg_rows = best;
  if (best_seg_rows) *best_seg_rows = best;
  if (best_ms_per_step) *best_ms_per_step = best_ms;
  return E2D_OK;
}

int e2d_synchronize(e2d_run *r) {
  int rc = check_run(r);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(r->stream));
  return E2D_OK;
}

// ------------------------------------------------------------------ batch
}  // extern "C"

struct e2d_batch {
  std::vector<e2d_run *> runs;
  int precision = 0;
  int seg_rows = 14;
  int device = 0;
  hipStream_t stream = nullptr;
  e2d::BatchInst *insts = nullptr;   // device
  e2d::StepCtl *ctls = nullptr;      // device, one per instance
  int phase = 0;                     // phase of the next launch in the control rings
  bool ready = false;                // e2d_batch_init has set every instance's loop state
  int nwaves = 0;
  long long cells = 0;
  hipGraphExec_t graph = nullptr;
  int graph_steps = 0;
  int graph_tile = 0;
  int floor_sqrt = 1;  // every instance allows the short wave-speed square root
  int minmod = 1;      // every instance has slope_type 1
  // persistent launch (E2D_BATCH_LAUNCH_PERSISTENT)
  int launch = E2D_BATCH_LAUNCH_PERSISTENT;
  double queue_imbalance = 1.0;          // cells of the fullest XCD queue / the mean
  long long max_items = (1LL << 31) - 1;  // work items of one persistent launch, at most
  bool launch_chosen = false;            // e2d_batch_set_launch has fixed the mode
  unsigned spin_limit = 0;
  int nxcd = 0;                          // XCDs seen by the probe at creation
  int xmap[e2d::kMaxXcds];               // XCC id -> queue (-1: not seen)
  int ncu = 0;                           // CUs of the device
  int *queue_ints = nullptr;             // device: order, wave_prefix, xcd_begin, per_step
  e2d::BatchQueue queue{};
  std::vector<int> queue_per_step;       // host copy of queue.per_step
  e2d::BatchCounters *counters = nullptr;  // device, followed by done[count * kCounterStride]
};

namespace {

// steps per captured graph: a multiple of 3, so that a replay that starts at
// phase 0 of the control rings ends there
constexpr int kGraphSteps = 30;

int batch_check(e2d_batch *b) {
  if (!b) return fail(E2D_ERR_ARG, "null e2d_batch handle");
  hipError_t e = hipSetDevice(b->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  return E2D_OK;
}

void batch_drop_graph(e2d_batch *b) {
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  b->graph = nullptr;
  b->graph_steps = 0;
}

// default bound of a persistent dependency wait, in polls: each is a
// device-scope load of the completion count (about a microsecond) and a sleep
// of about 64 cycles, about a second in all
constexpr unsigned kDefaultSpinLimit = 1u << 20;

// largest queue imbalance (fullest XCD queue's cells over the mean) at which
// a new batch defaults to the persistent launch
constexpr double kPersistentImbalance = 1.10;

// Work queues of the persistent launch for the current tile: instances go to
// the XCD with the least work so far, largest first (cells as the measure);
// within an XCD the largest instances come first in each step.
int batch_upload_queue(e2d_batch *b, const std::vector<int> &waves) {
  const int n = (int)b->runs.size(), nx = b->nxcd;
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  auto cells = [&](int i) { return (long long)b->runs[i]->p.nx * b->runs[i]->p.ny; };
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int c) { return cells(a) > cells(c); });
  std::vector<long long> load(nx, 0);
  std::vector<std::vector<int>> group(nx);
  for (int i : idx) {
    int x = 0;
    for (int k = 1; k < nx; ++k)
      if (load[k] < load[x]) x = k;
    load[x] += cells(i);
    group[x].push_back(i);
  }
  std::vector<int> order, prefix, begin(nx + 1, 0), per(nx, 0);
  for (int x = 0; x < nx; ++x) {
    begin[x] = (int)order.size();
    for (int i : group[x]) {
      order.push_back(i);
      prefix.push_back(per[x]);
      per[x] += waves[i];
    }
  }
  begin[nx] = (int)order.size();
  b->queue_per_step = per;
  long long total = 0, most = 0;
  for (long long l : load) {
    total += l;
    most = std::max(most, l);
  }
  b->queue_imbalance = total > 0 ? (double)most * nx / (double)total : 1.0;
  std::vector<int> h;
  h.insert(h.end(), order.begin(), order.end());
  h.insert(h.end(), prefix.begin(), prefix.end());
  h.insert(h.end(), begin.begin(), begin.end());
  h.insert(h.end(), per.begin(), per.end());
  HIP_TRY(hipMemcpyAsync(b->queue_ints, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice,
                         b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));  // h is read before it goes
  b->queue.order = b->queue_ints;
  b->queue.wave_prefix = b->queue_ints + n;
  b->queue.xcd_begin = b->queue_ints + 2 * n;
  b->queue.per_step = b->queue_ints + 2 * n + nx + 1;
  b->queue.nxcd = nx;
  for (int k = 0; k < e2d::kMaxXcds; ++k) b->queue.xmap[k] = b->xmap[k];
  return E2D_OK;
}

// instance table for the current tile
int batch_upload_table(e2d_batch *b) {
  std::vector<e2d::BatchInst> h(b->runs.size());
  int waves = 0;
  long long cells = 0;
  std::vector<int> inst_waves(b->runs.size());
  for (size_t i = 0; i < b->runs.size(); ++i) {
    e2d_run *r = b->runs[i];
    e2d::BatchInst &bi = h[i];
    std::memset(&bi, 0, sizeof(bi));
    bi.kp = r->kp;
    bi.U = r->U;
    bi.U2 = r->U2;
    bi.wave_begin = waves;
    inst_waves[i] = e2d::batch_waves(r->kp, b->seg_rows, &bi.nstrips);
    waves += inst_waves[i];
    cells += (long long)r->p.nx * r->p.ny;
  }
  b->nwaves = waves;
  b->cells = cells;
  HIP_TRY(hipMemcpyAsync(b->insts, h.data(), h.size() * sizeof(e2d::BatchInst),
                         hipMemcpyHostToDevice, b->stream));
  int rc = batch_upload_queue(b, inst_waves);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(b->stream));
  batch_drop_graph(b);
  return E2D_OK;
}

int batch_launch(e2d_batch *b) {
  e2d::KParams k0 = b->runs[0]->kp;
  k0.s.floor_sqrt = b->floor_sqrt;
  for (const e2d_run *r : b->runs) k0.wide_addr |= r->kp.wide_addr;
  HIP_TRY(e2d::launch_step_batch(k0, b->insts, (int)b->runs.size(), b->ctls, b->phase,
                                 b->nwaves, b->seg_rows, b->stream, b->minmod != 0));
  b->phase = b->phase == 2 ? 0 : b->phase + 1;
  return E2D_OK;
}

int batch_graph(e2d_batch *b) {
  if (b->graph && b->graph_tile == b->seg_rows) return E2D_OK;
  batch_drop_graph(b);
  hipGraph_t g = nullptr;
  // the graph's launches take the phases 0, 1, 2, 0, ... (it is replayed only
  // at phase 0); capturing runs nothing, so the phase is restored afterwards
  const int phase = b->phase;
  b->phase = 0;
  HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  for (int k = 0; k < kGraphSteps; ++k) {
    int rc = batch_launch(b);
    if (rc) {
      b->phase = phase;
      (void)hipStreamEndCapture(b->stream, &g);
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
  }
  b->phase = phase;
  HIP_TRY(hipStreamEndCapture(b->stream, &g));
  hipError_t e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
  b->graph_steps = kGraphSteps;
  b->graph_tile = b->seg_rows;
  return E2D_OK;
}

// state of instance i after the launched steps
int batch_state(e2d_batch *b, int i, e2d::StepState &s) {
  HIP_TRY(hipMemcpyAsync(&s, &b->ctls[i].s[b->phase], sizeof(s), hipMemcpyDeviceToHost,
                         b->stream));
  return E2D_OK;
}

int batch_active(e2d_batch *b, int *active) {
  // one strided copy gathers entry `phase` of every instance's rings
  std::vector<e2d::StepState> s(b->runs.size());
  HIP_TRY(hipMemcpy2DAsync(s.data(), sizeof(e2d::StepState), &b->ctls[0].s[b->phase],
                           sizeof(e2d::StepCtl), sizeof(e2d::StepState), s.size(),
                           hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  int n = 0;
  for (size_t i = 0; i < s.size(); ++i) n += loop_view(b->runs[i]->p, s[i]).active ? 1 : 0;
  *active = n;
  return E2D_OK;
}

// All max_steps steps in one persistent launch (E2D_BATCH_LAUNCH_PERSISTENT),
// or in as few as keep the launch's item count (steps x the largest XCD queue
// step) below 2^31.
int batch_run_persistent(e2d_batch *b, int max_steps, double *elapsed_ms, int *active_left) {
  e2d::KParams k0 = b->runs[0]->kp;
  k0.s.floor_sqrt = b->floor_sqrt;
  for (const e2d_run *r : b->runs) k0.wide_addr |= r->kp.wide_addr;
  const size_t cbytes =
      sizeof(e2d::BatchCounters) + sizeof(unsigned) * e2d::kCounterStride * b->runs.size();
  unsigned *done = reinterpret_cast<unsigned *>(b->counters + 1);
  long long per_max = 1;
  for (int x = 0; x < b->nxcd; ++x) per_max = std::max(per_max, (long long)b->queue_per_step[x]);
  const int cap = (int)std::max<long long>(1, std::min<long long>(INT_MAX, b->max_items / per_max));
  Events<2> ev;
  HIP_TRY(ev.create());
  HIP_TRY(hipEventRecord(ev.e[0], b->stream));
  unsigned err = 0;
  for (int left = max_steps; left > 0 && !err;) {
    const int steps = std::min(left, cap);
    HIP_TRY(hipMemsetAsync(b->counters, 0, cbytes, b->stream));
    HIP_TRY(e2d::launch_step_batch_persistent(
        k0, b->insts, (int)b->runs.size(), b->ctls, b->phase, steps, b->queue, b->counters,
        done, b->seg_rows, b->spin_limit ? b->spin_limit : kDefaultSpinLimit, b->ncu,
        b->stream, b->minmod != 0));
    b->phase = (b->phase + steps) % 3;
    left -= steps;
    e2d::BatchCounters c{};
    HIP_TRY(hipMemcpyAsync(&c, b->counters, sizeof(c), hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    err = c.err;
    // every queue was drained (an XCD that ran no wavefront would leave its
    // instances behind)
    for (int x = 0; x < b->nxcd && !err; ++x)
      if (c.head[x][0] < (unsigned)steps * (unsigned)b->queue_per_step[x]) err = 2;
  }
  HIP_TRY(hipEventRecord(ev.e[1], b->stream));
  HIP_TRY(hipEventSynchronize(ev.e[1]));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, ev.e[0], ev.e[1]));
  if (err == 2) {
    b->ready = false;
    return fail(E2D_ERR_HIP, "persistent batch launch: an XCD queue was not drained "
                             "(the instances are part-way; run e2d_batch_init)");
  }
  if (err) {
    b->ready = false;
    return fail(E2D_ERR_HIP, "persistent batch launch: a dependency wait ran past its limit "
                             "(the instances are part-way; run e2d_batch_init)");
  }
  if (elapsed_ms) *elapsed_ms = ms;
  int act = 0;
  int rc = batch_active(b, &act);
  if (rc) return rc;
  if (active_left) *active_left = act;
  return E2D_OK;
}

}  // namespace

extern "C" {

int e2d_batch_set_launch(e2d_batch *b, int mode, int spin_limit) {
  int rc = batch_check(b);
  if (rc) return rc;
  if (mode != E2D_BATCH_LAUNCH_STEPS && mode != E2D_BATCH_LAUNCH_PERSISTENT)
    return fail(E2D_ERR_ARG, "mode must be E2D_BATCH_LAUNCH_STEPS or _PERSISTENT");
  b->launch = mode;
  b->launch_chosen = true;
  b->spin_limit = spin_limit > 0 ? (unsigned)spin_limit : 0u;
  return E2D_OK;
}

int e2d_batch_get_launch(e2d_batch *b, int *mode) {
  if (!b || !mode) return fail(E2D_ERR_ARG, "null argument");
  *mode = b->launch;
  return E2D_OK;
}


int e2d_batch_queue_info(e2d_batch *b, int *nxcd, double *imbalance) {
  int rc = batch_check(b);
  if (rc) return rc;
  if (!nxcd || !imbalance) return fail(E2D_ERR_ARG, "null argument");
  *nxcd = b->nxcd;
  *imbalance = b->queue_imbalance;
  return E2D_OK;
}

int e2d_batch_create(const e2d_params *params, int count, int precision, e2d_batch **out) {
  if (!params || !out || count <= 0) return fail(E2D_ERR_ARG, "bad argument");
  *out = nullptr;
  return guarded([&]() -> int {
    auto *b = new e2d_batch();
    b->precision = precision;
    for (int i = 0; i < count; ++i) {
      e2d_run *r = nullptr;
      int rc = e2d_create(params + i, precision, &r);
      if (rc) {
        const std::string msg = "instance " + std::to_string(i) + ": " + g_last_error;
        e2d_batch_destroy(b);
        return fail(rc, msg);
      }
      b->runs.push_back(r);
      b->floor_sqrt = b->floor_sqrt && r->kp.s.floor_sqrt;
      b->minmod = b->minmod && r->kp.s.slope_type == 1.0;
      if (r->kp.s.riemann != b->runs[0]->kp.s.riemann) {
        e2d_batch_destroy(b);
        return fail(E2D_ERR_UNSUPPORTED, "all instances of a batch must use the same Riemann solver");
      }
    }
    b->device = b->runs[0]->device;
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      e2d_batch_destroy(b);
      return hip_fail(e, "hipStreamCreate");
    }
    void *p1 = nullptr, *p2 = nullptr;
    if (hipMalloc(&p1, sizeof(e2d::BatchInst) * count) != hipSuccess ||
        hipMalloc(&p2, sizeof(e2d::StepCtl) * count) != hipSuccess) {
      if (p1) (void)hipFree(p1);
      e2d_batch_destroy(b);
      return fail(E2D_ERR_ALLOC, "batch tables");
    }
    b->insts = static_cast<e2d::BatchInst *>(p1);
    b->ctls = static_cast<e2d::StepCtl *>(p2);
    // persistent launch: queue tables, counters, and the XCDs the device has
    // (the highest XCC id seen by 4096 one-wavefront blocks, plus one)
    void *p3 = nullptr, *p4 = nullptr, *p5 = nullptr;
    const size_t qints = 2 * (size_t)count + 2 * e2d::kMaxXcds + 1;
    if (hipMalloc(&p3, qints * sizeof(int)) != hipSuccess ||
        hipMalloc(&p4, sizeof(e2d::BatchCounters) +
                           sizeof(unsigned) * e2d::kCounterStride * count) != hipSuccess ||
        hipMalloc(&p5, 4096 * sizeof(unsigned)) != hipSuccess) {
      for (void *q : {p3, p4, p5})
        if (q) (void)hipFree(q);
      e2d_batch_destroy(b);
      return fail(E2D_ERR_ALLOC, "batch queue tables");
    }
    b->queue_ints = static_cast<int *>(p3);
    b->counters = static_cast<e2d::BatchCounters *>(p4);
    std::vector<unsigned> xcc(4096, 0);
    e = e2d::launch_xcc_probe(static_cast<unsigned *>(p5), 4096, b->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(xcc.data(), p5, xcc.size() * sizeof(unsigned), hipMemcpyDeviceToHost,
                         b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    (void)hipFree(p5);
    if (e != hipSuccess) {
      e2d_batch_destroy(b);
      return hip_fail(e, "XCC probe");
    }
    // one queue per XCC id that blocks ran on
    for (int k = 0; k < e2d::kMaxXcds; ++k) b->xmap[k] = -1;
    for (unsigned x : xcc) b->xmap[x & (e2d::kMaxXcds - 1)] = 0;
    b->nxcd = 0;
    for (int k = 0; k < e2d::kMaxXcds; ++k)
      if (b->xmap[k] == 0) b->xmap[k] = b->nxcd++;
    int ncu = 0;
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, b->device);
    b->ncu = std::max(ncu, 1);
    e = hipMemsetAsync(b->ctls, 0, sizeof(e2d::StepCtl) * count, b->stream);
    if (e != hipSuccess) {
      e2d_batch_destroy(b);
      return hip_fail(e, "hipMemsetAsync(batch tables)");
    }
    int rc = batch_upload_table(b);
    if (rc) {
      const std::string msg = g_last_error;
      e2d_batch_destroy(b);
      return fail(rc, msg);
    }
    // E2D_PERSISTENT_MAX_ITEMS lowers the item count of one persistent launch
    // (below 2^31 anyway): a run then takes several launches, which tests use
    const char *mi = std::getenv("E2D_PERSISTENT_MAX_ITEMS");
    if (mi && *mi) {
      const long long v = std::atoll(mi);
      if (v > 0 && v < b->max_items) b->max_items = v;
    }
    // The persistent launch keeps each instance on one XCD: it is the default
    // when the instances fill the XCD queues evenly, and one launch per step
    // (every instance over the whole GPU) otherwise, e.g. for a few large grids.
    b->launch = b->queue_imbalance <= kPersistentImbalance ? E2D_BATCH_LAUNCH_PERSISTENT
                                                           : E2D_BATCH_LAUNCH_STEPS;
    *out = b;
    return E2D_OK;
  });
}

int e2d_batch_destroy(e2d_batch *b) {
  if (!b) return fail(E2D_ERR_ARG, "null e2d_batch handle");
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  batch_drop_graph(b);
  for (e2d_run *r : b->runs) e2d_destroy(r);
  if (b->insts) (void)hipFree(b->insts);
  if (b->ctls) (void)hipFree(b->ctls);
  if (b->queue_ints) (void)hipFree(b->queue_ints);
  if (b->counters) (void)hipFree(b->counters);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return E2D_OK;
}

int e2d_batch_init(e2d_batch *b) {
  int rc = batch_check(b);
  if (rc) return rc;
  // phase 0 of every instance's rings: t = 0, nStep = 0, the rest cleared
  b->ready = false;
  b->phase = 0;
  HIP_TRY(hipMemsetAsync(b->ctls, 0, sizeof(e2d::StepCtl) * b->runs.size(), b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  for (size_t i = 0; i < b->runs.size(); ++i) {
    e2d_run *r = b->runs[i];
    rc = e2d_init_problem(r);
    if (rc) return rc;
    HIP_TRY(e2d::launch_boundaries(r->kp, r->U, r->stream));
    HIP_TRY(e2d::launch_boundaries(r->kp, r->U2, r->stream));
    HIP_TRY(e2d::launch_dt_reduce(r->kp, r->U, &b->ctls[i].rate[0][0], r->stream));
  }
  for (e2d_run *r : b->runs) HIP_TRY(hipStreamSynchronize(r->stream));
  b->ready = true;
  return E2D_OK;
}

int e2d_batch_set_tile(e2d_batch *b, int seg_rows) {
  int rc = batch_check(b);
  if (rc) return rc;
  if (seg_rows < 1 || seg_rows > 4096) return fail(E2D_ERR_ARG, "seg_rows must be in [1, 4096]");
  b->seg_rows = seg_rows;
  return batch_upload_table(b);
}

int e2d_batch_run(e2d_batch *b, int max_steps, int use_graph, double *elapsed_ms,
                  int *active_left) {
  int rc = batch_check(b);
  if (rc) return rc;
  if (max_steps <= 0) return fail(E2D_ERR_ARG, "max_steps must be > 0");
  if (!b->ready) return fail(E2D_ERR_ARG, "e2d_batch_init has not run on this batch");
  if (b->launch == E2D_BATCH_LAUNCH_PERSISTENT) return batch_run_persistent(b, max_steps,
                                                                            elapsed_ms, active_left);
  if (use_graph) {
    rc = batch_graph(b);
    if (rc) return rc;
  }
  Events<2> ev;
  HIP_TRY(ev.create());
  HIP_TRY(hipEventRecord(ev.e[0], b->stream));
  int done = 0;
  while (done < max_steps) {
    if (use_graph && b->phase == 0 && max_steps - done >= b->graph_steps) {
      HIP_TRY(hipGraphLaunch(b->graph, b->stream));
      done += b->graph_steps;
    } else {
      rc = batch_launch(b);
      if (rc) return rc;
      done += 1;
    }
  }
  HIP_TRY(hipEventRecord(ev.e[1], b->stream));
  HIP_TRY(hipEventSynchronize(ev.e[1]));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, ev.e[0], ev.e[1]));
  if (elapsed_ms) *elapsed_ms = ms;
  int act = 0;
  rc = batch_active(b, &act);
  if (rc) return rc;
  if (active_left) *active_left = act;
  return E2D_OK;
}

int e2d_batch_status(e2d_batch *b, int instance, double *t, int *nStep) {
  int rc = batch_check(b);
  if (rc) return rc;
  if (instance < 0 || instance >= (int)b->runs.size()) return fail(E2D_ERR_ARG, "bad instance");
  e2d::StepState s{};
  rc = batch_state(b, instance, s);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(b->stream));
  if (t) *t = s.t;
  if (nStep) *nStep = s.nstep;
  return E2D_OK;
}

int e2d_batch_copy_state(e2d_batch *b, int instance, double *host, int layout) {
  double t;
  int n = 0;
  int rc = e2d_batch_status(b, instance, &t, &n);
  if (rc) return rc;
  return e2d_copy_state(b->runs[instance], n % 2 == 0 ? E2D_U : E2D_U2, host, layout);
}

int e2d_batch_info(e2d_batch *b, long long *cells, int *waves, int *count) {
  if (!b) return fail(E2D_ERR_ARG, "null e2d_batch handle");
  if (cells) *cells = b->cells;
  if (waves) *waves = b->nwaves;
  if (count) *count = (int)b->runs.size();
  return E2D_OK;
}

int e2d_batch_autotune(e2d_batch *b, const int *candidates, int ncand, int steps,
                       int *best_seg_rows, double *best_ms_per_step) {
  int rc = batch_check(b);
  if (rc) return rc;
  if (!candidates || ncand <= 0 || steps <= 0) return fail(E2D_ERR_ARG, "bad argument");
  if (!b->ready) return fail(E2D_ERR_ARG, "e2d_batch_init has not run on this batch");
  // the launch modes to try: the one e2d_batch_set_launch fixed, else both
  // where the queues are even enough for the persistent launch
  std::vector<int> modes{b->launch};
  if (!b->launch_chosen && b->queue_imbalance <= kPersistentImbalance)
    modes = {E2D_BATCH_LAUNCH_PERSISTENT, E2D_BATCH_LAUNCH_STEPS};
  const int nm = (int)modes.size();
  // two interleaved passes; the faster timing of each (mode, tile) counts
  std::vector<double> cand_ms((size_t)ncand * nm, 1e300);
  for (int pass = 0; pass < 2; ++pass) {
    for (int m = 0; m < nm; ++m) {
      b->launch = modes[(size_t)m];
      for (int c = 0; c < ncand; ++c) {
        rc = e2d_batch_set_tile(b, candidates[c]);
        if (rc) return rc;
        double ms = 0.0;
        int act = 0;
        for (int attempt = 0; attempt < 2; ++attempt) {
          rc = e2d_batch_run(b, 2, 0, &ms, &act);  // warm-up
          if (rc) return rc;
          rc = e2d_batch_run(b, steps, 0, &ms, &act);
          if (rc) return rc;
          // a run in which instances reached the end of their loops timed
          // launches that returned at once: start the batch again and repeat it
          if (act == (int)b->runs.size() || attempt == 1) break;
          rc = e2d_batch_init(b);
          if (rc) return rc;
        }
        double &slot = cand_ms[(size_t)m * ncand + c];
        slot = std::min(slot, ms / steps);
      }
    }
  }
  int best = b->seg_rows, best_mode = modes[0];
  double best_ms = 1e300;
  for (int m = 0; m < nm; ++m)
    for (int c = 0; c < ncand; ++c)
      if (cand_ms[(size_t)m * ncand + c] < best_ms) {
        best_ms = cand_ms[(size_t)m * ncand + c];
        best = candidates[c];
        best_mode = modes[(size_t)m];
      }
  b->launch = best_mode;
  rc = e2d_batch_set_tile(b, best);
  if (rc) return rc;
  if (best_seg_rows) *best_seg_rows = best;
  if (best_ms_per_step) *best_ms_per_step = best_ms;
  return E2D_OK;
}

}  // extern "C"
