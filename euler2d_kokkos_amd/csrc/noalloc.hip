// noalloc.hip -- e2d_run_loop on caller-owned state arrays, with every scratch
// word of the loop in a file-scope device array: no device or host allocation.
//
// The run loop's kernels (loop_begin_kernel, step_fused_kernel) need device
// memory beyond the state arrays: the control rings (StepCtl: t / nStep and
// the CFL-rate slots of three launch phases, plus the loop start's arrival
// counter) and the loop start's per-block partial maxima.  The handle API
// allocates them in e2d_create; here they live in `scratch`, one copy per GPU,
// held until the process ends.  Host code never passes `scratch` itself to a
// kernel: its device address comes from hipGetSymbolAddress, once per device.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "euler2d_hip.h"
#include "kernels.hpp"
#include "status.hpp"

// Zero at module load; the loop start and the steps initialise every word
// they read (ring phases 0 and 1, the partials), and the arrival counter is
// cleared by a fill kernel before each loop start.
static __device__ __align__(256) unsigned char scratch[E2D_NOALLOC_SCRATCH_BYTES];

namespace {

using e2d::Grid;

// What the loop uses of the scratch array.
struct NoallocScratch {
  e2d::StepCtl ctl;
  double partials[e2d::kBeginBlocks];
};
static_assert(sizeof(NoallocScratch) <= E2D_NOALLOC_SCRATCH_BYTES,
              "the no-allocation loop's scratch outgrew its static array");

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
void *g_scratch_dev[kMaxDevices] = {};

int fail(int code, const std::string &msg) { return e2d::report(code, msg); }

int hip_fail(hipError_t e, const char *what) {
  return fail(E2D_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define NA_TRY(expr)                                  \
  do {                                                \
    hipError_t e_ = (expr);                           \
    if (e_ != hipSuccess) return hip_fail(e_, #expr); \
  } while (0)

// Device address of this device's copy of `scratch` (the current device must
// be `dev`).
int scratch_address(int dev, NoallocScratch **out) {
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_scratch_dev[dev]) {
    void *ptr = nullptr;
    NA_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(E2D_ERR_HIP, "hipGetSymbolAddress(scratch) returned null");
    g_scratch_dev[dev] = ptr;
  }
  *out = static_cast<NoallocScratch *>(g_scratch_dev[dev]);
  return E2D_OK;
}

bool aligned(const void *p, unsigned a) {
  return (reinterpret_cast<unsigned long long>(p) & (a - 1)) == 0;
}

}  // namespace

extern "C" int e2d_run_loop_noalloc(const e2d_params *p, int precision, void *U, void *U2,
                                    long long pitch, long long plane, int seg_rows,
                                    int max_steps, double *t, int *nStep, void *stream) {
  if (!p || !U || !U2 || !t || !nStep) return fail(E2D_ERR_ARG, "null argument");
  int rc = e2d_params_validate(p);
  if (rc) return rc;
  if (precision != E2D_PRECISION_FP64 && precision != E2D_PRECISION_FP32_SPLIT &&
      precision != E2D_PRECISION_FP32)
    return fail(E2D_ERR_ARG, "precision must be one of E2D_PRECISION_FP64 / _FP32_SPLIT / _FP32");
  if (U == U2) return fail(E2D_ERR_ARG, "U and U2 must be distinct arrays");
  if (*nStep < 0) return fail(E2D_ERR_ARG, "nStep must be >= 0");
  if (seg_rows < 0 || seg_rows > 4096) return fail(E2D_ERR_ARG, "seg_rows must be in [0, 4096]");
  // the state layout: 4 planes of isize x jsize cells, rows pitch elements
  // apart, planes `plane` elements apart
  const int esize = precision == E2D_PRECISION_FP32 ? 4 : 8;
  if (pitch < p->isize || pitch > (1LL << 30))
    return fail(E2D_ERR_ARG, "pitch must be at least isize (nx + 4) elements");
  if (plane < pitch * p->jsize)
    return fail(E2D_ERR_ARG, "plane must be at least pitch * jsize (ny + 4) elements");
  if (!aligned(U, esize) || !aligned(U2, esize))
    return fail(E2D_ERR_ARG, "U and U2 must be aligned to their element size");

  int dev = 0;
  NA_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(E2D_ERR_HIP, "device index out of range");
  NoallocScratch *sc = nullptr;
  rc = scratch_address(dev, &sc);
  if (rc) return rc;
  int ncu = 0;
  NA_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));

  e2d::KParams kp = e2d::make_kparams(*p, precision);
  kp.simds = 4 * ncu;
  kp.wide_addr = (unsigned long long)plane * (unsigned long long)esize > (1ull << 32);
  const Grid gU{U, (int)pitch, plane};
  const Grid gU2{U2, (int)pitch, plane};
  const int tile = seg_rows > 0 ? seg_rows : 16;
  hipStream_t st = static_cast<hipStream_t>(stream);

  // loop start: arrival counter cleared, then the CFL rate of the step's input
  // (and at nStep == 0 the ghost cells of both arrays) into ring phase 0
  const int n0 = *nStep;
  const bool even0 = n0 % 2 == 0;
  NA_TRY(e2d::launch_fill_u32(&sc->ctl.arrivals, 0u, st));
  NA_TRY(e2d::launch_loop_begin(kp, even0 ? gU : gU2, even0 ? gU2 : gU, n0 == 0, &sc->ctl, *t,
                                n0, sc->partials, st));
  // main.cpp:100-145 with dt on the device; past the end of the loop a launch
  // returns at once.  Without a budget the loop state is read every 256
  // launches.
  int phase = 0, launched = 0;
  e2d::StepState s{};
  for (;;) {
    int todo = max_steps > 0 ? max_steps - launched : 256;
    const long long left = (long long)p->nStepmax - (n0 + launched);
    if (left < todo) todo = left > 0 ? (int)left : 0;
    for (int k = 0; k < todo; ++k) {
      const int n = n0 + launched + k;
      const Grid &in = n % 2 == 0 ? gU : gU2;
      const Grid &out = n % 2 == 0 ? gU2 : gU;
      NA_TRY(e2d::launch_step_fused(kp, in, out, 0.0, &sc->ctl, phase, tile, st));
      phase = phase == 2 ? 0 : phase + 1;
    }
    launched += todo;
    NA_TRY(hipMemcpyAsync(&s, &sc->ctl.s[phase], sizeof(s), hipMemcpyDeviceToHost, st));
    NA_TRY(hipStreamSynchronize(st));
    const bool active = s.t < p->tEnd && s.nstep < p->nStepmax;
    if (todo <= 0 || !active || (max_steps > 0 && launched >= max_steps)) break;
  }
  *t = s.t;
  *nStep = s.nstep;
  return E2D_OK;
}
