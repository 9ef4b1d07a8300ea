// kernels.hpp -- host-side declarations of the HIP kernels' launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "euler2d_hip.h"
#include "numerics.hpp"

namespace e2d {

// One state array on the device, as SoA planes [plane][j][i] with a row pitch
// and a plane stride in elements of the storage scalar:
//   FP64        4 planes of double, plane v = variable v
//   FP32 split  4 planes of (hi, lo) float pairs, plane v = variable v
//               (value = hi + lo, |lo| <= ulp(hi)/2; hi at the lower address)
//   FP32        4 planes of float
struct Grid {
  void *p;
  int pitch;
  long long plane;
};

// Everything a kernel needs from HydroParams, by value.
struct KParams {
  Scheme s;             // FP64 constants
  SchemeT<float> sf;    // the same constants rounded to float (split path)
  int prec;             // E2D_PRECISION_*
  int nx, ny, isize, jsize;
  int bcxmin, bcxmax, bcymin, bcymax;
  int nStepmax;
  double dx, dy, cfl, tEnd;
  float dxf, dyf;
  RateT<double> rd;     // cell sizes of the CFL rate (FP64 / FP32 arithmetic)
  RateT<float> rf;
  // split path: primitive conversion redone in FP64 where the internal energy
  // is below this fraction of the total energy (p < kin_frac * (gamma-1) * E)
  float kin_frac;
  // 1: the run-loop and batched step kernels address rows with 64-bit scalar
  // arithmetic even where every plane is at most 4 GiB (set for planes above
  // 4 GiB, or through E2D_WIDE_ADDRESSING=1 to exercise that form)
  int wide_addr;
  // SIMDs of the device (4 per CU), for the fused step's issue-priority rule
  int simds;
};

// Run-loop control block kept in device memory: the time loop of
// main.cpp:100-145 without a host round trip per step.  Launch L of a loop
// (phase ph = L % 3) reads s[ph], the time and step count before its step, and
// rate[ph], the CFL rate of its input, from which every wavefront derives the
// same dt (HydroRun.h:246, with the tEnd clamp) and while condition.  One
// wavefront writes s[(ph + 1) % 3] and clears rate[(ph + 2) % 3]; every
// wavefront folds the rate of its part of the step's output into
// rate[(ph + 1) % 3].  No launch reads what it writes, so nothing waits for
// another wavefront at the end of a step.
struct StepState {
  double t;   // time before the step
  int nstep;  // steps executed before it
  int pad;
};
constexpr int kBeginBlocks = 512;  // blocks (and partials) of loop_begin_kernel
constexpr int kRateSlots = 16;   // wavefront w folds into slot w % kRateSlots
constexpr int kSlotStride = 16;  // one 128-byte line per slot
struct StepCtl {
  StepState s[3];
  // host-visible copy of s (pinned host memory, or null): the step's writer
  // stores there too, so that the host reads t and nStep after a stream
  // synchronisation without a copy command
  StepState *host_s;
  unsigned arrivals;  // loop_begin_kernel's block counter, 0 between launches
  unsigned pad32;
  unsigned long long pad[8];
  // max over the slots of (c+|u|)/dx+(c+|v|)/dy as IEEE bits (0 = none yet)
  unsigned long long rate[3][kRateSlots * kSlotStride];
};

// One instance of a batch as the batched step kernel sees it.
struct BatchInst {
  KParams kp;
  Grid U, U2;
  int wave_begin;  // first global wavefront index of this instance
  int nstrips;
  int pad[2];
};

// Work queues of the persistent batched step.  Every instance is assigned to
// one XCD (accelerator complex die, with its own L2); the wavefronts running on
// that XCD take the instance's work items, so the rows one step writes are
// read by the next from the same L2.  Items of a queue are ordered step-major:
// step n of every instance of the XCD, then step n + 1.  An item is one
// wavefront's march (strip, segment) of one instance.
constexpr int kMaxXcds = 16;  // the XCC id is a 4-bit field
struct BatchQueue {
  const int *order;        // instance ids grouped by XCD (xcd_begin[x] .. xcd_begin[x + 1])
  const int *wave_prefix;  // per position of `order`: wavefronts of the XCD's earlier instances
  const int *xcd_begin;    // nxcd + 1 offsets into order
  const int *per_step;     // per XCD: items (wavefronts) of one step of all its instances
  int nxcd;
  int xmap[kMaxXcds];      // XCC id -> queue (-1: no queue; the probe saw no block there)
};
// Counters of one persistent run, zeroed before the launch: the next item of
// each XCD's queue, an error flag (a dependency wait ran past its limit), and
// then, per instance, the number of its items completed (done[i * kCounterStride]).
// Every counter has a 256-byte line of its own: device-scope atomics on one
// line serialise, and each queue head takes one atomic per item taken.
constexpr int kCounterStride = 64;
struct BatchCounters {
  unsigned head[kMaxXcds][kCounterStride];
  unsigned err;
  unsigned pad[kCounterStride - 1];
};

// Geometry of the radial profile (ComputeRadialProfileFunctor.h:83-100)
struct RadialGeom {
  double xmin, ymin, cx, cy, max_radial_distance;
  int nbins;
};

// fused step geometry: one wave of 64 lanes covers 60 output columns (plus a
// 2-column halo on each side) and marches over `seg_rows` output rows
constexpr int kLanes = 64;
constexpr int kOutCols = 60;
constexpr int kBlock = 256;

// Work map of the fused step.  A wavefront covers one segment of one strip.
// The edge strips (0 and nstrips - 1) also write ghost columns and take about
// 10% longer per row, so their segments are shorter: tile_e = (tile + 4) *
// 10 / 11 - 4 rows, which with the 4 warm-up rows of every segment makes an
// edge wavefront end with the interior ones.  Interior-strip wavefronts come
// first in the launch.
struct FusedMap {
  int tile, tile_e;  // rows per segment: interior strips, edge strips
  int nstrips, nedge;  // strips of 60 columns; edge strips (min(nstrips, 2))
  int n_int;           // wavefronts of the interior strips
  int nwaves;          // all wavefronts
  // 1 when every wavefront is resident at once, two per SIMD: a wavefront then
  // lowers its issue priority as its march progresses (s_setprio), so that the
  // two on a SIMD finish together instead of the younger running alone last
  int prio;
};
FusedMap fused_map(const KParams &kp, int seg_rows);

KParams make_kparams(const e2d_params &p, int precision);
// planes of one state array (one per conserved variable) and bytes of one
// element: a double (FP64), a (hi, lo) pair of floats (split) or a float
int grid_planes(int precision);
int grid_scalar_bytes(int precision);

hipError_t launch_init(const KParams &kp, const e2d_params &p, Grid U, unsigned int *count_dev,
                       hipStream_t st);
hipError_t launch_blast_energy(const KParams &kp, const e2d_params &p, Grid U, double e_inside,
                               hipStream_t st);
hipError_t launch_boundaries(const KParams &kp, Grid U, hipStream_t st);
hipError_t launch_dt_reduce(const KParams &kp, Grid U, unsigned long long *invdt_bits,
                            hipStream_t st);
// start of a run loop from (t, nstep): CFL rate of `in` into ring phase 0 of
// *ctl, with `bc` also the ghost cells of `in` and `other`; partials holds
// kBeginBlocks doubles
hipError_t launch_loop_begin(const KParams &kp, Grid in, Grid other, int bc, StepCtl *ctl,
                             double t, int nstep, double *partials, hipStream_t st);
// count of interior cells below the kinetic-energy threshold kp.kin_frac (added to *count)
hipError_t launch_kinetic_count(const KParams &kp, Grid U, unsigned long long *count,
                                hipStream_t st);
// reference pass schedule (FP64 only)
hipError_t launch_prim(const KParams &kp, Grid U, Grid Q, hipStream_t st);
hipError_t launch_fluxes(const KParams &kp, Grid Q, Grid Fx, Grid Fy, double dtdx, double dtdy,
                         hipStream_t st);
hipError_t launch_update(const KParams &kp, Grid U, Grid Fx, Grid Fy, hipStream_t st);
// implementation version 1: stored slopes, then trace + flux and update per
// direction (dir 0 = x, 1 = y), FP64 only
hipError_t launch_slopes(const KParams &kp, Grid Q, Grid Sx, Grid Sy, hipStream_t st);
hipError_t launch_trace_fluxes(const KParams &kp, int dir, Grid Q, Grid Sx, Grid Sy, Grid F,
                               double dtdx, double dtdy, hipStream_t st);
hipError_t launch_update_dir(const KParams &kp, int dir, Grid U, Grid F, hipStream_t st);
// implementation version 2: fluxes added to both neighbours with FP64 atomics
hipError_t launch_fluxes_atomic(const KParams &kp, Grid Q, Grid U, double dtdx, double dtdy,
                                hipStream_t st);
// ctl == nullptr: explicit dt (HydroRun::godunov_unsplit); else run-loop mode,
// launch phase ph (0..2) of the loop's control rings
hipError_t launch_step_fused(const KParams &kp, Grid in, Grid out, double dt, StepCtl *ctl,
                             int ph, int seg_rows, hipStream_t st);
hipError_t launch_sums(const KParams &kp, Grid U, double *partials, double *out4, hipStream_t st);
int sums_partials_needed(const KParams &kp);
// wavefronts the fused step launches for one grid (and its strip count)
int batch_waves(const KParams &kp, int seg_rows, int *nstrips);
// one step of every active instance; kp0 gives the precision and solver of all
// (kp0.s.floor_sqrt must hold for every instance; kp0.wide_addr must be set
// when any instance has a plane above 4 GiB; minmod: every instance has
// slope_type 1)
hipError_t launch_step_batch(const KParams &kp0, const BatchInst *insts_dev, int ninst,
                             StepCtl *ctls_dev, int ph, int nwaves, int seg_rows, hipStream_t st,
                             bool minmod = false);
// All `steps` steps of every instance in one persistent launch: as many blocks
// of 256 threads as the kernel fits on `ncu` CUs at once.  counters and done
// (one word every kCounterStride per instance) must be zero; steps times the
// largest per_step must stay below 2^31.
// spin_limit bounds every dependency wait (polls); a wait that runs
// past it sets counters->err and the launch ends early.
hipError_t launch_step_batch_persistent(const KParams &kp0, const BatchInst *insts_dev, int ninst,
                                        StepCtl *ctls_dev, int ph0, int steps, BatchQueue q,
                                        BatchCounters *counters, unsigned *done, int seg_rows,
                                        unsigned spin_limit, int ncu, hipStream_t st,
                                        bool minmod = false);
// XCC id of the XCD each of `blocks` one-wavefront blocks ran on, into out[block]
hipError_t launch_xcc_probe(unsigned *out, int blocks, hipStream_t st);
// per-bin cell counts and density sums of the radial profile (nbins entries each)
hipError_t launch_radial_profile(const KParams &kp, Grid U, const RadialGeom &geo,
                                 unsigned long long *counts, double *sums, hipStream_t st);
// *p = v, one thread (a kernel, so that it may be captured in a graph)
hipError_t launch_fill_u32(unsigned *p, unsigned v, hipStream_t st);
// whole-array copy (all planes, ghosts and padding included)
hipError_t launch_copy(const KParams &kp, Grid src, Grid dst, hipStream_t st);

}  // namespace e2d
