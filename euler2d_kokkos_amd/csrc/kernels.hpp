// kernels.hpp -- host-side declarations of the FP64 HIP kernels' launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "euler2d_hip.h"
#include "numerics.hpp"

namespace e2d {

// One state array on the device: four SoA planes [var][j][i], row pitch in
// elements, plane stride in elements.  Element (v,i,j) at p[v*plane + j*pitch + i].
struct Grid {
  double *p;
  int pitch;
  long long plane;
};

// Everything a kernel needs from HydroParams, by value.
struct KParams {
  Scheme s;
  int nx, ny, isize, jsize;
  int bcxmin, bcxmax, bcymin, bcymax;
  int nStepmax;
  double dx, dy, cfl, tEnd;
};

// Run-loop control block kept in device memory (time loop of main.cpp:100-145
// without a host round trip per step).
struct StepCtl {
  double t;                     // time at the start of the pending step
  double dt;                    // dt of the pending step (already clamped to tEnd)
  unsigned long long invdt;     // max of (c+|u|)/dx+(c+|v|)/dy as IEEE bits
  int nstep;                    // steps executed so far
  int active;                   // pending step runs iff t < tEnd && nstep < nStepmax
  unsigned int blocks_done;     // last-block counter of the fused step kernel
  int pad;
};

// fused step geometry: one wave of 64 lanes covers 60 output columns (plus a
// 2-column halo on each side) and marches over `seg_rows` output rows
constexpr int kLanes = 64;
constexpr int kOutCols = 60;
constexpr int kBlock = 256;

KParams make_kparams(const e2d_params &p);

hipError_t launch_init(const e2d_params &p, Grid U, unsigned int *count_dev, hipStream_t st);
hipError_t launch_blast_energy(const e2d_params &p, Grid U, double e_inside, hipStream_t st);
hipError_t launch_boundaries(const KParams &kp, Grid U, hipStream_t st);
hipError_t launch_dt_reduce(const KParams &kp, Grid U, unsigned long long *invdt_bits,
                            hipStream_t st);
hipError_t launch_dt_finalize(const KParams &kp, StepCtl *ctl, hipStream_t st);
hipError_t launch_prim(const KParams &kp, Grid U, Grid Q, hipStream_t st);
hipError_t launch_fluxes(const KParams &kp, Grid Q, Grid Fx, Grid Fy, double dtdx, double dtdy,
                         hipStream_t st);
hipError_t launch_update(const KParams &kp, Grid U, Grid Fx, Grid Fy, hipStream_t st);
// ctl == nullptr: explicit dt (HydroRun::godunov_unsplit); else run-loop mode
hipError_t launch_step_fused(const KParams &kp, Grid in, Grid out, double dt, StepCtl *ctl,
                             int seg_rows, hipStream_t st);
hipError_t launch_sums(const KParams &kp, Grid U, double *partials, double *out4, hipStream_t st);
int sums_partials_needed(const KParams &kp);
// whole-array copy (all four planes, ghosts and padding included)
hipError_t launch_copy(Grid src, Grid dst, hipStream_t st);

}  // namespace e2d
