// main.cpp -- `euler2d` command-line driver over the C ABI.
//
// Mirrors the reference driver (pkestene/euler2d_kokkos src/main.cpp:24-212):
// one positional .ini path, HydroParams::setup + print, HydroRun construction,
// startup dt and boundary fills, the time loop, the timing table and the Perf
// line.  Field output (VTK/HDF5) is not produced by this build.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "euler2d_hip.h"

static int die(const char *what) {
  std::fprintf(stderr, "%s: %s\n", what, e2d_last_error());
  return EXIT_FAILURE;
}

int main(int argc, char *argv[]) {
  const char *ini = nullptr;
  int schedule = E2D_SCHEDULE_FUSED;
  for (int a = 1; a < argc; ++a) {
    if (!std::strcmp(argv[a], "--schedule=passes"))
      schedule = E2D_SCHEDULE_PASSES;
    else if (!std::strcmp(argv[a], "--schedule=fused"))
      schedule = E2D_SCHEDULE_FUSED;
    else if (!ini)
      ini = argv[a];
    else
      ini = nullptr, a = argc;
  }
  if (!ini) {
    std::fprintf(stderr,
                 "Error: wrong number of argument; input filename must be "
                 "the only parameter on the command line\n");
    return EXIT_FAILURE;
  }

  char dev[256] = "";
  int ncu = 0;
  if (e2d_device_info(dev, sizeof(dev), &ncu)) return die("device");
  std::printf("##########################\nHIP device: %s, %d CUs\n##########################\n",
              dev, ncu);

  e2d_params params;
  int perr = 0;
  e2d_params_from_ini(ini, &params, &perr);  // missing file -> defaults, as the reference
  std::printf("Using Euler implementation version %d\n", params.implementationVersion);
  e2d_params_print(&params);

  e2d_run *run = nullptr;
  if (e2d_create(&params, E2D_PRECISION_FP64, &run)) return die("create");
  if (e2d_set_schedule(run, schedule)) return die("schedule");
  if (e2d_init_problem(run)) return die("init");

  double t = 0.0;
  int nStep = 0;
  std::printf("Start computation....\n");
  const auto t0 = std::chrono::steady_clock::now();
  // 10-step chunks reproduce the reference's progress print (main.cpp:103-106)
  while (t < params.tEnd && nStep < params.nStepmax) {
    std::printf("time step=%7d (t=% 10.8f)\n", nStep, t);
    if (e2d_run_loop(run, 10, &t, &nStep)) return die("run_loop");
  }
  if (e2d_synchronize(run)) return die("sync");
  const double t_tot =
    std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

  double tm[5];
  e2d_timers(run, tm);
  std::printf("total           time : %5.3f secondes\n", t_tot);
  std::printf("godunov         time : %5.3f secondes %5.2f%%\n", tm[1], 100 * tm[1] / t_tot);
  std::printf("primitive       time : %5.3f secondes %5.2f%%\n", tm[2], 100 * tm[2] / t_tot);
  std::printf("compute fluxes  time : %5.3f secondes %5.2f%%\n", tm[3], 100 * tm[3] / t_tot);
  std::printf("update hydro    time : %5.3f secondes %5.2f%%\n", tm[4], 100 * tm[4] / t_tot);
  std::printf("boundaries      time : %5.3f secondes %5.2f%%\n", tm[0], 100 * tm[0] / t_tot);
  // reference figure: ghost cells included, whole loop (main.cpp:202-203)
  std::printf("Perf                 : %10.2f number of Mcell-updates/s\n",
              1.0 * nStep * params.isize * params.jsize / t_tot * 1e-6);
  std::printf("Perf (interior)      : %10.2f number of Mcell-updates/s\n",
              1.0 * nStep * params.nx * params.ny / t_tot * 1e-6);
  double sums[4];
  if (e2d_conserved_sums(run, nStep % 2 == 0 ? E2D_U : E2D_U2, sums) == 0)
    std::printf("final t=%.17g nStep=%d sums: rho %.17g E %.17g mx %.17g my %.17g\n", t, nStep,
                sums[0], sums[1], sums[2], sums[3]);
  e2d_destroy(run);
  return EXIT_SUCCESS;
}
