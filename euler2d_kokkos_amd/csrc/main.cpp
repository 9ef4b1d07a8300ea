// main.cpp -- `euler2d` command-line driver over the C ABI.
//
// Mirrors the reference driver (pkestene/euler2d_kokkos src/main.cpp:24-212):
// one positional .ini path, HydroParams::setup + print, HydroRun construction,
// startup dt and boundary fills, the time loop with the VTK output every
// nOutput steps and at the end (main.cpp:109-161), the radial density profile
// of a Sedov blast (main.cpp:175-179), the timing table and the Perf line.
// HDF5/XDMF output is not built (no HDF5 library): with use_HDF5 set, a note
// is printed instead.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <optional>
#include <string>
#include <vector>

#include "euler2d_hip.h"
#include "trace.hpp"

static int die(const char *what) {
  std::fprintf(stderr, "%s: %s\n", what, e2d_last_error());
  return EXIT_FAILURE;
}

int main(int argc, char *argv[]) {
  const char *ini = nullptr;
  int schedule = -1;  // -1: the one implementationVersion selects
  for (int a = 1; a < argc; ++a) {
    if (!std::strcmp(argv[a], "--schedule=passes"))
      schedule = E2D_SCHEDULE_PASSES;
    else if (!std::strcmp(argv[a], "--schedule=fused"))
      schedule = E2D_SCHEDULE_FUSED;
    else if (!std::strcmp(argv[a], "--schedule=directional"))
      schedule = E2D_SCHEDULE_DIRECTIONAL;
    else if (!std::strcmp(argv[a], "--schedule=atomic"))
      schedule = E2D_SCHEDULE_ATOMIC;
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

  // output location, read where HydroRun::saveVTK reads it (HydroRun.h:530-531)
  std::string out_dir = "./", out_prefix = "output";
  {
    e2d_config *cfg = nullptr;
    int cerr = 0;
    if (e2d_config_open(ini, &cfg, &cerr) == E2D_OK) {
      char buf[4096];
      // returns the value's full length (negative on error)
      if (e2d_config_get_string(cfg, "output", "outputDir", "./", buf, sizeof(buf)) >= 0)
        out_dir = buf;
      if (e2d_config_get_string(cfg, "output", "outputPrefix", "output", buf, sizeof(buf)) >= 0)
        out_prefix = buf;
      e2d_config_close(cfg);
    }
  }

  e2d_run *run = nullptr;
  if (e2d_create(&params, E2D_PRECISION_FP64, &run)) return die("create");
  if (schedule >= 0 && e2d_set_schedule(run, schedule)) return die("schedule");
  if (e2d_init_problem(run)) return die("init");

  const bool output = params.enableOutput && params.nOutput > 0;
  bool hdf5_note = false;
  double t_io = 0.0;
  // HydroRun::saveData (HydroRun.h:486-497) of the current state array
  auto save = [&](int step, double t) -> int {
    e2d::Region reg("output");
    std::printf("Output results at time t=%g step %d\n", t, step);
    const auto s0 = std::chrono::steady_clock::now();
    if (params.ioVTK &&
        e2d_save_vtk(run, step % 2 == 0 ? E2D_U : E2D_U2, out_dir.c_str(), out_prefix.c_str(),
                     step))
      return -1;
    if (params.ioHDF5 && !hdf5_note) {
      std::printf("use_HDF5: HDF5 output is not part of this build; skipped\n");
      hdf5_note = true;
    }
    t_io += std::chrono::duration<double>(std::chrono::steady_clock::now() - s0).count();
    return 0;
  };

  double t = 0.0;
  int nStep = 0;
  std::printf("Start computation....\n");
  const auto t0 = std::chrono::steady_clock::now();
  std::optional<e2d::Region> main_loop;
  main_loop.emplace("main_loop");
  // chunks end at every 10th step (the reference's progress print,
  // main.cpp:103-106) and at every output step
  while (t < params.tEnd && nStep < params.nStepmax) {
    if (nStep % 10 == 0) std::printf("time step=%7d (t=% 10.8f)\n", nStep, t);
    if (output && nStep % params.nOutput == 0 && save(nStep, t)) return die("output");
    int chunk = 10 - nStep % 10;
    if (output) chunk = std::min(chunk, params.nOutput - nStep % params.nOutput);
    if (e2d_run_loop(run, chunk, &t, &nStep)) return die("run_loop");
  }
  if (output && save(nStep, t)) return die("output");  // last state (main.cpp:148-161)
  if (e2d_synchronize(run)) return die("sync");
  main_loop.reset();
  const double t_tot =
    std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

  // Sedov blast post-processing: radial profile of U (always U, as the
  // reference passes hydro->U), written as two .npy files (main.cpp:175-179)
  if (params.problemType == 1 && params.blast_total_energy_inside > 0) {
    const int nb = params.blast_nbins;
    std::vector<double> dist(nb), dens(nb);
    if (e2d_radial_profile(run, E2D_U, nb, dist.data(), dens.data(), nullptr) ||
        e2d_write_npy("sedov_blast_radial_distances.npy", dist.data(), nb) ||
        e2d_write_npy("sedov_blast_density_profile.npy", dens.data(), nb))
      return die("radial profile");
  }

  double tm[5];
  e2d_timers(run, tm);
  std::printf("total           time : %5.3f secondes\n", t_tot);
  std::printf("godunov         time : %5.3f secondes %5.2f%%\n", tm[1], 100 * tm[1] / t_tot);
  std::printf("primitive       time : %5.3f secondes %5.2f%%\n", tm[2], 100 * tm[2] / t_tot);
  std::printf("compute fluxes  time : %5.3f secondes %5.2f%%\n", tm[3], 100 * tm[3] / t_tot);
  std::printf("update hydro    time : %5.3f secondes %5.2f%%\n", tm[4], 100 * tm[4] / t_tot);
  std::printf("boundaries      time : %5.3f secondes %5.2f%%\n", tm[0], 100 * tm[0] / t_tot);
  std::printf("io              time : %5.3f secondes %5.2f%%\n", t_io, 100 * t_io / t_tot);
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
