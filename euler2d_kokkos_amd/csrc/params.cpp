// params.cpp -- HydroParams::setup / init / print on the e2d_params POD.
//
// Follows src/HydroParams.cpp:19-224 of pkestene/euler2d_kokkos: the same keys,
// defaults and fall-backs, every real read through the float-returning
// ConfigMap::getFloat and widened to double.
#include "params.hpp"

#include <cstdio>
#include <cstring>

namespace e2d {

namespace {
// the reference passes double literals to float-typed default arguments
constexpr float F(double x) { return static_cast<float>(x); }
}  // namespace

void params_defaults(e2d_params &p) {
  // HydroParams() and HydroSettings() constructors, src/HydroParams.h:131-140,227-266
  std::memset(&p, 0, sizeof(p));
  p.enableOutput = 1;
  p.ghostWidth = 2;
  p.xmax = 1.0;
  p.ymax = 1.0;
  p.boundary_type_xmin = p.boundary_type_xmax = 1;
  p.boundary_type_ymin = p.boundary_type_ymax = 1;
  p.ioVTK = 1;
  p.gamma0 = 1.4;
  p.gamma6 = 1.0;
  p.cfl = 1.0;
  p.slope_type = 2.0;
  p.smallr = 1e-8;
  p.smallc = 1e-8;
  p.smallp = 1e-6;
  p.smallpp = 1e-6;
  p.niter_riemann = 10;
  p.blast_radius = 10.0;
  p.blast_density_in = 1.0;
  p.blast_density_out = 1.2;
  p.blast_pressure_in = 10.0;
  p.blast_pressure_out = 0.1;
  p.blast_nbins = 100;
  p.implode_pressure_in = 1.0;
  p.kin_frac = F(0.1);
}

void params_init(e2d_params &p) {
  // HydroParams::init, src/HydroParams.cpp:160-176
  p.imax = p.nx - 1 + 2 * p.ghostWidth;
  p.jmax = p.ny - 1 + 2 * p.ghostWidth;
  p.isize = p.imax - p.imin + 1;
  p.jsize = p.jmax - p.jmin + 1;
  p.dx = (p.xmax - p.xmin) / p.nx;
  p.dy = (p.ymax - p.ymin) / p.ny;
  p.smallp = p.smallc * p.smallc / p.gamma0;
  p.smallpp = p.smallr * p.smallp;
  p.gamma6 = (p.gamma0 + 1.0) / (2.0 * p.gamma0);
}

void params_setup(const ConfigMap &cfg, e2d_params &p) {
  params_defaults(p);
  // [run]  HydroParams.cpp:47-59
  p.nStepmax = static_cast<int32_t>(cfg.getInteger("run", "nstepmax", 1000));
  p.tEnd = cfg.getFloat("run", "tend", F(0.0));
  p.nOutput = static_cast<int32_t>(cfg.getInteger("run", "noutput", 100));
  if (p.nOutput == -1) p.enableOutput = 0;
  p.ioHDF5 = cfg.getBool("run", "use_HDF5", false) ? 1 : 0;
  p.ioVTK = cfg.getBool("run", "use_VTK", true) ? 1 : 0;

  // [mesh]  :62-78
  p.nx = static_cast<int32_t>(cfg.getInteger("mesh", "nx", 2));
  p.ny = static_cast<int32_t>(cfg.getInteger("mesh", "ny", 2));
  p.xmin = cfg.getFloat("mesh", "xmin", F(0.0));
  p.ymin = cfg.getFloat("mesh", "ymin", F(0.0));
  p.xmax = cfg.getFloat("mesh", "xmax", F(1.0));
  p.ymax = cfg.getFloat("mesh", "ymax", F(1.0));
  p.boundary_type_xmin = static_cast<int32_t>(cfg.getInteger("mesh", "boundary_type_xmin", 1));
  p.boundary_type_xmax = static_cast<int32_t>(cfg.getInteger("mesh", "boundary_type_xmax", 1));
  p.boundary_type_ymin = static_cast<int32_t>(cfg.getInteger("mesh", "boundary_type_ymin", 1));
  p.boundary_type_ymax = static_cast<int32_t>(cfg.getInteger("mesh", "boundary_type_ymax", 1));

  // [hydro]  :80-134
  p.gamma0 = cfg.getFloat("hydro", "gamma0", F(1.4));
  p.cfl = cfg.getFloat("hydro", "cfl", F(0.5));
  p.slope_type = cfg.getFloat("hydro", "slope_type", F(1.0));
  p.smallc = cfg.getFloat("hydro", "smallc", F(1e-10));
  p.smallr = cfg.getFloat("hydro", "smallr", F(1e-10));
  p.niter_riemann = static_cast<int32_t>(cfg.getInteger("hydro", "niter_riemann", 10));

  const std::string riemann = cfg.getString("hydro", "riemann", "approx");
  if (riemann == "hll")
    p.riemannSolverType = 1;
  else if (riemann == "hllc")
    p.riemannSolverType = 2;
  else
    p.riemannSolverType = 0;  // "approx", and the fall-back of :100-105

  const std::string problem = cfg.getString("hydro", "problem", "unknown");
  if (problem == "blast")
    p.problemType = 1;
  else if (problem == "four_quadrant")
    p.problemType = 2;
  else if (problem == "discontinuity")
    p.problemType = 3;
  else if (problem == "shocked_bubble")
    p.problemType = 4;
  else
    p.problemType = 0;  // "implode", and the fall-back of :129-134

  // readShockedBubbleParams, :19-37 (read for every problem; only used by one)
  p.sb_bubble_radius = cfg.getFloat("shocked_bubble", "bubble_radius", F(0.025));
  p.sb_bubble_center_x = cfg.getFloat("shocked_bubble", "bubble_center_x", F(0.225));
  p.sb_bubble_center_y = cfg.getFloat("shocked_bubble", "bubble_center_y", F(0.0445));
  p.sb_bubble_density = cfg.getFloat("shocked_bubble", "bubble_density", F(3.863));
  p.sb_bubble_pressure = cfg.getFloat("shocked_bubble", "bubble_pressure", F(1.0132e5));
  p.sb_preshock_density = cfg.getFloat("shocked_bubble", "preshock_density", F(1.225));
  p.sb_preshock_pressure = cfg.getFloat("shocked_bubble", "preshock_pressure", F(1.0132e5));
  p.sb_postshock_density = cfg.getFloat("shocked_bubble", "postshock_density", F(1.686));
  p.sb_postshock_pressure = cfg.getFloat("shocked_bubble", "postshock_pressure", F(1.59e5));
  p.sb_postshock_velocity = cfg.getFloat("shocked_bubble", "postshock_velocity", F(113.5));
  p.sb_shock_loc = cfg.getFloat("shocked_bubble", "shock_loc", F(0.170));

  // [blast]  :136-144; defaults are computed in double, then rounded to float
  // by the float-typed default argument
  p.blast_radius =
    cfg.getFloat("blast", "radius", static_cast<float>((p.xmin + p.xmax) / 2.0 / 10));
  p.blast_center_x = cfg.getFloat("blast", "center_x", static_cast<float>((p.xmin + p.xmax) / 2));
  p.blast_center_y = cfg.getFloat("blast", "center_y", static_cast<float>((p.ymin + p.ymax) / 2));
  p.blast_density_in = cfg.getFloat("blast", "density_in", F(1.0));
  p.blast_density_out = cfg.getFloat("blast", "density_out", F(1.2));
  p.blast_pressure_in = cfg.getFloat("blast", "pressure_in", F(10.0));
  p.blast_pressure_out = cfg.getFloat("blast", "pressure_out", F(0.1));
  p.blast_total_energy_inside = cfg.getFloat("blast", "total_energy_inside", F(0.0));
  p.blast_nbins = static_cast<int32_t>(cfg.getInteger("blast", "nbins", 100));

  // [other]  :146-152 (float read, truncated to int)
  p.implementationVersion = static_cast<int32_t>(cfg.getFloat("OTHER", "implementationVersion", 0));
  if (p.implementationVersion != 0 && p.implementationVersion != 1 && p.implementationVersion != 2)
    p.implementationVersion = 0;

  // extension keys; their defaults reproduce the reference
  p.honor_riemann = cfg.getBool("hydro", "honor_riemann", false) ? 1 : 0;
  p.implode_pressure_in = cfg.getFloat("implode", "pressure_in", F(1.0));
  p.kin_frac = cfg.getFloat("hydro", "kin_frac", F(0.1));

  params_init(p);
}

void params_print(const e2d_params &p) {
  // HydroParams::print, src/HydroParams.cpp:195-224
  std::printf("##########################\n");
  std::printf("Simulation run parameters:\n");
  std::printf("##########################\n");
  std::printf("nx         : %d\n", p.nx);
  std::printf("ny         : %d\n", p.ny);
  std::printf("dx         : %f\n", p.dx);
  std::printf("dy         : %f\n", p.dy);
  std::printf("imin       : %d\n", p.imin);
  std::printf("imax       : %d\n", p.imax);
  std::printf("jmin       : %d\n", p.jmin);
  std::printf("jmax       : %d\n", p.jmax);
  std::printf("nStepmax   : %d\n", p.nStepmax);
  std::printf("tEnd       : %f\n", p.tEnd);
  std::printf("nOutput    : %d\n", p.nOutput);
  std::printf("gamma0     : %f\n", p.gamma0);
  std::printf("cfl        : %f\n", p.cfl);
  std::printf("smallr     : %12.10f\n", p.smallr);
  std::printf("smallc     : %12.10f\n", p.smallc);
  std::printf("slope_type : %f\n", p.slope_type);
  std::printf("riemann    : %d\n", p.riemannSolverType);
  std::printf("problem    : %d\n", p.problemType);
  std::printf("implementation version : %d\n", p.implementationVersion);
  std::printf("##########################\n");
  std::fflush(stdout);
}

const char *params_check(const e2d_params &p) {
  if (p.ghostWidth != 2) return "ghostWidth must be 2";
  if (p.nx < 2 || p.ny < 2) return "nx and ny must be >= 2 (ghost maps read interior cells only)";
  if (p.isize != p.nx + 4 || p.jsize != p.ny + 4 || p.imin != 0 || p.jmin != 0)
    return "derived sizes inconsistent: call e2d_params_init";
  if (!(p.dx > 0.0) || !(p.dy > 0.0)) return "dx and dy must be positive (xmax > xmin, ymax > ymin)";
  if (!(p.gamma0 > 1.0)) return "gamma0 must be > 1";
  if (!(p.slope_type == 0.0 || p.slope_type == 1.0 || p.slope_type == 2.0))
    return "slope_type must be 0, 1 or 2 (other values leave the reference's slopes unset)";
  if (p.riemannSolverType < 0 || p.riemannSolverType > 2) return "riemannSolverType must be 0, 1 or 2";
  if (!(p.cfl > 0.0)) return "cfl must be positive";
  if (!(p.kin_frac >= 0.0 && p.kin_frac <= 1.0e30))
    return "kin_frac must be a finite number >= 0";
  if (static_cast<long long>(p.isize) * p.jsize > (1LL << 31)) return "grid too large";
  return nullptr;
}

}  // namespace e2d
