/*
 * euler2d_oracle.h -- CPU oracle for the unsplit MUSCL-Hancock Godunov step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (pkestene/euler2d_kokkos, implementation version 0) used as the
 * checker for the HIP product path.  Only tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py may load it.  The product library never
 * links, loads or calls anything in this directory.
 *
 * Layout: the reference's OpenMP layout, LayoutRight AoS [i][j][var]
 * (src/kokkos_shared.h:20-21), element (i,j,v) at ((i*jsize)+j)*4+v.
 * Variable order ID=0 (rho), IP=1 (E / p), IU=2, IV=3 (src/HydroParams.h:27-34).
 *
 * Pinning status of each function is listed in the header of
 * euler2d_oracle.c.
 */
#ifndef EULER2D_ORACLE_H_
#define EULER2D_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Identical layout to e2d_params in include/euler2d_hip.h (checked by a test). */
typedef struct ora_params {
  /* [run] */
  int32_t nStepmax;
  int32_t nOutput;
  int32_t enableOutput;
  int32_t ioVTK;
  int32_t ioHDF5;
  int32_t ghostWidth;
  double tEnd;
  /* [mesh] */
  int32_t nx, ny;
  int32_t imin, imax, jmin, jmax;
  int32_t isize, jsize;
  int32_t boundary_type_xmin, boundary_type_xmax;
  int32_t boundary_type_ymin, boundary_type_ymax;
  double xmin, xmax, ymin, ymax, dx, dy;
  /* [hydro] (HydroSettings) */
  double gamma0, gamma6, cfl, slope_type, smallr, smallc, smallp, smallpp;
  int32_t niter_riemann;
  int32_t riemannSolverType; /* 0 approx, 1 hll, 2 hllc */
  int32_t problemType;       /* 0 implode 1 blast 2 four_quadrant 3 discontinuity 4 shocked_bubble */
  int32_t implementationVersion;
  /* extension: 0 = reference behaviour (HLLC always), 1 = honour riemannSolverType */
  int32_t honor_riemann;
  int32_t blast_nbins;
  /* [blast] */
  double blast_radius, blast_center_x, blast_center_y;
  double blast_density_in, blast_density_out;
  double blast_pressure_in, blast_pressure_out;
  double blast_total_energy_inside;
  /* [shocked_bubble] */
  double sb_bubble_radius, sb_bubble_center_x, sb_bubble_center_y;
  double sb_bubble_density, sb_bubble_pressure;
  double sb_preshock_density, sb_preshock_pressure;
  double sb_postshock_density, sb_postshock_pressure, sb_postshock_velocity;
  double sb_shock_loc;
  /* extension: [implode] pressure_in, high-state pressure (reference hard-codes 1) */
  double implode_pressure_in;
  /* extension: [hydro] kin_frac, split-precision threshold of the product (unused here) */
  double kin_frac;
} ora_params;

int ora_params_size(void);

/* ---- configuration (ConfigMap semantics) ---- */
void *ora_config_open(const char *path, int *parse_error);
void ora_config_close(void *cfg);
int ora_config_get_string(void *cfg, const char *section, const char *name,
                          const char *default_value, char *out, int out_len);
long ora_config_get_integer(void *cfg, const char *section, const char *name, long default_value);
float ora_config_get_float(void *cfg, const char *section, const char *name, float default_value);
int ora_config_get_bool(void *cfg, const char *section, const char *name, int default_value);
void ora_params_defaults(ora_params *p);
void ora_params_setup(void *cfg, ora_params *p);
void ora_params_init(ora_params *p);

/* ---- pointwise numerics (HydroBaseFunctor) ---- */
void ora_compute_primitives(const ora_params *p, const double u[4], double *c, double q[4]);
void ora_slope_unsplit_hydro_2d(const ora_params *p, const double q[4], const double qpx[4],
                                const double qmx[4], const double qpy[4], const double qmy[4],
                                double dqx[4], double dqy[4]);
void ora_trace_unsplit_2d_along_dir(const ora_params *p, const double q[4], const double dqx[4],
                                    const double dqy[4], double dtdx, double dtdy, int face,
                                    double qface[4]);
void ora_riemann_hllc(const ora_params *p, const double ql[4], const double qr[4], double flux[4]);
void ora_riemann_hll(const ora_params *p, const double ql[4], const double qr[4], double flux[4]);
void ora_riemann_approx(const ora_params *p, const double ql[4], const double qr[4],
                        double qgdnv[4], double flux[4]);

/* ---- the run (HydroRun, implementation version 0) ---- */
typedef struct ora_run ora_run;
int ora_create(const ora_params *p, ora_run **out);
void ora_destroy(ora_run *r);
void ora_set_threads(int n);
int ora_get_threads(void);
void ora_init_problem(ora_run *r);
double ora_compute_dt(ora_run *r, int useU);
void ora_make_boundaries(ora_run *r, int which);
void ora_convert_to_primitives(ora_run *r, int which);
void ora_compute_and_store_fluxes(ora_run *r, double dtdx, double dtdy);
void ora_update(ora_run *r, int which);
void ora_godunov_unsplit(ora_run *r, int nStep, double dt);
/* main.cpp loop; runs at most max_steps iterations of the while loop (<=0: no cap
 * besides nStepmax/tEnd).  t and nStep carry the loop state in and out. */
void ora_run_loop(ora_run *r, double *t, int *nStep, int max_steps);
/* which: 0 U, 1 U2, 2 Q, 3 Fluxes_x, 4 Fluxes_y; full isize*jsize*4 AoS arrays */
int ora_get_array(ora_run *r, int which, double *out);
int ora_set_array(ora_run *r, int which, const double *in);

/* output formats */
/* HydroRun::saveVTK (HydroRun.h:502-609) of array `which` (0 U, 1 U2) */
int ora_save_vtk(ora_run *r, int which, int use_float32, const char *output_dir,
                 const char *output_prefix, int step);
/* ComputeRadialProfileFunctor::apply (ComputeRadialProfileFunctor.h:83-165)
 * of array `which`, sequential sums; cells past the last bin are skipped */
int ora_radial_profile(ora_run *r, int which, int nbins, double *distances, double *density,
                       int64_t *counts);
/* cnpy::npy_save of a 1-D double array (cnpy.h:129-187, 343-376) */
int ora_write_npy(const char *path, const double *data, int64_t n);

#ifdef __cplusplus
}
#endif

#endif
