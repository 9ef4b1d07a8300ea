/*
 * euler2d_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY, never shipped).
 *
 * Plain-C restatement of pkestene/euler2d_kokkos, implementation version 0,
 * double precision (USE_DOUBLE).  It keeps the reference's pass structure
 * (boundaries -> copy -> primitives -> store fluxes -> update, separate dt
 * reduction), its OpenMP memory layout (LayoutRight AoS [i][j][var]) and its
 * left-to-right expression order.  Build with -ffp-contract=off and without
 * -ffast-math so that every operation is a single correctly rounded IEEE op.
 *
 * Who may use it: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg.  The product library (euler2d_kokkos_amd/) never links or loads it.
 *
 * How each part is pinned (tests named in brackets):
 *   INI / ConfigMap parsing  -> the reference's own config/ sources compiled
 *                               into oracle/_ref [test_config_parity.py]
 *   HydroParams setup/init   -> float-rounding and derived-value closed forms
 *                               [test_oracle_params.py]; smallp by the floor
 *                               sound speed, smallpp by the empty-cell
 *                               pressure floor
 *   computePrimitives        -> exact inverse of the textbook prim->cons map
 *                               [test_oracle_pointwise.py]
 *   slope limiter            -> minmod / MC definitions on monotone, extremum
 *                               and flat data [test_oracle_pointwise.py]
 *   Hancock trace            -> a linear (smooth) state: the predictor equals a
 *                               half-step Taylor expansion of the Euler
 *                               equations [test_oracle_pointwise.py]
 *   riemann_hllc             -> Toro (2009) 10.4 HLLC written independently
 *                               from the book [test_oracle_pointwise.py]
 *   riemann_hll              -> Toro 10.3, two algebraically different forms
 *                               of the HLL flux [test_oracle_pointwise.py]
 *   riemann_approx           -> exact two-shock Riemann solution (Toro 4.2)
 *                               on shock-shock data [test_oracle_pointwise.py]
 *   make_boundaries          -> closed-form ghost maps brute-forced on small
 *                               grids [test_oracle_step.py]
 *   whole step / loop        -> conservation (reflecting / periodic walls),
 *                               uniform state, x<->y transpose and mirror
 *                               symmetry, Sod tube convergence to the exact
 *                               solution, dt closed form [test_oracle_step.py]
 *   Init* functors           -> closed-form states and region tests
 *                               [test_oracle_step.py]
 *   saveVTK                  -> XML structure parsed independently, values
 *                               read back at 6 significant digits, file name
 *                               rule [test_output_formats.py]
 *   npy writer               -> numpy's own loader [test_output_formats.py]
 *   radial profile           -> brute-force numpy binning and a uniform field
 *                               whose profile is the field value
 *                               [test_output_formats.py]
 */
#include "euler2d_oracle.h"

#include <ctype.h>
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { ID = 0, IP = 1, IU = 2, IV = 3 };
enum { FACE_XMIN = 0, FACE_XMAX = 1, FACE_YMIN = 2, FACE_YMAX = 3 };
enum { BC_DIRICHLET = 1, BC_NEUMANN = 2 };
enum { RIEMANN_APPROX = 0, RIEMANN_HLL = 1, RIEMANN_HLLC = 2 };

int ora_params_size(void) { return (int)sizeof(ora_params); }

/* ======================================================================
 * INI file -> key/value map.
 * Semantics of config/inih/ini.cpp:16-152 (INI_ALLOW_MULTILINE=1,
 * ini.h:37-39) and config/inih/INIReader.cpp:36-100.
 * ====================================================================== */
typedef struct {
  char *key;
  char *value;
} ora_kv;

typedef struct {
  ora_kv *items;
  int n, cap;
  int error;
} ora_cfg;

static char *dupstr(const char *s) {
  size_t n = strlen(s) + 1;
  char *d = (char *)malloc(n);
  memcpy(d, s, n);
  return d;
}

/* lower-cased "section.name" (INIReader.cpp:80-88) */
static char *make_key(const char *section, const char *name) {
  size_t ls = strlen(section), ln = strlen(name);
  char *k = (char *)malloc(ls + ln + 2);
  for (size_t i = 0; i < ls; i++) k[i] = (char)tolower((unsigned char)section[i]);
  k[ls] = '.';
  for (size_t i = 0; i < ln; i++) k[ls + 1 + i] = (char)tolower((unsigned char)name[i]);
  k[ls + ln + 1] = '\0';
  return k;
}

/* later assignments to a key replace earlier ones (INIReader.cpp:96-99) */
static void cfg_set(ora_cfg *c, const char *section, const char *name, const char *value) {
  char *k = make_key(section, name);
  for (int i = 0; i < c->n; i++) {
    if (strcmp(c->items[i].key, k) == 0) {
      free(c->items[i].value);
      c->items[i].value = dupstr(value);
      free(k);
      return;
    }
  }
  if (c->n == c->cap) {
    c->cap = c->cap ? 2 * c->cap : 32;
    c->items = (ora_kv *)realloc(c->items, (size_t)c->cap * sizeof(ora_kv));
  }
  c->items[c->n].key = k;
  c->items[c->n].value = dupstr(value);
  c->n++;
}

static const char *cfg_lookup(const ora_cfg *c, const char *section, const char *name) {
  char *k = make_key(section, name);
  const char *v = NULL;
  for (int i = 0; i < c->n; i++)
    if (strcmp(c->items[i].key, k) == 0) v = c->items[i].value;
  free(k);
  return v;
}

/* trailing whitespace removed in place */
static char *trim_right(char *s) {
  char *e = s + strlen(s);
  while (e > s && isspace((unsigned char)e[-1])) *--e = '\0';
  return s;
}

static char *skip_left(char *s) {
  while (*s && isspace((unsigned char)*s)) s++;
  return s;
}

/* first `stop` char, or a ';' that follows whitespace, or the terminator */
static char *scan_to(char *s, char stop) {
  int after_space = 0;
  for (; *s; s++) {
    if (*s == stop) break;
    if (after_space && *s == ';') break;
    after_space = isspace((unsigned char)*s) ? 1 : 0;
  }
  return s;
}

static void copy_bounded(char *dst, const char *src, size_t cap) {
  strncpy(dst, src, cap);
  dst[cap - 1] = '\0';
}

/* Reads `path` line by line into `c`.  Returns 0, the first bad line number,
 * or -1 when the file cannot be opened.  Lines are read in chunks of at most
 * 199 characters, sections and continuation names are clipped to 49. */
static int cfg_parse(ora_cfg *c, const char *path) {
  char buf[200];
  char section[50] = "";
  char last_name[50] = "";
  int lineno = 0, first_error = 0;
  FILE *f = fopen(path, "r");
  if (!f) return -1;
  while (fgets(buf, (int)sizeof(buf), f)) {
    lineno++;
    char *s = skip_left(trim_right(buf));
    if (last_name[0] && *s && s > buf) {
      /* indented non-empty line after a name: continuation, replaces value */
      cfg_set(c, section, last_name, s);
      continue;
    }
    if (*s == ';' || *s == '#') continue;
    if (*s == '[') {
      char *close = scan_to(s + 1, ']');
      if (*close == ']') {
        *close = '\0';
        copy_bounded(section, s + 1, sizeof(section));
        last_name[0] = '\0';
      } else if (!first_error) {
        first_error = lineno;
      }
      continue;
    }
    if (!*s) continue;
    char *eq = scan_to(s, '=');
    if (*eq != '=') {
      if (!first_error) first_error = lineno;
      continue;
    }
    *eq = '\0';
    char *name = trim_right(s);
    char *val = skip_left(eq + 1);
    char *cm = scan_to(val, '\0');
    if (*cm == ';') *cm = '\0';
    trim_right(val);
    copy_bounded(last_name, name, sizeof(last_name));
    cfg_set(c, section, name, val);
  }
  fclose(f);
  return first_error;
}

void *ora_config_open(const char *path, int *parse_error) {
  ora_cfg *c = (ora_cfg *)calloc(1, sizeof(ora_cfg));
  c->error = cfg_parse(c, path);
  if (parse_error) *parse_error = c->error;
  return c;
}

void ora_config_close(void *cfg) {
  ora_cfg *c = (ora_cfg *)cfg;
  if (!c) return;
  for (int i = 0; i < c->n; i++) {
    free(c->items[i].key);
    free(c->items[i].value);
  }
  free(c->items);
  free(c);
}

int ora_config_get_string(void *cfg, const char *section, const char *name,
                          const char *default_value, char *out, int out_len) {
  const char *v = cfg_lookup((ora_cfg *)cfg, section, name);
  if (!v) v = default_value;
  int n = (int)strlen(v);
  if (out && out_len > 0) {
    strncpy(out, v, (size_t)out_len);
    out[out_len - 1] = '\0';
  }
  return n;
}

/* INIReader::getInteger, INIReader.cpp:54-64: strtol base 0 */
long ora_config_get_integer(void *cfg, const char *section, const char *name, long default_value) {
  const char *v = cfg_lookup((ora_cfg *)cfg, section, name);
  if (!v) v = "";
  char *end;
  long n = strtol(v, &end, 0);
  return end > v ? n : default_value;
}

/* ConfigMap::getFloat, config/ConfigMap.cpp:31-40: strtof, i.e. float rounding */
float ora_config_get_float(void *cfg, const char *section, const char *name, float default_value) {
  const char *v = cfg_lookup((ora_cfg *)cfg, section, name);
  if (!v) v = "";
  char *end;
  float f = strtof(v, &end);
  return end > v ? f : default_value;
}

/* ConfigMap::getBool, config/ConfigMap.cpp:57-76 */
int ora_config_get_bool(void *cfg, const char *section, const char *name, int default_value) {
  const char *v = cfg_lookup((ora_cfg *)cfg, section, name);
  if (!v) v = "";
  int val = default_value;
  if (!strcmp(v, "1") || !strcmp(v, "yes") || !strcmp(v, "true") || !strcmp(v, "on")) val = 1;
  if (!strcmp(v, "0") || !strcmp(v, "no") || !strcmp(v, "false") || !strcmp(v, "off")) val = 0;
  if (!v[0]) val = default_value;
  return val;
}

/* ======================================================================
 * HydroParams (src/HydroParams.h:155-265, src/HydroParams.cpp:19-190)
 * ====================================================================== */
void ora_params_defaults(ora_params *p) {
  /* HydroParams() constructor, HydroParams.h:227-266 and HydroSettings() :131-140 */
  memset(p, 0, sizeof(*p));
  p->enableOutput = 1;
  p->ghostWidth = 2;
  p->xmax = 1.0;
  p->ymax = 1.0;
  p->boundary_type_xmin = p->boundary_type_xmax = 1;
  p->boundary_type_ymin = p->boundary_type_ymax = 1;
  p->ioVTK = 1;
  p->gamma0 = 1.4;
  p->gamma6 = 1.0;
  p->cfl = 1.0;
  p->slope_type = 2.0;
  p->smallr = 1e-8;
  p->smallc = 1e-8;
  p->smallp = 1e-6;
  p->smallpp = 1e-6;
  p->niter_riemann = 10;
  p->blast_radius = 10.0;
  p->blast_density_in = 1.0;
  p->blast_density_out = 1.2;
  p->blast_pressure_in = 10.0;
  p->blast_pressure_out = 0.1;
  p->blast_nbins = 100;
  p->implode_pressure_in = 1.0;
  p->kin_frac = 0.1f;
}

/* HydroParams::init, HydroParams.cpp:160-190 */
void ora_params_init(ora_params *p) {
  p->imax = p->nx - 1 + 2 * p->ghostWidth;
  p->jmax = p->ny - 1 + 2 * p->ghostWidth;
  p->isize = p->imax - p->imin + 1;
  p->jsize = p->jmax - p->jmin + 1;
  p->dx = (p->xmax - p->xmin) / p->nx;
  p->dy = (p->ymax - p->ymin) / p->ny;
  p->smallp = p->smallc * p->smallc / p->gamma0;
  p->smallpp = p->smallr * p->smallp;
  p->gamma6 = (p->gamma0 + 1.0) / (2.0 * p->gamma0);
}

/* HydroParams::setup, HydroParams.cpp:42-156; readShockedBubbleParams :19-37 */
void ora_params_setup(void *cfg, ora_params *p) {
  char s[256];
  ora_params_defaults(p);
  p->nStepmax = (int32_t)ora_config_get_integer(cfg, "run", "nstepmax", 1000);
  p->tEnd = ora_config_get_float(cfg, "run", "tend", 0.0f);
  p->nOutput = (int32_t)ora_config_get_integer(cfg, "run", "noutput", 100);
  if (p->nOutput == -1) p->enableOutput = 0;
  p->ioHDF5 = ora_config_get_bool(cfg, "run", "use_HDF5", 0);
  p->ioVTK = ora_config_get_bool(cfg, "run", "use_VTK", 1);

  p->nx = (int32_t)ora_config_get_integer(cfg, "mesh", "nx", 2);
  p->ny = (int32_t)ora_config_get_integer(cfg, "mesh", "ny", 2);
  p->xmin = ora_config_get_float(cfg, "mesh", "xmin", 0.0f);
  p->ymin = ora_config_get_float(cfg, "mesh", "ymin", 0.0f);
  p->xmax = ora_config_get_float(cfg, "mesh", "xmax", 1.0f);
  p->ymax = ora_config_get_float(cfg, "mesh", "ymax", 1.0f);
  p->boundary_type_xmin = (int32_t)ora_config_get_integer(cfg, "mesh", "boundary_type_xmin", 1);
  p->boundary_type_xmax = (int32_t)ora_config_get_integer(cfg, "mesh", "boundary_type_xmax", 1);
  p->boundary_type_ymin = (int32_t)ora_config_get_integer(cfg, "mesh", "boundary_type_ymin", 1);
  p->boundary_type_ymax = (int32_t)ora_config_get_integer(cfg, "mesh", "boundary_type_ymax", 1);

  p->gamma0 = ora_config_get_float(cfg, "hydro", "gamma0", (float)1.4);
  p->cfl = ora_config_get_float(cfg, "hydro", "cfl", (float)0.5);
  p->slope_type = ora_config_get_float(cfg, "hydro", "slope_type", 1.0f);
  p->smallc = ora_config_get_float(cfg, "hydro", "smallc", (float)1e-10);
  p->smallr = ora_config_get_float(cfg, "hydro", "smallr", (float)1e-10);
  p->niter_riemann = (int32_t)ora_config_get_integer(cfg, "hydro", "niter_riemann", 10);

  ora_config_get_string(cfg, "hydro", "riemann", "approx", s, (int)sizeof(s));
  if (!strcmp(s, "approx")) p->riemannSolverType = RIEMANN_APPROX;
  else if (!strcmp(s, "hll")) p->riemannSolverType = RIEMANN_HLL;
  else if (!strcmp(s, "hllc")) p->riemannSolverType = RIEMANN_HLLC;
  else p->riemannSolverType = RIEMANN_APPROX;

  ora_config_get_string(cfg, "hydro", "problem", "unknown", s, (int)sizeof(s));
  if (!strcmp(s, "implode")) p->problemType = 0;
  else if (!strcmp(s, "blast")) p->problemType = 1;
  else if (!strcmp(s, "four_quadrant")) p->problemType = 2;
  else if (!strcmp(s, "discontinuity")) p->problemType = 3;
  else if (!strcmp(s, "shocked_bubble")) p->problemType = 4;
  else p->problemType = 0;

  /* readShockedBubbleParams defaults (read unconditionally here; the reference
   * leaves them unset for other problems, where they are never used) */
  p->sb_bubble_radius = ora_config_get_float(cfg, "shocked_bubble", "bubble_radius", (float)0.025);
  p->sb_bubble_center_x = ora_config_get_float(cfg, "shocked_bubble", "bubble_center_x", (float)0.225);
  p->sb_bubble_center_y = ora_config_get_float(cfg, "shocked_bubble", "bubble_center_y", (float)0.0445);
  p->sb_bubble_density = ora_config_get_float(cfg, "shocked_bubble", "bubble_density", (float)3.863);
  p->sb_bubble_pressure = ora_config_get_float(cfg, "shocked_bubble", "bubble_pressure", (float)1.0132e5);
  p->sb_preshock_density = ora_config_get_float(cfg, "shocked_bubble", "preshock_density", (float)1.225);
  p->sb_preshock_pressure = ora_config_get_float(cfg, "shocked_bubble", "preshock_pressure", (float)1.0132e5);
  p->sb_postshock_density = ora_config_get_float(cfg, "shocked_bubble", "postshock_density", (float)1.686);
  p->sb_postshock_pressure = ora_config_get_float(cfg, "shocked_bubble", "postshock_pressure", (float)1.59e5);
  p->sb_postshock_velocity = ora_config_get_float(cfg, "shocked_bubble", "postshock_velocity", (float)113.5);
  p->sb_shock_loc = ora_config_get_float(cfg, "shocked_bubble", "shock_loc", (float)0.170);

  /* defaults computed in double from the float-widened bounds, then passed
   * through the float-typed default argument */
  p->blast_radius = ora_config_get_float(cfg, "blast", "radius", (float)((p->xmin + p->xmax) / 2.0 / 10));
  p->blast_center_x = ora_config_get_float(cfg, "blast", "center_x", (float)((p->xmin + p->xmax) / 2));
  p->blast_center_y = ora_config_get_float(cfg, "blast", "center_y", (float)((p->ymin + p->ymax) / 2));
  p->blast_density_in = ora_config_get_float(cfg, "blast", "density_in", 1.0f);
  p->blast_density_out = ora_config_get_float(cfg, "blast", "density_out", (float)1.2);
  p->blast_pressure_in = ora_config_get_float(cfg, "blast", "pressure_in", 10.0f);
  p->blast_pressure_out = ora_config_get_float(cfg, "blast", "pressure_out", (float)0.1);
  p->blast_total_energy_inside = ora_config_get_float(cfg, "blast", "total_energy_inside", 0.0f);
  p->blast_nbins = (int32_t)ora_config_get_integer(cfg, "blast", "nbins", 100);

  /* int <- float conversion truncates, as in HydroParams.cpp:146 */
  p->implementationVersion = (int32_t)ora_config_get_float(cfg, "OTHER", "implementationVersion", 0.0f);
  if (p->implementationVersion != 0 && p->implementationVersion != 1 && p->implementationVersion != 2)
    p->implementationVersion = 0;

  /* extensions (keys unknown to the reference; defaults keep its behaviour) */
  p->honor_riemann = ora_config_get_bool(cfg, "hydro", "honor_riemann", 0);
  p->implode_pressure_in = ora_config_get_float(cfg, "implode", "pressure_in", 1.0f);
  p->kin_frac = ora_config_get_float(cfg, "hydro", "kin_frac", 0.1f);

  ora_params_init(p);
}

/* ======================================================================
 * Pointwise numerics (src/HydroBaseFunctor.h)
 * ====================================================================== */

/* computePrimitives, HydroBaseFunctor.h:76-102 */
void ora_compute_primitives(const ora_params *P, const double u[4], double *c, double q[4]) {
  const double gamma0 = P->gamma0, smallr = P->smallr, smallp = P->smallp;
  double d = fmax(u[ID], smallr);
  double ux = u[IU] / d;
  double uy = u[IV] / d;
  double eken = 0.5 * (ux * ux + uy * uy);
  double e = u[IP] / d - eken;
  double p = fmax((gamma0 - 1.0) * d * e, d * smallp);
  *c = sqrt(gamma0 * (p) / d);
  q[ID] = d;
  q[IP] = p;
  q[IU] = ux;
  q[IV] = uy;
}

/* slope_unsplit_hydro_2d_scalar, HydroBaseFunctor.h:419-455 (one direction) */
static double limited_slope(double slope_type, double q, double qplus, double qminus) {
  double dlft = slope_type * (q - qminus);
  double drgt = slope_type * (qplus - q);
  double dcen = 0.5 * (qplus - qminus);
  double dsgn = (dcen >= 0.0) ? 1.0 : -1.0;
  double slop = fmin(fabs(dlft), fabs(drgt));
  double dlim = slop;
  if ((dlft * drgt) <= 0.0) dlim = 0.0;
  return dsgn * fmin(dlim, fabs(dcen));
}

/* slope_unsplit_hydro_2d, HydroBaseFunctor.h:473-516.  slope_type values
 * other than 0, 1, 2 leave the reference's slopes unset; the oracle returns
 * zero for them and the product rejects them at creation. */
void ora_slope_unsplit_hydro_2d(const ora_params *P, const double q[4], const double qpx[4],
                                const double qmx[4], const double qpy[4], const double qmy[4],
                                double dqx[4], double dqy[4]) {
  const double st = P->slope_type;
  if (st == 1.0 || st == 2.0) {
    for (int v = 0; v < 4; v++) {
      dqx[v] = limited_slope(st, q[v], qpx[v], qmx[v]);
      dqy[v] = limited_slope(st, q[v], qpy[v], qmy[v]);
    }
  } else {
    for (int v = 0; v < 4; v++) dqx[v] = dqy[v] = 0.0;
  }
}

/* trace_unsplit_2d_along_dir, HydroBaseFunctor.h:214-291 */
void ora_trace_unsplit_2d_along_dir(const ora_params *P, const double q[4], const double dqX[4],
                                    const double dqY[4], double dtdx, double dtdy, int face,
                                    double qface[4]) {
  const double gamma0 = P->gamma0, smallr = P->smallr;
  double r = q[ID], p = q[IP], u = q[IU], v = q[IV];
  double drx = dqX[ID], dpx = dqX[IP], dux = dqX[IU], dvx = dqX[IV];
  double dry = dqY[ID], dpy = dqY[IP], duy = dqY[IU], dvy = dqY[IV];
  double sr0 = -u * drx - v * dry - (dux + dvy) * r;
  double sp0 = -u * dpx - v * dpy - (dux + dvy) * gamma0 * p;
  double su0 = -u * dux - v * duy - (dpx) / r;
  double sv0 = -u * dvx - v * dvy - (dpy) / r;
  if (face == FACE_XMIN) {
    qface[ID] = r - 0.5 * drx + sr0 * dtdx * 0.5;
    qface[IP] = p - 0.5 * dpx + sp0 * dtdx * 0.5;
    qface[IU] = u - 0.5 * dux + su0 * dtdx * 0.5;
    qface[IV] = v - 0.5 * dvx + sv0 * dtdx * 0.5;
    qface[ID] = fmax(smallr, qface[ID]);
  }
  if (face == FACE_XMAX) {
    qface[ID] = r + 0.5 * drx + sr0 * dtdx * 0.5;
    qface[IP] = p + 0.5 * dpx + sp0 * dtdx * 0.5;
    qface[IU] = u + 0.5 * dux + su0 * dtdx * 0.5;
    qface[IV] = v + 0.5 * dvx + sv0 * dtdx * 0.5;
    qface[ID] = fmax(smallr, qface[ID]);
  }
  if (face == FACE_YMIN) {
    qface[ID] = r - 0.5 * dry + sr0 * dtdy * 0.5;
    qface[IP] = p - 0.5 * dpy + sp0 * dtdy * 0.5;
    qface[IU] = u - 0.5 * duy + su0 * dtdy * 0.5;
    qface[IV] = v - 0.5 * dvy + sv0 * dtdy * 0.5;
    qface[ID] = fmax(smallr, qface[ID]);
  }
  if (face == FACE_YMAX) {
    qface[ID] = r + 0.5 * dry + sr0 * dtdy * 0.5;
    qface[IP] = p + 0.5 * dpy + sp0 * dtdy * 0.5;
    qface[IU] = u + 0.5 * duy + su0 * dtdy * 0.5;
    qface[IV] = v + 0.5 * dvy + sv0 * dtdy * 0.5;
    qface[ID] = fmax(smallr, qface[ID]);
  }
}

/* riemann_hllc, HydroBaseFunctor.h:704-809 */
void ora_riemann_hllc(const ora_params *P, const double qleft[4], const double qright[4],
                      double flux[4]) {
  const double gamma0 = P->gamma0, smallr = P->smallr, smallp = P->smallp, smallc = P->smallc;
  const double entho = 1.0 / (gamma0 - 1.0);

  double rl = fmax(qleft[ID], smallr);
  double pl = fmax(qleft[IP], rl * smallp);
  double ul = qleft[IU];
  double ecinl = 0.5 * rl * ul * ul;
  ecinl += 0.5 * rl * qleft[IV] * qleft[IV];
  double etotl = pl * entho + ecinl;
  double ptotl = pl;

  double rr = fmax(qright[ID], smallr);
  double pr = fmax(qright[IP], rr * smallp);
  double ur = qright[IU];
  double ecinr = 0.5 * rr * ur * ur;
  ecinr += 0.5 * rr * qright[IV] * qright[IV];
  double etotr = pr * entho + ecinr;
  double ptotr = pr;

  double cfastl = sqrt(fmax(gamma0 * pl / rl, smallc * smallc));
  double cfastr = sqrt(fmax(gamma0 * pr / rr, smallc * smallc));

  double SL = fmin(ul, ur) - fmax(cfastl, cfastr);
  double SR = fmax(ul, ur) + fmax(cfastl, cfastr);

  double rcl = rl * (ul - SL);
  double rcr = rr * (SR - ur);

  double ustar = (rcr * ur + rcl * ul + (ptotl - ptotr)) / (rcr + rcl);
  double ptotstar = (rcr * ptotl + rcl * ptotr + rcl * rcr * (ul - ur)) / (rcr + rcl);

  double rstarl = rl * (SL - ul) / (SL - ustar);
  double etotstarl = ((SL - ul) * etotl - ptotl * ul + ptotstar * ustar) / (SL - ustar);
  double rstarr = rr * (SR - ur) / (SR - ustar);
  double etotstarr = ((SR - ur) * etotr - ptotr * ur + ptotstar * ustar) / (SR - ustar);

  double ro, uo, ptoto, etoto;
  if (SL > 0.0) {
    ro = rl; uo = ul; ptoto = ptotl; etoto = etotl;
  } else if (ustar > 0.0) {
    ro = rstarl; uo = ustar; ptoto = ptotstar; etoto = etotstarl;
  } else if (SR > 0.0) {
    ro = rstarr; uo = ustar; ptoto = ptotstar; etoto = etotstarr;
  } else {
    ro = rr; uo = ur; ptoto = ptotr; etoto = etotr;
  }

  flux[ID] = ro * uo;
  flux[IU] = ro * uo * uo + ptoto;
  flux[IP] = (etoto + ptoto) * uo;
  if (flux[ID] > 0.0)
    flux[IV] = flux[ID] * qleft[IV];
  else
    flux[IV] = flux[ID] * qright[IV];
}

/* HLL (absent from the reference; HydroParams.h:49 declares the enum only).
 * Toro, "Riemann Solvers and Numerical Methods for Fluid Dynamics", 3rd ed.,
 * 2009, section 10.3, eq. (10.21), with the same floors and wave-speed
 * estimates SL, SR as riemann_hllc above. */
void ora_riemann_hll(const ora_params *P, const double qleft[4], const double qright[4],
                     double flux[4]) {
  const double gamma0 = P->gamma0, smallr = P->smallr, smallp = P->smallp, smallc = P->smallc;
  const double entho = 1.0 / (gamma0 - 1.0);
  double rl = fmax(qleft[ID], smallr);
  double pl = fmax(qleft[IP], rl * smallp);
  double ul = qleft[IU], vl = qleft[IV];
  double rr = fmax(qright[ID], smallr);
  double pr = fmax(qright[IP], rr * smallp);
  double ur = qright[IU], vr = qright[IV];
  double el = pl * entho + (0.5 * rl * ul * ul + 0.5 * rl * vl * vl);
  double er = pr * entho + (0.5 * rr * ur * ur + 0.5 * rr * vr * vr);
  double cfastl = sqrt(fmax(gamma0 * pl / rl, smallc * smallc));
  double cfastr = sqrt(fmax(gamma0 * pr / rr, smallc * smallc));
  double SL = fmin(ul, ur) - fmax(cfastl, cfastr);
  double SR = fmax(ul, ur) + fmax(cfastl, cfastr);
  /* conserved (rho, E, rho u, rho v) and physical x-fluxes */
  double UL[4] = {rl, el, rl * ul, rl * vl};
  double UR[4] = {rr, er, rr * ur, rr * vr};
  double FL[4] = {rl * ul, (el + pl) * ul, rl * ul * ul + pl, rl * ul * vl};
  double FR[4] = {rr * ur, (er + pr) * ur, rr * ur * ur + pr, rr * ur * vr};
  if (SL >= 0.0) {
    for (int v = 0; v < 4; v++) flux[v] = FL[v];
  } else if (SR <= 0.0) {
    for (int v = 0; v < 4; v++) flux[v] = FR[v];
  } else {
    double inv = 1.0 / (SR - SL);
    for (int v = 0; v < 4; v++)
      flux[v] = (SR * FL[v] - SL * FR[v] + SL * SR * (UR[v] - UL[v])) * inv;
  }
}

/* cmpflx, HydroBaseFunctor.h:523-547 */
static void cmpflx(const ora_params *P, const double qgdnv[4], double flux[4]) {
  const double gamma0 = P->gamma0;
  flux[ID] = qgdnv[ID] * qgdnv[IU];
  flux[IU] = flux[ID] * qgdnv[IU] + qgdnv[IP];
  flux[IV] = flux[ID] * qgdnv[IV];
  double entho = 1.0 / (gamma0 - 1.0);
  double ekin = 0.5 * qgdnv[ID] * (qgdnv[IU] * qgdnv[IU] + qgdnv[IV] * qgdnv[IV]);
  double etot = qgdnv[IP] * entho + ekin;
  flux[IP] = qgdnv[IU] * (etot + qgdnv[IP]);
}

/* riemann_approx, HydroBaseFunctor.h:558-693 (10 Newton iterations hard-coded
 * at :592, not niter_riemann) */
void ora_riemann_approx(const ora_params *P, const double qleft[4], const double qright[4],
                        double qgdnv[4], double flux[4]) {
  const double gamma0 = P->gamma0, gamma6 = P->gamma6, smallr = P->smallr;
  const double smallc = P->smallc, smallp = P->smallp, smallpp = P->smallpp;
  double rl = fmax(qleft[ID], smallr);
  double ul = qleft[IU];
  double pl = fmax(qleft[IP], rl * smallp);
  double rr = fmax(qright[ID], smallr);
  double ur = qright[IU];
  double pr = fmax(qright[IP], rr * smallp);
  double cl = gamma0 * pl * rl;
  double cr = gamma0 * pr * rr;
  double wl = sqrt(cl);
  double wr = sqrt(cr);
  double pstar = fmax(((wr * pl + wl * pr) + wl * wr * (ul - ur)) / (wl + wr), 0.0);
  double pold = pstar;
  double conv = 1.0;
  for (int iter = 0; (iter < 10) && (conv > 1e-6); ++iter) {
    double wwl = sqrt(cl * (1.0 + gamma6 * (pold - pl) / pl));
    double wwr = sqrt(cr * (1.0 + gamma6 * (pold - pr) / pr));
    double ql = 2.0 * wwl * wwl * wwl / (wwl * wwl + cl);
    double qr = 2.0 * wwr * wwr * wwr / (wwr * wwr + cr);
    double usl = ul - (pold - pl) / wwl;
    double usr = ur + (pold - pr) / wwr;
    double delp = fmax(qr * ql / (qr + ql) * (usl - usr), -pold);
    pold = pold + delp;
    conv = fabs(delp / (pold + smallpp));
  }
  pstar = pold;
  wl = sqrt(cl * (1.0 + gamma6 * (pstar - pl) / pl));
  wr = sqrt(cr * (1.0 + gamma6 * (pstar - pr) / pr));
  double ustar = 0.5 * (ul + (pl - pstar) / wl + ur - (pr - pstar) / wr);
  double sgnm = copysign(1.0, ustar);
  double ro, uo, po, wo;
  if (sgnm > 0.0) {
    ro = rl; uo = ul; po = pl; wo = wl;
  } else {
    ro = rr; uo = ur; po = pr; wo = wr;
  }
  double co = fmax(smallc, sqrt(fabs(gamma0 * po / ro)));
  double rstar = fmax((ro / (1.0 + ro * (po - pstar) / (wo * wo))), smallr);
  double cstar = fmax(smallc, sqrt(fabs(gamma0 * pstar / rstar)));
  double spout = co - sgnm * uo;
  double spin = cstar - sgnm * ustar;
  double ushock = wo / ro - sgnm * uo;
  if (pstar >= po) {
    spin = ushock;
    spout = ushock;
  }
  double scr = fmax(spout - spin, smallc + fabs(spout + spin));
  double frac = 0.5 * (1.0 + (spout + spin) / scr);
  if (frac != frac)
    frac = 0.0;
  else
    frac = frac >= 1.0 ? 1.0 : frac <= 0.0 ? 0.0 : frac;
  qgdnv[ID] = frac * rstar + (1.0 - frac) * ro;
  qgdnv[IU] = frac * ustar + (1.0 - frac) * uo;
  qgdnv[IP] = frac * pstar + (1.0 - frac) * po;
  if (spout < 0.0) {
    qgdnv[ID] = ro;
    qgdnv[IU] = uo;
    qgdnv[IP] = po;
  }
  if (spin > 0.0) {
    qgdnv[ID] = rstar;
    qgdnv[IU] = ustar;
    qgdnv[IP] = pstar;
  }
  qgdnv[IV] = (sgnm > 0.0) ? qleft[IV] : qright[IV];
  cmpflx(P, qgdnv, flux);
}

/* solver dispatch: the reference always calls riemann_hllc
 * (HydroRunFunctors.h:567,631); honor_riemann=1 selects by riemannSolverType */
static void riemann(const ora_params *P, const double ql[4], const double qr[4], double flux[4]) {
  if (P->honor_riemann) {
    if (P->riemannSolverType == RIEMANN_HLL) {
      ora_riemann_hll(P, ql, qr, flux);
      return;
    }
    if (P->riemannSolverType == RIEMANN_APPROX) {
      double qg[4];
      ora_riemann_approx(P, ql, qr, qg, flux);
      return;
    }
  }
  ora_riemann_hllc(P, ql, qr, flux);
}

/* ======================================================================
 * HydroRun, implementation version 0 (src/HydroRun.h:142-399)
 * ====================================================================== */
struct ora_run {
  ora_params p;
  double *arr[5]; /* U, U2, Q, Fluxes_x, Fluxes_y */
};

#define AT(a, i, j, v) (a)[(((size_t)(i) * (size_t)P->jsize + (size_t)(j)) << 2) + (size_t)(v)]

void ora_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int ora_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

int ora_create(const ora_params *p, ora_run **out) {
  if (!p || !out || p->nx < 2 || p->ny < 2) return -1;
  ora_run *r = (ora_run *)calloc(1, sizeof(ora_run));
  r->p = *p;
  size_t n = (size_t)p->isize * (size_t)p->jsize * 4;
  for (int k = 0; k < 5; k++) {
    r->arr[k] = (double *)calloc(n, sizeof(double));
    if (!r->arr[k]) {
      ora_destroy(r);
      return -2;
    }
  }
  *out = r;
  return 0;
}

void ora_destroy(ora_run *r) {
  if (!r) return;
  for (int k = 0; k < 5; k++) free(r->arr[k]);
  free(r);
}

int ora_get_array(ora_run *r, int which, double *out) {
  if (which < 0 || which > 4) return -1;
  memcpy(out, r->arr[which], (size_t)r->p.isize * r->p.jsize * 4 * sizeof(double));
  return 0;
}

int ora_set_array(ora_run *r, int which, const double *in) {
  if (which < 0 || which > 4) return -1;
  memcpy(r->arr[which], in, (size_t)r->p.isize * r->p.jsize * 4 * sizeof(double));
  return 0;
}

/* HydroRun::saveVTK, HydroRun.h:502-609.  ostream's default formatting of a
 * real is printf's %g (precision 6). */
int ora_save_vtk(ora_run *r, int which, int use_float32, const char *output_dir,
                 const char *output_prefix, int step) {
  static const char *const names[4] = {"rho", "E", "mx", "my"}; /* HydroParams.cpp:17 */
  const ora_params *P = &r->p;
  const double *U = r->arr[which];
  char fname[4096];
  const int one = 1;
  if (which < 0 || which > 1) return -1;
  snprintf(fname, sizeof(fname), "%s/%s_%07d.vti", output_dir, output_prefix, step);
  FILE *f = fopen(fname, "w");
  if (!f) return -2;
  fprintf(f, "<?xml version=\"1.0\"?>\n");
  fprintf(f, "<VTKFile type=\"ImageData\" version=\"0.1\" byte_order=\"%s\">\n",
          (*(const char *)&one) ? "LittleEndian" : "BigEndian");
  fprintf(f, "  <ImageData WholeExtent=\"%d %d %d %d %d %d\" Origin=\"%g %g %g\" "
             "Spacing=\"%g %g %g\">\n",
          0, P->nx, 0, P->ny, 0, 0, P->xmin, P->ymin, 0.0, P->dx, P->dy, 0.0);
  fprintf(f, "  <Piece Extent=\"%d %d %d %d %d %d \">\n", 0, P->nx, 0, P->ny, 0, 0);
  fprintf(f, "    <PointData>\n    </PointData>\n    <CellData>\n");
  const int gw = P->ghostWidth;
  for (int v = 0; v < 4; v++) {
    fprintf(f, "    <DataArray type=\"%s\" Name=\"%s\" format=\"ascii\" >\n",
            use_float32 ? "Float32" : "Float64", names[v]);
    for (int index = 0; index < P->isize * P->jsize; ++index) {
      const int j = index / P->isize;
      const int i = index - j * P->isize;
      if (j >= P->jmin + gw && j <= P->jmax - gw && i >= P->imin + gw && i <= P->imax - gw) {
        const double x = AT(U, i, j, v);
        fprintf(f, "%g ", use_float32 ? (double)(float)x : x);
      }
    }
    fprintf(f, "\n    </DataArray>\n");
  }
  fprintf(f, "    </CellData>\n  </Piece>\n  </ImageData>\n</VTKFile>\n");
  return fclose(f) == 0 ? 0 : -2;
}

/* ComputeRadialProfileFunctor, ComputeRadialProfileFunctor.h:83-165 */
int ora_radial_profile(ora_run *r, int which, int nbins, double *distances, double *density,
                       int64_t *counts) {
  const ora_params *P = &r->p;
  const double *U = r->arr[which];
  if (which < 0 || which > 1 || nbins < 1) return -1;
  const double cx = (P->xmin + P->xmax) / 2, cy = (P->ymin + P->ymax) / 2;
  const double Dx = (P->xmax - P->xmin) / 2, Dy = (P->ymax - P->ymin) / 2;
  const double maxr = sqrt(Dx * Dx + Dy * Dy);
  int64_t *cnt = (int64_t *)calloc((size_t)nbins, sizeof(int64_t));
  double *sum = (double *)calloc((size_t)nbins, sizeof(double));
  if (!cnt || !sum) {
    free(cnt);
    free(sum);
    return -3;
  }
  for (int i = 0; i < P->isize; i++)
    for (int j = 0; j < P->jsize; j++) {
      const double x = P->xmin + P->dx / 2 + (i - P->ghostWidth) * P->dx;
      const double y = P->ymin + P->dy / 2 + (j - P->ghostWidth) * P->dy;
      const double distance = sqrt((x - cx) * (x - cx) + (y - cy) * (y - cy));
      const int bin = (int)(distance / maxr * nbins);
      if (bin < 0 || bin >= nbins) continue;
      cnt[bin] += 1;
      sum[bin] += AT(U, i, j, ID);
    }
  const double dr = maxr / nbins;
  for (int b = 0; b < nbins; b++) {
    distances[b] = (b + 0.5) * dr;
    density[b] = sum[b] / (int)cnt[b];
    if (counts) counts[b] = cnt[b];
  }
  free(cnt);
  free(sum);
  return 0;
}

/* cnpy::create_npy_header<double> + npy_save, cnpy.h:129-187, 343-376 */
int ora_write_npy(const char *path, const double *data, int64_t n) {
  char dict[256];
  const int one = 1;
  int len = snprintf(dict, sizeof(dict), "{'descr': '%cf8', 'fortran_order': False, 'shape': (%lld,), }",
                     (*(const char *)&one) ? '<' : '>', (long long)n);
  const int pad = 16 - (10 + len) % 16;
  for (int k = 0; k < pad; k++) dict[len + k] = ' ';
  len += pad;
  dict[len - 1] = '\n';
  FILE *f = fopen(path, "wb");
  if (!f) return -2;
  const unsigned char magic[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
  const unsigned char hl[2] = {(unsigned char)(len & 0xff), (unsigned char)(len >> 8)};
  int ok = fwrite(magic, 1, 8, f) == 8 && fwrite(hl, 1, 2, f) == 2 &&
           fwrite(dict, 1, (size_t)len, f) == (size_t)len &&
           (n == 0 || fwrite(data, sizeof(double), (size_t)n, f) == (size_t)n);
  return (fclose(f) == 0 && ok) ? 0 : -2;
}

static void set4(double *U, const ora_params *P, int i, int j, double d, double e, double mx, double my) {
  AT(U, i, j, ID) = d;
  AT(U, i, j, IP) = e;
  AT(U, i, j, IU) = mx;
  AT(U, i, j, IV) = my;
}

/* Init*Functor, HydroRunFunctors.h:1347-1827; cell centres as at :1383-1384 */
void ora_init_problem(ora_run *r) {
  const ora_params *P = &r->p;
  double *U = r->arr[0];
  const int gw = P->ghostWidth;
  const double xmin = P->xmin, ymin = P->ymin, dx = P->dx, dy = P->dy, gamma0 = P->gamma0;
  const int isize = P->isize, jsize = P->jsize;

  if (P->problemType == 1) {
    /* InitBlastFunctor (:1402-1527); the volume sum is done in a fixed
     * sequential i-then-j order */
    const double radius2 = P->blast_radius * P->blast_radius;
    double volume = 0.0;
    for (int i = 0; i < isize; i++)
      for (int j = 0; j < jsize; j++) {
        double x = xmin + dx / 2 + (i - gw) * dx;
        double y = ymin + dy / 2 + (j - gw) * dy;
        double d2 = (x - P->blast_center_x) * (x - P->blast_center_x) +
                    (y - P->blast_center_y) * (y - P->blast_center_y);
        if (d2 < radius2) {
          set4(U, P, i, j, P->blast_density_in, P->blast_pressure_in / (gamma0 - 1.0), 0.0, 0.0);
          volume += dx * dy;
        } else {
          set4(U, P, i, j, P->blast_density_out, P->blast_pressure_out / (gamma0 - 1.0), 0.0, 0.0);
        }
      }
    if (P->blast_total_energy_inside > 0) {
      for (int i = 0; i < isize; i++)
        for (int j = 0; j < jsize; j++) {
          double x = xmin + dx / 2 + (i - gw) * dx;
          double y = ymin + dy / 2 + (j - gw) * dy;
          double d2 = (x - P->blast_center_x) * (x - P->blast_center_x) +
                      (y - P->blast_center_y) * (y - P->blast_center_y);
          if (d2 < radius2) AT(U, i, j, IP) = P->blast_total_energy_inside / volume;
        }
    }
  } else {
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < isize; i++)
      for (int j = 0; j < jsize; j++) {
        double x = xmin + dx / 2 + (i - gw) * dx;
        double y = ymin + dy / 2 + (j - gw) * dy;
        if (P->problemType == 2) {
          /* InitFourQuadrantFunctor :1531-1646 */
          double W[4][4] = {{1.5, 1.5, 0.0, 0.0},
                            {0.5323, 0.3, 1.206, 0.0},
                            {0.138, 0.029, 1.206, 1.206},
                            {0.5323, 0.3, 0.0, 1.206}};
          for (int k = 0; k < 4; k++) {
            double rho = W[k][ID], p = W[k][IP], u = W[k][IU], v = W[k][IV];
            W[k][IU] *= rho;
            W[k][IV] *= rho;
            W[k][IP] = p / (gamma0 - 1.0) + rho * (u * u + v * v) * 0.5;
          }
          int k = (x < 0.8) ? ((y < 0.8) ? 2 : 1) : ((y < 0.8) ? 3 : 0);
          set4(U, P, i, j, W[k][ID], W[k][IP], W[k][IU], W[k][IV]);
        } else if (P->problemType == 3) {
          /* InitDiscontinuityFunctor :1651-1712 */
          if (x + y < 1)
            set4(U, P, i, j, 1.0 + x * x, 1.0 / (gamma0 - 1.0), 0.0, 0.0);
          else
            set4(U, P, i, j, 0.25, 1.0 / (gamma0 - 1.0), 0.0, 0.0);
        } else if (P->problemType == 4) {
          /* InitShockedBubbleFunctor :1717-1827 */
          double d, mx, my, e;
          if (x < P->sb_shock_loc) {
            d = P->sb_postshock_density;
            mx = P->sb_postshock_density * P->sb_postshock_velocity;
            my = 0.0;
            double rho_eint = P->sb_postshock_pressure / (gamma0 - 1);
            e = rho_eint + 0.5 * (mx * mx + my * my) / d;
          } else {
            double radius = sqrt((x - P->sb_bubble_center_x) * (x - P->sb_bubble_center_x) +
                                 (y - P->sb_bubble_center_y) * (y - P->sb_bubble_center_y));
            if (radius < P->sb_bubble_radius) {
              d = P->sb_bubble_density;
              mx = 0.0;
              my = 0.0;
              double rho_eint = P->sb_bubble_pressure / (gamma0 - 1);
              e = rho_eint + 0.5 * (mx * mx + my * my) / d;
            } else {
              d = P->sb_preshock_density;
              mx = 0.0;
              my = 0.0;
              double rho_eint = P->sb_preshock_pressure / (gamma0 - 1);
              e = rho_eint + 0.5 * (mx * mx + my * my) / d;
            }
          }
          set4(U, P, i, j, d, e, mx, my);
        } else {
          /* InitImplodeFunctor :1347-1399 (high-state pressure is an
           * extension key; 1.0 reproduces 1.0/(gamma0-1.0) bit for bit) */
          double tmp = x + y * y;
          if (tmp > 0.5 && tmp < 1.5)
            set4(U, P, i, j, 1.0, P->implode_pressure_in / (gamma0 - 1.0), 0.0, 0.0);
          else
            set4(U, P, i, j, 0.125, 0.14 / (gamma0 - 1.0), 0.0, 0.0);
        }
      }
  }
  /* HydroRun.h:214: U2 = U */
  memcpy(r->arr[1], r->arr[0], (size_t)isize * jsize * 4 * sizeof(double));
}

/* ComputeDtFunctor, HydroRunFunctors.h:17-79 + HydroRun::compute_dt :227-251 */
double ora_compute_dt(ora_run *r, int useU) {
  const ora_params *P = &r->p;
  const double *U = useU == 0 ? r->arr[0] : r->arr[1];
  const int gw = P->ghostWidth, isize = P->isize, jsize = P->jsize;
  double invDt = -DBL_MAX;
#pragma omp parallel
  {
    double loc = -DBL_MAX;
#pragma omp for collapse(2) schedule(static)
    for (int i = gw; i < isize - gw; i++)
      for (int j = gw; j < jsize - gw; j++) {
        double u[4] = {AT(U, i, j, 0), AT(U, i, j, 1), AT(U, i, j, 2), AT(U, i, j, 3)};
        double q[4], c = 0.0;
        ora_compute_primitives(P, u, &c, q);
        double vx = c + fabs(q[IU]);
        double vy = c + fabs(q[IV]);
        loc = fmax(loc, vx / P->dx + vy / P->dy);
      }
#pragma omp critical
    invDt = fmax(invDt, loc);
  }
  return P->cfl / invDt;
}

/* MakeBoundariesFunctor, HydroRunFunctors.h:1832-2030, launched for XMIN,
 * XMAX, YMIN, YMAX in that order (HydroRun.h:388-399) */
void ora_make_boundaries(ora_run *r, int which) {
  const ora_params *P = &r->p;
  double *U = which == 0 ? r->arr[0] : r->arr[1];
  const int nx = P->nx, ny = P->ny, gw = P->ghostWidth;
  /* XMIN */
  for (int j = P->jmin; j <= P->jmax; j++)
    for (int i = 0; i < gw; i++)
      for (int v = 0; v < 4; v++) {
        double sign = 1.0;
        int i0;
        if (P->boundary_type_xmin == BC_DIRICHLET) {
          i0 = 2 * gw - 1 - i;
          if (v == IU) sign = -1.0;
        } else if (P->boundary_type_xmin == BC_NEUMANN) {
          i0 = gw;
        } else {
          i0 = nx + i;
        }
        AT(U, i, j, v) = AT(U, i0, j, v) * sign;
      }
  /* XMAX */
  for (int j = P->jmin; j <= P->jmax; j++)
    for (int i = nx + gw; i <= nx + 2 * gw - 1; i++)
      for (int v = 0; v < 4; v++) {
        double sign = 1.0;
        int i0;
        if (P->boundary_type_xmax == BC_DIRICHLET) {
          i0 = 2 * nx + 2 * gw - 1 - i;
          if (v == IU) sign = -1.0;
        } else if (P->boundary_type_xmax == BC_NEUMANN) {
          i0 = nx + gw - 1;
        } else {
          i0 = i - nx;
        }
        AT(U, i, j, v) = AT(U, i0, j, v) * sign;
      }
  /* YMIN */
  for (int i = P->imin; i <= P->imax; i++)
    for (int j = 0; j < gw; j++)
      for (int v = 0; v < 4; v++) {
        double sign = 1.0;
        int j0;
        if (P->boundary_type_ymin == BC_DIRICHLET) {
          j0 = 2 * gw - 1 - j;
          if (v == IV) sign = -1.0;
        } else if (P->boundary_type_ymin == BC_NEUMANN) {
          j0 = gw;
        } else {
          j0 = ny + j;
        }
        AT(U, i, j, v) = AT(U, i, j0, v) * sign;
      }
  /* YMAX */
  for (int i = P->imin; i <= P->imax; i++)
    for (int j = ny + gw; j <= ny + 2 * gw - 1; j++)
      for (int v = 0; v < 4; v++) {
        double sign = 1.0;
        int j0;
        if (P->boundary_type_ymax == BC_DIRICHLET) {
          j0 = 2 * ny + 2 * gw - 1 - j;
          if (v == IV) sign = -1.0;
        } else if (P->boundary_type_ymax == BC_NEUMANN) {
          j0 = ny + gw - 1;
        } else {
          j0 = j - ny;
        }
        AT(U, i, j, v) = AT(U, i, j0, v) * sign;
      }
}

/* ConvertToPrimitivesFunctor, HydroRunFunctors.h:84-143 (all cells) */
void ora_convert_to_primitives(ora_run *r, int which) {
  const ora_params *P = &r->p;
  const double *U = which == 0 ? r->arr[0] : r->arr[1];
  double *Q = r->arr[2];
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 0; i < P->isize; i++)
    for (int j = 0; j < P->jsize; j++) {
      double u[4] = {AT(U, i, j, 0), AT(U, i, j, 1), AT(U, i, j, 2), AT(U, i, j, 3)};
      double q[4], c;
      ora_compute_primitives(P, u, &c, q);
      for (int v = 0; v < 4; v++) AT(Q, i, j, v) = q[v];
    }
}

static void load4(const double *Q, const ora_params *P, int i, int j, double q[4]) {
  for (int v = 0; v < 4; v++) q[v] = AT(Q, i, j, v);
}

/* ComputeAndStoreFluxesFunctor, HydroRunFunctors.h:412-651 */
void ora_compute_and_store_fluxes(ora_run *r, double dtdx, double dtdy) {
  const ora_params *P = &r->p;
  const double *Q = r->arr[2];
  double *Fx = r->arr[3], *Fy = r->arr[4];
  const int gw = P->ghostWidth;
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = gw; i <= P->isize - gw; i++)
    for (int j = gw; j <= P->jsize - gw; j++) {
      double qLoc[4], qN0[4], qN1[4], qN2[4], qN3[4], qNb[4];
      double dqX[4], dqY[4], dqXn[4], dqYn[4], qleft[4], qright[4], flux[4];

      load4(Q, P, i, j, qLoc);
      load4(Q, P, i + 1, j, qN0);
      load4(Q, P, i - 1, j, qN1);
      load4(Q, P, i, j + 1, qN2);
      load4(Q, P, i, j - 1, qN3);
      ora_slope_unsplit_hydro_2d(P, qLoc, qN0, qN1, qN2, qN3, dqX, dqY);

      /* left neighbour along X */
      load4(Q, P, i - 1, j, qNb);
      load4(Q, P, i, j, qN0);
      load4(Q, P, i - 2, j, qN1);
      load4(Q, P, i - 1, j + 1, qN2);
      load4(Q, P, i - 1, j - 1, qN3);
      ora_slope_unsplit_hydro_2d(P, qNb, qN0, qN1, qN2, qN3, dqXn, dqYn);
      ora_trace_unsplit_2d_along_dir(P, qLoc, dqX, dqY, dtdx, dtdy, FACE_XMIN, qright);
      ora_trace_unsplit_2d_along_dir(P, qNb, dqXn, dqYn, dtdx, dtdy, FACE_XMAX, qleft);
      riemann(P, qleft, qright, flux);
      for (int v = 0; v < 4; v++) AT(Fx, i, j, v) = flux[v] * dtdx;

      /* left neighbour along Y */
      load4(Q, P, i, j - 1, qNb);
      load4(Q, P, i + 1, j - 1, qN0);
      load4(Q, P, i - 1, j - 1, qN1);
      load4(Q, P, i, j, qN2);
      load4(Q, P, i, j - 2, qN3);
      ora_slope_unsplit_hydro_2d(P, qNb, qN0, qN1, qN2, qN3, dqXn, dqYn);
      ora_trace_unsplit_2d_along_dir(P, qLoc, dqX, dqY, dtdx, dtdy, FACE_YMIN, qright);
      ora_trace_unsplit_2d_along_dir(P, qNb, dqXn, dqYn, dtdx, dtdy, FACE_YMAX, qleft);
      double t;
      t = qleft[IU]; qleft[IU] = qleft[IV]; qleft[IV] = t;
      t = qright[IU]; qright[IU] = qright[IV]; qright[IV] = t;
      riemann(P, qleft, qright, flux);
      t = flux[IU]; flux[IU] = flux[IV]; flux[IV] = t;
      for (int v = 0; v < 4; v++) AT(Fy, i, j, v) = flux[v] * dtdy;
    }
}

/* UpdateFunctor, HydroRunFunctors.h:656-723 */
void ora_update(ora_run *r, int which) {
  const ora_params *P = &r->p;
  double *U = which == 0 ? r->arr[0] : r->arr[1];
  const double *Fx = r->arr[3], *Fy = r->arr[4];
  const int gw = P->ghostWidth;
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = gw; i < P->isize - gw; i++)
    for (int j = gw; j < P->jsize - gw; j++)
      for (int v = 0; v < 4; v++) {
        AT(U, i, j, v) += AT(Fx, i, j, v);
        AT(U, i, j, v) -= AT(Fx, i + 1, j, v);
        AT(U, i, j, v) += AT(Fy, i, j, v);
        AT(U, i, j, v) -= AT(Fy, i, j + 1, v);
      }
}

/* HydroRun::godunov_unsplit / _impl, HydroRun.h:258-364 (version 0) */
void ora_godunov_unsplit(ora_run *r, int nStep, double dt) {
  const ora_params *P = &r->p;
  int in = (nStep % 2 == 0) ? 0 : 1;
  int out = 1 - in;
  double dtdx = dt / P->dx;
  double dtdy = dt / P->dy;
  ora_make_boundaries(r, in);
  memcpy(r->arr[out], r->arr[in], (size_t)P->isize * P->jsize * 4 * sizeof(double));
  ora_convert_to_primitives(r, in);
  ora_compute_and_store_fluxes(r, dtdx, dtdy);
  ora_update(r, out);
}

/* main loop, src/main.cpp:87-145 (no output) */
void ora_run_loop(ora_run *r, double *t_io, int *nStep_io, int max_steps) {
  const ora_params *P = &r->p;
  double t = *t_io, dt;
  int nStep = *nStep_io;
  int done = 0;
  if (nStep == 0) {
    /* startup: dt from U, boundaries of U and U2 (main.cpp:87-91) */
    dt = ora_compute_dt(r, 0);
    (void)dt;
    ora_make_boundaries(r, 0);
    ora_make_boundaries(r, 1);
  }
  while (t < P->tEnd && nStep < P->nStepmax && (max_steps <= 0 || done < max_steps)) {
    dt = ora_compute_dt(r, nStep % 2);
    if (t + dt > P->tEnd) dt = P->tEnd - t;
    ora_godunov_unsplit(r, nStep, dt);
    nStep++;
    t += dt;
    done++;
  }
  *t_io = t;
  *nStep_io = nStep;
}
