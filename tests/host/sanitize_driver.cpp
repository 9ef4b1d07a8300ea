// Host-side driver for the sanitizer test (tests/test_host_sanitizers.py):
// runs the product's host code -- the .ini reader, HydroParams setup / init /
// check, the .vti and .npy writers -- on the decks given on the command line,
// built with -fsanitize=address,undefined so that an out-of-bounds access, a
// leak or undefined behaviour aborts the run.  Test infrastructure only.
#include <cstdio>
#include <string>
#include <vector>

#include "config.hpp"
#include "euler2d_hip.h"
#include "params.hpp"
#include "status.hpp"

// the C-ABI error channel lives in api.cpp, which needs the HIP runtime
int e2d::report(int code, const std::string &) { return code; }

int main(int argc, char **argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: sanitize_driver <outdir> deck.ini...\n");
    return 2;
  }
  const std::string outdir = argv[1];
  int checked = 0;
  for (int a = 2; a < argc; ++a) {
    e2d::ConfigMap cfg(argv[a]);
    // every getter on keys that exist and keys that do not
    const char *sections[] = {"run", "mesh", "hydro", "blast", "implode", "output", "OTHER",
                              "missing", ""};
    const char *names[] = {"tEnd", "nStepmax", "nx", "ny", "cfl", "problem", "riemann",
                           "gamma0", "radius", "outputDir", "implementationVersion", "x", ""};
    for (const char *s : sections)
      for (const char *n : names) {
        (void)cfg.getString(s, n, "dflt");
        (void)cfg.getInteger(s, n, 7);
        (void)cfg.getFloat(s, n, 1.5f);
        (void)cfg.getBool(s, n, true);
      }
    e2d_params p;
    e2d::params_defaults(p);
    e2d::params_setup(cfg, p);
    e2d::params_init(p);
    (void)e2d::params_check(p);
    // small output files from a synthetic state
    if (p.nx > 0 && p.ny > 0 && (long long)p.nx * p.ny <= 1 << 20) {
      std::vector<double> st((size_t)4 * p.nx * p.ny);
      for (size_t k = 0; k < st.size(); ++k) st[k] = 1.0 + 1e-3 * (double)(k % 97) - 0.05;
      if (e2d_write_vti(&p, st.data(), a % 2, outdir.c_str(), "san", a) != E2D_OK) return 3;
      const std::string npy = outdir + "/san_" + std::to_string(a) + ".npy";
      if (e2d_write_npy(npy.c_str(), st.data(), (int64_t)st.size()) != E2D_OK) return 4;
    }
    ++checked;
  }
  std::printf("checked %d decks\n", checked);
  return 0;
}
