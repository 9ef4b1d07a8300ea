// trace.hpp -- named host ranges for rocprofv3 --marker-trace (roctx), with
// the names of the reference's Kokkos profiling regions
// (src/HydroRun.h:242-357, src/main.cpp:93-165): main_loop, output,
// compute_dt, make_boundaries, compute_primitives, hydro_impl0/1/2,
// compute_fluxes, update_hydro.  Without a tool attached a range costs a
// function call.
#pragma once

#include <rocprofiler-sdk-roctx/roctx.h>

namespace e2d {

struct Region {
  explicit Region(const char *name) { roctxRangePushA(name); }
  ~Region() { roctxRangePop(); }
  Region(const Region &) = delete;
  Region &operator=(const Region &) = delete;
};

}  // namespace e2d
