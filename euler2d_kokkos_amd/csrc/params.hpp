// params.hpp -- HydroParams on the C-ABI POD (include/euler2d_hip.h).
#pragma once

#include "config.hpp"
#include "euler2d_hip.h"

namespace e2d {

void params_defaults(e2d_params &p);
void params_init(e2d_params &p);
void params_setup(const ConfigMap &cfg, e2d_params &p);
void params_print(const e2d_params &p);
// nullptr when the kernels can run these params, else the reason
const char *params_check(const e2d_params &p);

}  // namespace e2d
