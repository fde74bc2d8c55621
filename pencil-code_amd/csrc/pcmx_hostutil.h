// pcmx_hostutil.h -- host helpers shared by the host translation units
// (pcmx_host.hip: single box; pcmx_ens.hip: ensembles).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/pcmx.h"
#include "pcmx_physics.h"

namespace pcmx {

// Records the message returned by pcmx_last_error() and returns code.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
const char* last_error();

// Williamson (1980) 2N-RK3 coefficients, timestep.f90:43-44.
extern const double ALPHA_TS[3];
extern const double BETA_TS[3];

// Validates run parameters (grid, variable slots, EOS, Courant factors).
int check_params(const pcmx_params* p);

// Derived constants of one box, in the reference's operation order.
Phys make_phys(const pcmx_params& p);

// The layer physics of samples/conv-slab is on: Coriolis force, K-profile
// conduction, heating or cooling.
inline bool layers_on(const pcmx_params& p) {
  return p.omega != 0.0 || p.lheatc_kprof != 0 || p.luminosity != 0.0 || p.cool != 0.0;
}

}  // namespace pcmx

#define HIPCHECK(call)                                                                                  \
  do {                                                                                                  \
    hipError_t e_ = (call);                                                                             \
    if (e_ != hipSuccess) return pcmx::fail(PCMX_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
  } while (0)
