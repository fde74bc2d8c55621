// pcmx_kern_f64.hip -- the double entry points: the single-box RHS dispatch
// (its kernel variants are instantiated in pcmx_kern_phys.hip) and the
// ensemble kernels.
#include "pcmx_dispatch.h"

namespace pcmx {

PCMX_EXTERN_PHYS(double)

bool launch_rhs_f64(const LaunchCfg& c, const KArgs& a, const Phys& P, hipStream_t s) {
  return launch_rhs_t<double>(c, a, P, s);
}

bool launch_ens_f64(const LaunchCfg& c, int stage, const EnsBox* boxes, const EnsWork* work, int nwork,
                    double* partials, int nblocks, hipStream_t s) {
  return launch_ens_t<double>(c, stage, boxes, work, nwork, partials, nblocks, s);
}

}  // namespace pcmx
