// pcmx_kern_phys.hip -- instantiations of the RHS kernel variants of one
// precision and one physics (magnetic or not, entropy or not).  The Makefile
// compiles this file once per combination, selected by PCMX_PREC (64 or 32),
// PCMX_MAG and PCMX_ENT (0 or 1), so the eight units build in parallel.
#include "pcmx_dispatch.h"

#if !defined(PCMX_PREC) || !defined(PCMX_MAG) || !defined(PCMX_ENT)
#error "pcmx_kern_phys.hip needs PCMX_PREC, PCMX_MAG and PCMX_ENT"
#endif

namespace pcmx {

#if PCMX_PREC == 64
template bool launch_phys<double, PCMX_MAG != 0, PCMX_ENT != 0>(const LaunchCfg&, const KArgs&, const Phys&,
                                                                hipStream_t);
#elif PCMX_PREC == 32
template bool launch_phys<float, PCMX_MAG != 0, PCMX_ENT != 0>(const LaunchCfg&, const KArgs&, const Phys&,
                                                               hipStream_t);
#else
#error "PCMX_PREC must be 64 or 32"
#endif

}  // namespace pcmx
