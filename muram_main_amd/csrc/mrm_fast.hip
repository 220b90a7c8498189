// mrm_fast.hip -- the fused-step kernels built with FMA contraction
// (-ffp-contract=fast): same source and operation order as the exact build,
// with a*b+c allowed to fuse.  Selected at run time (MRM_FMA=1; off by default).
#define MRM_NS fast
#include "mrm_launch.h"

namespace fast {
hipError_t (*p_launch_stage)(hipStream_t, unsigned, int, const StageArgs&, const double*, const double*,
                             const double*, double*, double*, const double*) = launch_stage;
hipError_t (*p_launch_tvd_x_lines)(hipStream_t, unsigned, unsigned, size_t, const TvdLineArgs&, double*,
                                   const double*) = launch_tvd_x_lines;
hipError_t (*p_launch_tvd_sweep)(hipStream_t, int, bool, const TvdEpilogue*, unsigned, const Geom&,
                                 const TvdParams&, double*, const double*) = launch_tvd_sweep;
}  // namespace fast
