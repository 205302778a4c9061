// walk_mixed.hip -- mixed-precision walk (FP32 side tests with an FP64
// fallback filter).  Not built yet.
#include "sdtfe_host.h"

namespace sdtfe {
extern const bool kHaveMixed = false;
void launch_map_mixed(sdtfe_tess *, const MapGeom &, double *, int *, int) {}
void launch_trace_mixed(sdtfe_tess *, const MapGeom &, int, const double *, const double *,
                        const int *, const uint64_t *, const uint32_t *, double *, int *) {}
}  // namespace sdtfe
