// kde_density_q2.hip -- k_density with at most 2 queries per lane.
#include "kde_density.hpp"

namespace ent {
template int density_occupancy_q<2>(bool, bool, int*);
template void launch_density_q<2>(hipStream_t, bool, bool, unsigned, const DJob*, const DTile*, const uint32_t*,
                                   uint32_t*, const float*, const float*, const void*, const uint32_t*,
                                   const uint32_t*, const float4*, const float*, double*, unsigned long long*);
}  // namespace ent
