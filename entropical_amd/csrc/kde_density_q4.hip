// kde_density_q4.hip -- k_density with at most 4 queries per lane: plain path and 32-bit cell keys.
#include "kde_density.hpp"

namespace ent {
ENT_DENSITY_VARIANT(4, true, uint32_t, 2)
ENT_DENSITY_VARIANT(4, false, uint32_t, 2)
}  // namespace ent
