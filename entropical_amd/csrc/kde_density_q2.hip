// kde_density_q2.hip -- k_density with at most 2 queries per lane (all variants).
#include "kde_density.hpp"

namespace ent {
ENT_DENSITY_VARIANT(2, true, uint32_t, 2)
ENT_DENSITY_VARIANT(2, false, uint32_t, 2)
ENT_DENSITY_VARIANT(2, false, uint64_t, 2)
}  // namespace ent
