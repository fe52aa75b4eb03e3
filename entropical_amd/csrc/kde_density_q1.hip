// kde_density_q1.hip -- k_density with at most 1 query per lane (all variants).
#include "kde_density.hpp"

namespace ent {
ENT_DENSITY_VARIANT(1, true, uint32_t, 2)
ENT_DENSITY_VARIANT(1, false, uint32_t, 2)
ENT_DENSITY_VARIANT(1, false, uint64_t, 2)
}  // namespace ent
