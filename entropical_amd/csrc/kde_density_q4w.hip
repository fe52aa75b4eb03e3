// kde_density_q4w.hip -- k_density with at most 4 queries per lane: 64-bit cell keys.
#include "kde_density.hpp"

namespace ent {
ENT_DENSITY_VARIANT(4, false, uint64_t, 2)
}  // namespace ent
