// kde_density_q1s3.hip -- k_density with one query per lane whose 3-D jobs
// run as sub-tiles (cell-list variants).
#include "kde_density.hpp"

namespace ent {
ENT_DENSITY_VARIANT(1, false, uint32_t, 3)
ENT_DENSITY_VARIANT(1, false, uint64_t, 3)
}  // namespace ent
