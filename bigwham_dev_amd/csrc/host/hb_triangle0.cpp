// 3DT0-H kernel (3-D P0 triangle): placeholder until the restatement of
// src/elasticity/fullspace_iso_3d_triangle/elastic_3dT0_element.cc lands.
#include "hb_core.h"

namespace bwh {

std::unique_ptr<Kernel> make_triangle0_h(double, double) { return nullptr; }

}  // namespace bwh
