// Frame elements (DispBeamColumn2d/3d over fiber sections) on gfx950.
#pragma once
#include <memory>

#include "xg_domain.h"

namespace xg {

std::unique_ptr<Group> make_frame_group(Domain& d);
void frame_group_add(Domain& d, Group& g, int tag, const EleDef& ed);

}  // namespace xg
