// Frame elements (DispBeamColumn2d/3d, ForceBeamColumn2d/3d over fiber
// sections) on gfx950.
#pragma once
#include <memory>

#include "xg_domain.h"
#include "xg_frame_dev.h"

namespace xg {

// force = true: ForceBeamColumn2d/3d elements (Lobatto sections)
std::unique_ptr<Group> make_frame_group(Domain& d, bool force = false);
void frame_group_add(Domain& d, Group& g, int tag, const EleDef& ed);
// device view of a frame group (throws for any other element group)
FrameDev frame_dev(Group& g);

}  // namespace xg
