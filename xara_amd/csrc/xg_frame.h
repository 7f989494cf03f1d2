// Frame elements (DispBeamColumn2d/3d, ForceBeamColumn2d/3d over fiber
// sections) on gfx950.
#pragma once
#include <memory>
#include <vector>

#include "xg_domain.h"
#include "xg_frame_dev.h"

namespace xg {

// force = true: ForceBeamColumn2d/3d elements (Lobatto sections)
std::unique_ptr<Group> make_frame_group(Domain& d, bool force = false);
void frame_group_add(Domain& d, Group& g, int tag, const EleDef& ed);
// device view of a frame group (throws for any other element group)
FrameDev frame_dev(Group& g);
// regroups the displacement-based update's sections by (part, definition),
// the parts being element ranges [ele_off[p], ele_off[p+1]); returns the
// first wavefront item of each part and the end (empty if one wavefront per
// section is in use)
std::vector<int> frame_rows_by_parts(Group& g, const std::vector<int>& ele_off, hipStream_t s);

}  // namespace xg
