#include "xg_frame.h"
namespace xg {
std::unique_ptr<Group> make_frame_group(Domain&) { throw Error(XG_ERR_ARG, "dispBeamColumn: not available"); }
void frame_group_add(Domain&, Group&, int, const EleDef&) {}
}
