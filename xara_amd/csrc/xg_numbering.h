// Host-side equation numbering and CSR pattern (see xg_numbering.cpp).
#pragma once
#include <vector>

#include "xg_common.h"

namespace xg {

struct Numbering {
  int neq = 0;
  std::vector<int> eq;      // nnode*ndf, -1 fixed
  std::vector<int> rowptr;  // neq+1
  std::vector<int> colind;  // nnz
};

// fixed: nnode*ndf flags; elem_nodes: node indices of every element in
// ascending element-tag order.
void number_and_pattern(int nnode, int ndf, const std::vector<char>& fixed,
                        const std::vector<std::vector<int>>& elem_nodes, bool use_rcm, Numbering& out);

}  // namespace xg
