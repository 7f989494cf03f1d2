// Equation numbering and CSR pattern of the fixed-pattern system, computed on
// the host at setup (domainChanged).  The results are bit-identical to the
// reference chain PlainHandler::handle (PlainHandler.cpp:60-247) ->
// AnalysisModel::getDOFGroupGraph (AnalysisModel.cpp:355-398) ->
// RCM::number(graph, -1) (RCM.cpp:67-259, GPS = false) ->
// DOF_Numberer::numberDOF (DOF_Numberer.cpp:93-205) ->
// AnalysisModel::getDOFGraph (AnalysisModel.cpp:286-352) ->
// SparseGenRowLinSOE::setSize (SparseGenRowLinSOE.cpp:129-234), but built
// with sorted edge lists instead of per-vertex insertion so a 274k-node mesh
// numbers in well under a second.
#include "xg_numbering.h"

#include <algorithm>
#include <cstdint>

namespace xg {

// Node adjacency (sorted, unique, no self loops) from element connectivity.
static void node_graph(int nnode, const std::vector<std::vector<int>>& elem_nodes, std::vector<int>& ptr,
                       std::vector<int>& adj, std::vector<char>& in_element) {
  std::vector<uint64_t> edges;
  size_t est = 0;
  for (auto& en : elem_nodes) est += en.size() * (en.size() - 1);
  edges.reserve(est);
  in_element.assign(nnode, 0);
  for (auto& en : elem_nodes) {
    for (int a : en) in_element[a] = 1;
    for (size_t i = 0; i < en.size(); i++)
      for (size_t j = 0; j < en.size(); j++)
        if (en[i] != en[j]) edges.push_back(((uint64_t)(uint32_t)en[i] << 32) | (uint32_t)en[j]);
  }
  std::sort(edges.begin(), edges.end());
  edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
  ptr.assign(nnode + 1, 0);
  adj.resize(edges.size());
  for (size_t k = 0; k < edges.size(); k++) {
    ptr[(edges[k] >> 32) + 1]++;
    adj[k] = (int)(edges[k] & 0xffffffffu);
  }
  for (int v = 0; v < nnode; v++) ptr[v + 1] += ptr[v];
}

// Reverse Cuthill-McKee exactly as RCM::number(graph, -1): start from the
// first vertex of the (ascending-tag) iterator, visit neighbours in
// ascending tag, fill the result from the back, and on exhausting a
// component continue with the next unvisited vertex of the iterator.
static std::vector<int> rcm(int n, const std::vector<int>& ptr, const std::vector<int>& adj) {
  std::vector<int> result(n), tmp(n, -1);
  if (n == 0) return result;
  int iter = 0;
  int currentMark = n - 1, nextMark = currentMark - 1;
  result[currentMark] = 0;
  tmp[0] = currentMark;
  while (nextMark >= 0) {
    const int v = result[currentMark];
    for (int k = ptr[v]; k < ptr[v + 1]; k++) {
      const int w = adj[k];
      if (tmp[w] == -1) {
        tmp[w] = nextMark;
        result[nextMark--] = w;
      }
    }
    currentMark--;
    if (currentMark == nextMark && currentMark >= 0) {
      int vp = -1;
      while (iter < n) {
        const int c = iter++;
        if (tmp[c] == -1) {
          vp = c;
          break;
        }
      }
      nextMark--;
      tmp[vp] = currentMark;
      result[currentMark] = vp;
    }
  }
  return result;
}

void number_and_pattern(int nnode, int ndf, const std::vector<char>& fixed,
                        const std::vector<std::vector<int>>& elem_nodes, bool use_rcm, Numbering& out) {
  std::vector<int> ptr, adj;
  std::vector<char> in_el;
  node_graph(nnode, elem_nodes, ptr, adj, in_el);
  std::vector<int> order;
  if (use_rcm) order = rcm(nnode, ptr, adj);
  else {
    order.resize(nnode);
    for (int i = 0; i < nnode; i++) order[i] = i;
  }
  out.eq.assign((size_t)nnode * ndf, -1);
  int eqn = 0;
  for (int i = 0; i < nnode; i++) {
    const int v = order[i];
    for (int j = 0; j < ndf; j++)
      if (!fixed[(size_t)v * ndf + j]) out.eq[(size_t)v * ndf + j] = eqn++;
  }
  out.neq = eqn;
  // equation -> owning node
  std::vector<int> eq_node(eqn);
  for (int v = 0; v < nnode; v++)
    for (int j = 0; j < ndf; j++) {
      const int a = out.eq[(size_t)v * ndf + j];
      if (a >= 0) eq_node[a] = v;
    }
  // Row a: sorted({a} U free equations of every node sharing an element with
  // node(a), including node(a) itself).  A node in no element couples only
  // with itself.
  out.rowptr.assign(eqn + 1, 0);
  for (int a = 0; a < eqn; a++) {
    const int v = eq_node[a];
    long long cnt = 0;
    if (in_el[v]) {
      for (int j = 0; j < ndf; j++) cnt += out.eq[(size_t)v * ndf + j] >= 0;
      for (int k = ptr[v]; k < ptr[v + 1]; k++)
        for (int j = 0; j < ndf; j++) cnt += out.eq[(size_t)adj[k] * ndf + j] >= 0;
    } else
      cnt = 1;
    out.rowptr[a + 1] = (int)cnt;
  }
  long long total = 0;
  for (int a = 0; a < eqn; a++) {
    total += out.rowptr[a + 1];
    if (total > 2147483647LL) throw Error(-1, "CSR pattern exceeds 2^31 nonzeros");
    out.rowptr[a + 1] = (int)total;
  }
  out.colind.resize(total);
  for (int a = 0; a < eqn; a++) {
    const int v = eq_node[a];
    int* c = out.colind.data() + out.rowptr[a];
    int m = 0;
    if (in_el[v]) {
      for (int j = 0; j < ndf; j++) {
        const int b = out.eq[(size_t)v * ndf + j];
        if (b >= 0) c[m++] = b;
      }
      for (int k = ptr[v]; k < ptr[v + 1]; k++)
        for (int j = 0; j < ndf; j++) {
          const int b = out.eq[(size_t)adj[k] * ndf + j];
          if (b >= 0) c[m++] = b;
        }
      std::sort(c, c + m);
    } else
      c[m++] = a;
  }
}

}  // namespace xg
