// Block-symmetric layout of a node-blocked CSR pattern (host side, plain
// C++ so that it can be checked without a GPU).
#pragma once
#include <vector>

namespace xg {

// Stored blocks are the upper M x M blocks (I, J >= I), numbered block row
// by block row, J ascending.  For block row I, entries ptr[I] .. ptr[I+1]
// list its neighbour blocks J ascending (col) with the stored block holding
// each (blk): (I, J) itself for J >= I, (J, I) for J < I.  Stored block k
// takes CSR entries src[k] + r*len[k] + c (r, c < M) of rows M*I + r.
struct SymLayout {
  std::vector<int> ptr, col, blk, src, len;
  int nup = 0;
};

// rowptr: CSR row pointers of neq = M*nbr equations whose block row I lists
// its neighbour blocks as ascending runs of M columns, the same in each of
// its M rows (rowptr[M*I] = M*M * first entry of I in nbc); nbc: one block
// column per (block row, neighbour).  Returns false when the block pattern
// is not symmetric (J a neighbour of I but not I of J).
inline bool sym_layout(int M, int nbr, const std::vector<int>& rowptr, const std::vector<int>& nbc,
                       SymLayout& out) {
  out = SymLayout{};
  if (M < 1 || nbr < 0 || rowptr.size() < (size_t)M * nbr + 1) return false;
  auto first = [&](int I) { return rowptr[(size_t)M * I] / (M * M); };
  auto rowlen = [&](int I) { return rowptr[(size_t)M * I + 1] - rowptr[(size_t)M * I]; };
  for (int I = 0; I < nbr; I++)
    if (rowlen(I) % M || (size_t)(first(I) + rowlen(I) / M) > nbc.size()) return false;
  std::vector<int> ukptr(nbr + 1, 0);
  for (int I = 0; I < nbr; I++) {
    int up = 0;
    for (int q = first(I); q < first(I) + rowlen(I) / M; q++) up += nbc[q] >= I;
    ukptr[I + 1] = ukptr[I] + up;
  }
  const int nup = ukptr[nbr];
  out.src.assign(nup, 0);
  out.len.assign(nup, 0);
  // lower partners of each block row: (J, stored block of (J, I)), J < I ascending
  std::vector<int> lcnt(nbr + 1, 0);
  for (int J = 0; J < nbr; J++)
    for (int q = first(J); q < first(J) + rowlen(J) / M; q++) {
      const int I = nbc[q];
      if (I < 0 || I >= nbr) return false;
      if (I > J) lcnt[I + 1]++;
    }
  for (int I = 0; I < nbr; I++) lcnt[I + 1] += lcnt[I];
  std::vector<int> lJ(lcnt[nbr]), lk(lcnt[nbr]), lfill(lcnt.begin(), lcnt.end() - 1);
  for (int J = 0; J < nbr; J++) {
    const int f = first(J), L = rowlen(J);
    int k = ukptr[J];
    for (int q = f; q < f + L / M; q++) {
      if (nbc[q] < J) continue;
      out.src[k] = rowptr[(size_t)M * J] + M * (q - f);
      out.len[k] = L;
      if (nbc[q] > J) {
        const int I = nbc[q], w = lfill[I]++;
        lJ[w] = J;
        lk[w] = k;
      }
      k++;
    }
  }
  out.ptr.assign(nbr + 1, 0);
  size_t e = 0;
  for (int I = 0; I < nbr; I++) {
    const int f = first(I);
    int w = lcnt[I], k = ukptr[I];
    for (int q = f; q < f + rowlen(I) / M; q++, e++) {
      const int J = nbc[q];
      out.col.push_back(J);
      if (J < I) {
        if (w >= lcnt[I + 1] || lJ[w] != J) return false;
        out.blk.push_back(lk[w++]);
      } else {
        out.blk.push_back(k++);
      }
    }
    if (w != lcnt[I + 1]) return false;
    out.ptr[I + 1] = (int)e;
  }
  out.nup = nup;
  return true;
}

}  // namespace xg
