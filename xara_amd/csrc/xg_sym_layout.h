// Block-symmetric layout of a node-blocked CSR pattern (host side, plain
// C++ so that it can be checked without a GPU).
#pragma once
#include <algorithm>
#include <cstdint>
#include <vector>

namespace xg {

struct SymLayout {
  std::vector<int> ptr, col, blk, src, len;
  int nup = 0;
};

// The block rows are swept in an order ord (ord[position] = block row; the
// identity when ord is empty), and the layout is stored in that order: for
// the block row I at position p, entries ptr[p] .. ptr[p+1] list its
// neighbour blocks J in ascending J (col) with the stored block holding each
// (blk).  The stored blocks are the (I, J) blocks with pos(J) >= pos(I),
// numbered position by position in the order of their entries, so that a
// sweep reads them front to back; an entry with pos(J) < pos(I) refers to its
// partner's stored block (J, I), to be used transposed, marked by the sign
// bit of blk (blk = k | 0x80000000).  Stored block k takes CSR entries
// src[k] + r*len[k] + c (r, c < M) of rows M*I + r of its own block row I.
// Returns false when the block pattern is not symmetric (J a neighbour of I
// but not I of J) or ord is not a permutation of the block rows.
constexpr int kSymLower = (int)0x80000000u;
inline bool sym_layout(int M, int nbr, const std::vector<int>& rowptr, const std::vector<int>& nbc,
                       SymLayout& out, const std::vector<int>& ord = {}) {
  out = SymLayout{};
  if (M < 1 || nbr < 0 || rowptr.size() < (size_t)M * nbr + 1) return false;
  if (!ord.empty() && ord.size() != (size_t)nbr) return false;
  auto first = [&](int I) { return rowptr[(size_t)M * I] / (M * M); };
  auto rowlen = [&](int I) { return rowptr[(size_t)M * I + 1] - rowptr[(size_t)M * I]; };
  auto at = [&](int p) { return ord.empty() ? p : ord[p]; };
  std::vector<int> pos(nbr, -1);
  for (int p = 0; p < nbr; p++) {
    const int I = at(p);
    if (I < 0 || I >= nbr || pos[I] >= 0) return false;
    pos[I] = p;
  }
  for (int I = 0; I < nbr; I++) {
    if (rowlen(I) % M || (size_t)(first(I) + rowlen(I) / M) > nbc.size()) return false;
    for (int q = first(I); q < first(I) + rowlen(I) / M; q++)
      if (nbc[q] < 0 || nbc[q] >= nbr) return false;
  }
  // stored blocks per position, and the partners that will read each
  // transposed: (reader block row, stored block) lists per reader
  std::vector<int> ukptr(nbr + 1, 0), lcnt(nbr + 1, 0);
  for (int p = 0; p < nbr; p++) {
    const int I = at(p);
    int up = 0;
    for (int q = first(I); q < first(I) + rowlen(I) / M; q++) {
      up += pos[nbc[q]] >= p;
      if (pos[nbc[q]] > p) lcnt[nbc[q] + 1]++;
    }
    ukptr[p + 1] = ukptr[p] + up;
  }
  for (int I = 0; I < nbr; I++) lcnt[I + 1] += lcnt[I];
  const int nup = ukptr[nbr];
  out.src.assign(nup, 0);
  out.len.assign(nup, 0);
  std::vector<int> lJ(lcnt[nbr]), lk(lcnt[nbr]), lfill(lcnt.begin(), lcnt.end() - 1);
  for (int p = 0; p < nbr; p++) {
    const int I = at(p), f = first(I), L = rowlen(I);
    int k = ukptr[p];
    for (int q = f; q < f + L / M; q++) {
      const int J = nbc[q];
      if (pos[J] < p) continue;
      out.src[k] = rowptr[(size_t)M * I] + M * (q - f);
      out.len[k] = L;
      if (pos[J] > p) {
        const int w = lfill[J]++;
        lJ[w] = I;
        lk[w] = k;
      }
      k++;
    }
  }
  out.ptr.assign(nbr + 1, 0);
  out.col.reserve(nbc.size());
  out.blk.reserve(nbc.size());
  size_t e = 0;
  for (int p = 0; p < nbr; p++) {
    const int I = at(p), f = first(I);
    int k = ukptr[p], found = 0;
    for (int q = f; q < f + rowlen(I) / M; q++, e++) {
      const int J = nbc[q];
      out.col.push_back(J);
      if (pos[J] < p) {
        int w = lcnt[I];
        while (w < lcnt[I + 1] && lJ[w] != J) w++;
        if (w == lcnt[I + 1]) return false;
        out.blk.push_back(lk[w] | kSymLower);
        found++;
      } else {
        out.blk.push_back(k++);
      }
    }
    if (found != lcnt[I + 1] - lcnt[I]) return false;
    out.ptr[p + 1] = (int)e;
  }
  out.nup = nup;
  return true;
}

// Traversal order of the block rows for the symmetric product: block rows
// sorted by the Morton (Z-order) key of their node's coordinates, so that
// block rows close in space are close in the sweep.  A stored block is read
// twice per product, by its own block row and by its partner's; with rows
// swept in this order (and every XCD sweeping one front of them) the second
// read mostly finds the block still in the XCD's L2, where in the
// equation-number order the partner lies a whole grid plane away.  xyz:
// dim (1..3) coordinates per block row; coordinates are quantised to 21 bits
// per axis over their bounding box; equal keys keep block-row order.
// Returns ord with ord[position] = block row.
inline std::vector<int> morton_order(int nbr, int dim, const std::vector<double>& xyz) {
  std::vector<int> ord(std::max(nbr, 0));
  for (int i = 0; i < nbr; i++) ord[i] = i;
  if (nbr <= 1 || dim < 1 || dim > 3 || xyz.size() < (size_t)nbr * dim) return ord;
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  for (int d = 0; d < dim; d++) {
    lo[d] = hi[d] = xyz[d];
    for (int i = 1; i < nbr; i++) {
      lo[d] = std::min(lo[d], xyz[(size_t)i * dim + d]);
      hi[d] = std::max(hi[d], xyz[(size_t)i * dim + d]);
    }
  }
  constexpr int kBits = 21;
  constexpr double kMax = (double)((1u << kBits) - 1);
  std::vector<uint64_t> key(nbr, 0);
  for (int i = 0; i < nbr; i++) {
    uint64_t k = 0;
    for (int d = 0; d < dim; d++) {
      const double span = hi[d] - lo[d];
      const double f = span > 0 ? (xyz[(size_t)i * dim + d] - lo[d]) / span : 0.0;
      const uint64_t q = (uint64_t)std::min(kMax, std::max(0.0, f * kMax + 0.5));
      for (int b = 0; b < kBits; b++) k |= ((q >> b) & 1ull) << (dim * b + d);
    }
    key[i] = k;
  }
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return key[a] < key[b]; });
  return ord;
}

}  // namespace xg
