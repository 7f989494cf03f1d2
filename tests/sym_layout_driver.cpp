// Runs xg::sym_layout on a pattern read from argv[1] and writes the layout to
// argv[2] (int32 binary), for the CPU checks of the block-symmetric SpMV
// layout.  Input: M, nbr, n_nbc, rowptr[M*nbr + 1], nbc[n_nbc].  Output: ok,
// nup, ptr[nbr + 1], col[n_nbc], blk[n_nbc], src[nup], len[nup].
#include <cstdio>
#include <vector>

#include "../xara_amd/csrc/xg_sym_layout.h"

static bool rd(FILE* f, std::vector<int>& v, size_t n) {
  v.resize(n);
  return n == 0 || fread(v.data(), sizeof(int), n, f) == n;
}

int main(int argc, char** argv) {
  if (argc != 3) return 2;
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 2;
  std::vector<int> head, rowptr, nbc;
  if (!rd(f, head, 3) || !rd(f, rowptr, (size_t)head[0] * head[1] + 1) || !rd(f, nbc, head[2])) return 2;
  fclose(f);
  xg::SymLayout L;
  const int ok = xg::sym_layout(head[0], head[1], rowptr, nbc, L) ? 1 : 0;
  FILE* g = fopen(argv[2], "wb");
  if (!g) return 2;
  const int h2[2] = {ok, L.nup};
  fwrite(h2, sizeof(int), 2, g);
  for (auto* v : {&L.ptr, &L.col, &L.blk, &L.src, &L.len}) fwrite(v->data(), sizeof(int), v->size(), g);
  fclose(g);
  return 0;
}
