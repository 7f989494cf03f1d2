// CPU driver for the host code of the block-symmetric SpMV layout
// (xg_sym_layout.h).
//   sym_layout_driver IN OUT [ORD]: runs xg::sym_layout on a pattern read
//     from IN and writes the layout to OUT (int32 binary).  IN: M, nbr,
//     n_nbc, rowptr[M*nbr + 1], nbc[n_nbc]; ORD (optional): nbr int32, the
//     sweep order.  OUT: ok, nup, ptr[nbr + 1], col[n_nbc], blk[n_nbc],
//     src[nup], len[nup].
//   sym_layout_driver --morton IN OUT: IN: nbr, dim (int32), then nbr*dim
//     doubles; OUT: xg::morton_order, nbr int32.
#include <cstdio>
#include <cstring>
#include <vector>

#include "../xara_amd/csrc/xg_sym_layout.h"

template <class T>
static bool rd(FILE* f, std::vector<T>& v, size_t n) {
  v.resize(n);
  return n == 0 || fread(v.data(), sizeof(T), n, f) == n;
}

static int morton(const char* in, const char* outp) {
  FILE* f = fopen(in, "rb");
  if (!f) return 2;
  std::vector<int> head;
  std::vector<double> xyz;
  if (!rd(f, head, 2) || !rd(f, xyz, (size_t)head[0] * head[1])) return 2;
  fclose(f);
  const std::vector<int> ord = xg::morton_order(head[0], head[1], xyz);
  FILE* g = fopen(outp, "wb");
  if (!g) return 2;
  fwrite(ord.data(), sizeof(int), ord.size(), g);
  fclose(g);
  return 0;
}

int main(int argc, char** argv) {
  if (argc == 4 && !strcmp(argv[1], "--morton")) return morton(argv[2], argv[3]);
  if (argc != 3 && argc != 4) return 2;
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 2;
  std::vector<int> head, rowptr, nbc, ord;
  if (!rd(f, head, 3) || !rd(f, rowptr, (size_t)head[0] * head[1] + 1) || !rd(f, nbc, head[2])) return 2;
  fclose(f);
  if (argc == 4) {
    FILE* o = fopen(argv[3], "rb");
    if (!o || !rd(o, ord, head[1])) return 2;
    fclose(o);
  }
  xg::SymLayout L;
  const int ok = xg::sym_layout(head[0], head[1], rowptr, nbc, L, ord) ? 1 : 0;
  FILE* g = fopen(argv[2], "wb");
  if (!g) return 2;
  const int h2[2] = {ok, L.nup};
  fwrite(h2, sizeof(int), 2, g);
  for (auto* v : {&L.ptr, &L.col, &L.blk, &L.src, &L.len}) fwrite(v->data(), sizeof(int), v->size(), g);
  fclose(g);
  return 0;
}
