// Reads a symmetric matrix (int32 n, then n*n doubles row-major) from argv[1],
// writes its eigenvalues (n doubles) then eigenvectors (n*n, row-major,
// vectors as columns) to argv[2] with abh::sym_eig.
#include <cstdio>
#include <vector>

#include "../../aboria_amd/csrc/symeig.hpp"

int main(int argc, char **argv) {
  if (argc != 3) return 2;
  FILE *f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  int n = 0;
  if (std::fread(&n, 4, 1, f) != 1) return 2;
  std::vector<double> a(size_t(n) * n), w, v;
  if (std::fread(a.data(), 8, a.size(), f) != a.size()) return 2;
  std::fclose(f);
  if (!abh::sym_eig(n, a, w, v)) return 3;
  FILE *g = std::fopen(argv[2], "wb");
  std::fwrite(w.data(), 8, w.size(), g);
  std::fwrite(v.data(), 8, v.size(), g);
  std::fclose(g);
  return 0;
}
