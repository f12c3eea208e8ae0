// CPU harness for sparseimaging_amd/csrc/deal.h: reads "nblk n w_0 .. w_{n-1}"
// from stdin and prints the dealt table as "ok nb_0 .. nb_{n-1}" followed by
// one "problem index" line per workgroup (or "fail").
#include <cstdio>
#include <vector>

#include "../../sparseimaging_amd/csrc/deal.h"

int main() {
  int nblk = 0, n = 0;
  if (scanf("%d %d", &nblk, &n) != 2) return 2;
  std::vector<double> w(n);
  for (int i = 0; i < n; i++)
    if (scanf("%lf", &w[i]) != 1) return 2;
  std::vector<std::pair<int, int>> blk;
  std::vector<int> nb;
  if (!mf::deal_workgroups(w, nblk, blk, nb)) {
    printf("fail\n");
    return 0;
  }
  printf("ok");
  for (int v : nb) printf(" %d", v);
  printf("\n");
  for (const auto& p : blk) printf("%d %d\n", p.first, p.second);
  return 0;
}
