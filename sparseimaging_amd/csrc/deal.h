// deal.h -- dealing the workgroups of one cooperative launch to the problems
// of a batch (host-only, no HIP types, so that a CPU test can compile it).
//
// Workgroup b of a launch runs on XCD b mod kXcds (the dispatcher deals
// workgroups to the XCDs round-robin).  Each problem goes whole to the XCD with
// the least work so far, so its data stays in one L2, and each XCD's
// workgroups are split among its problems in proportion to their work per
// iteration (at least one each).  With fewer problems than XCDs each problem
// spans several XCDs, interleaved.
#pragma once

#include <algorithm>
#include <utility>
#include <vector>

namespace mf {

constexpr int kXcds = 8;

// work[i]: work per iteration of problem i (any positive unit); nblk >= the
// number of problems.  On return blk[b] = (problem, index within the problem)
// or (-1, 0) for an undealt workgroup, and nb[i] = workgroups of problem i.
// Returns false (nothing dealt) if there are more problems than workgroups.
inline bool deal_workgroups(const std::vector<double>& work, int nblk, std::vector<std::pair<int, int>>& blk,
                            std::vector<int>& nb) {
  const int pn = (int)work.size();
  blk.assign(nblk > 0 ? nblk : 0, std::make_pair(-1, 0));
  nb.assign(pn, 0);
  if (pn == 0) return true;
  if (nblk < pn) return false;
  std::vector<int> order(pn);
  for (int i = 0; i < pn; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work[a] > work[b]; });
  if (pn >= kXcds) {
    // the slots of XCD x: workgroups x, x + kXcds, ...; their count is the cap
    // on the problems that XCD can take (nblk need not be a multiple of kXcds)
    std::vector<std::vector<int>> slots(kXcds);
    for (int b = 0; b < nblk; b++) slots[b % kXcds].push_back(b);
    std::vector<double> load(kXcds, 0.0);
    std::vector<int> cnt(kXcds, 0), xcd(pn, -1);
    for (int i : order) {
      int best = -1;
      for (int x = 0; x < kXcds; x++)
        if (cnt[x] < (int)slots[x].size() && (best < 0 || load[x] < load[best])) best = x;
      if (best < 0) return false;  // cannot happen: the caps add up to nblk >= pn
      xcd[i] = best;
      load[best] += work[i];
      cnt[best]++;
    }
    for (int x = 0; x < kXcds; x++) {
      std::vector<int> mine;
      for (int i = 0; i < pn; i++)
        if (xcd[i] == x) mine.push_back(i);
      if (mine.empty()) continue;
      const int avail = (int)slots[x].size();
      const int left = avail - (int)mine.size();
      int used = 0;
      for (int i : mine) {
        nb[i] = 1 + (int)((double)left * work[i] / std::max(load[x], 1e-300));
        used += nb[i];
      }
      // rounding down leaves used <= avail; hand out the rest one by one
      for (size_t k = 0; used < avail; k = (k + 1) % mine.size(), used++) nb[mine[k]]++;
      int s0 = 0;
      for (int i : mine) {
        for (int j = 0; j < nb[i]; j++) blk[slots[x][s0 + j]] = std::make_pair(i, j);
        s0 += nb[i];
      }
    }
  } else {
    double wt = 0;
    for (int i = 0; i < pn; i++) wt += work[i];
    int used = 0;
    for (int i = 0; i < pn; i++) {
      nb[i] = std::max(1, (int)((double)nblk * work[i] / std::max(wt, 1e-300)));
      used += nb[i];
    }
    while (used > nblk)
      for (int i = 0; i < pn && used > nblk; i++)
        if (nb[i] > 1) nb[i]--, used--;
    std::vector<int> next(pn, 0);
    int b = 0;
    for (bool any = true; any;) {
      any = false;
      for (int i = 0; i < pn; i++)
        if (next[i] < nb[i]) {
          blk[b++] = std::make_pair(i, next[i]++);
          any = true;
        }
    }
  }
  return true;
}

}  // namespace mf
