// Host-side partition of the batch's work lists (plain C++ on HIP's vector
// types, so that it can be checked on the CPU).
//
// A list holds (solve, tile) items in solve order, each with an estimated
// cost.  It is split into 8 contiguous chunks of about equal cost, one per
// XCD; workgroup b of an XCD's G8 workgroups takes positions b, b + G8,
// b + 2 G8, ... of its chunk, so the workgroups of an XCD sweep a solve's
// neighbouring tiles together and share their halos in that XCD's L2.
#pragma once
#include <algorithm>
#include <numeric>
#include <vector>

#include <hip/hip_runtime.h>

namespace dgmi {

// cut[0..parts]: `parts` contiguous ranges of [0, n) of about equal cost
inline void cost_cuts(const double* cost, int n, int parts, int* cut) {
    cut[0] = 0;
    for (int k = 1; k <= parts; k++) cut[k] = n;
    const double total = std::accumulate(cost, cost + n, 0.0);
    double run = 0;
    int x = 1;
    for (int i = 0; i < n && x < parts; i++) {
        run += cost[i];
        while (x < parts && run >= total * x / parts) cut[x++] = i + 1;
    }
}

// A list appended to out in its 8 XCD chunks; bounds[0..8] = the chunks'
// start positions in out and the end of the last.
inline void xcd_chunks(const std::vector<int2>& items, const std::vector<double>& cost, std::vector<int2>& out,
                       int* bounds) {
    int cut[9];
    cost_cuts(cost.data(), (int)items.size(), 8, cut);
    const int base = (int)out.size();
    for (int k = 0; k <= 8; k++) bounds[k] = base + cut[k];
    out.insert(out.end(), items.begin(), items.end());
}

}  // namespace dgmi
