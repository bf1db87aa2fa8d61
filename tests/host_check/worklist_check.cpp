// CPU check of the batch's work-list partition (dg_worklist.h).  It walks
// each XCD chunk as the persistent kernels do (workgroup b of G8 takes
// positions b, b + G8, ...) and checks that every item is taken exactly
// once, that the list keeps its order and that the chunks are contiguous
// and of about equal cost.
#include <cstdio>
#include <random>

#include "dg_worklist.h"

using namespace dgmi;

static int fail(const char* what, int n, int G8) {
    std::printf("FAIL %s (n=%d G8=%d)\n", what, n, G8);
    return 1;
}

static int check(int n, int G8, std::mt19937& rng) {
    std::vector<int2> items;
    std::vector<double> cost;
    int slot = 0;
    while ((int)items.size() < n) {
        const int len = 1 + (int)(rng() % 300);
        const double c = (rng() & 1) ? 1.0 : 2.7;
        for (int t = 0; t < len && (int)items.size() < n; t++) {
            items.push_back(int2{slot, t});
            cost.push_back(c);
        }
        slot++;
    }
    std::vector<int2> out;
    out.push_back(int2{-7, -7});  // a list is appended after earlier lists
    int bounds[9];
    xcd_chunks(items, cost, out, bounds);
    if (bounds[0] != 1 || bounds[8] != (int)out.size() || (int)out.size() != n + 1) return fail("bounds", n, G8);
    const double total = std::accumulate(cost.begin(), cost.end(), 0.0);
    std::vector<int> seen(n, 0);
    for (int x = 0; x < 8; x++) {
        const int i0 = bounds[x], i1 = bounds[x + 1];
        if (i0 > i1) return fail("chunk order", n, G8);
        double chunk = 0;
        for (int it = i0; it < i1; it++) {
            if (out[it].x != items[it - 1].x || out[it].y != items[it - 1].y) return fail("order kept", n, G8);
            chunk += cost[it - 1];
        }
        if (chunk > total / 8 + 2.7 + 1e-9) return fail("chunk cost", n, G8);
        for (int b = 0; b < G8; b++)
            for (int it = i0 + b; it < i1; it += G8) seen[it - 1]++;
    }
    for (int i = 0; i < n; i++)
        if (seen[i] != 1) return fail("each item once", n, G8);
    return 0;
}

int main() {
    std::mt19937 rng(12345);
    int bad = 0, cases = 0;
    for (int n : {0, 1, 7, 8, 9, 100, 777, 5000, 17273})
        for (int G8 : {1, 2, 8, 64, 96, 128, 192, 224}) {
            bad += check(n, G8, rng);
            cases++;
        }
    if (bad) return 1;
    std::printf("worklist partition ok: %d cases\n", cases);
    return 0;
}
