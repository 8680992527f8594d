// Host check of the bisection-tree index helpers used by the surrogate order
// (mpt_amd/csrc/bisect_dev.hpp), for trees of 1..12 levels:
//   * path_preorder inverts preorder_path;
//   * bfs_path covers the top min(h, kTopLevels) levels, deep_path the rest,
//     and together they visit every node exactly once;
//   * deep_path yields the deeper nodes in ascending pre-order.
// Prints "bad=N" and exits non-zero when a check fails.
#include <cstdio>
#include <vector>

#include "bisect_dev.hpp"

int main() {
    using namespace mptg;
    int bad = 0;
    for (int h = 1; h <= 12; ++h) {
        const uint32_t W = (1u << h) - 1u;
        std::vector<int> seen(W, 0);
        for (uint32_t o = 0; o < W; ++o) {
            uint32_t p;
            int d;
            preorder_path(h, o, &p, &d);
            if (path_preorder(h, p, d) != o) ++bad;
        }
        const int tl = h < kTopLevels ? h : kTopLevels;
        const uint32_t T = (1u << tl) - 1u;
        for (uint32_t i = 0; i < T; ++i) {
            uint32_t p;
            int d;
            bfs_path(i, &p, &d);
            const uint32_t o = path_preorder(h, p, d);
            if (d >= tl || o >= W) {
                ++bad;
                continue;
            }
            seen[o]++;
        }
        long prev = -1;
        for (uint32_t k = 0; k + T < W; ++k) {
            uint32_t p;
            int d;
            deep_path(h, k, &p, &d);
            const uint32_t o = path_preorder(h, p, d);
            if (d < kTopLevels || d >= h || o >= W || (long)o <= prev) {
                ++bad;
                continue;
            }
            prev = (long)o;
            seen[o]++;
        }
        for (uint32_t o = 0; o < W; ++o)
            if (seen[o] != 1) ++bad;
    }
    std::printf("bad=%d\n", bad);
    return bad != 0;
}
