// Host check of the bisection-tree index helpers used by the surrogate order
// (mpt_amd/csrc/bisect_dev.hpp), for trees of 1..12 levels:
//   * path_preorder inverts preorder_path;
//   * bfs_path enumerates the top min(h, kPickLevels) levels, each node once,
//     level by level and left to right (pre-order ordinals rise within a
//     level);
//   * skip_picks(k, picks) is the k-th reference ordinal outside the picks,
//     against a brute-force walk, for pick sets drawn over the whole tree.
// Prints "bad=N" and exits non-zero when a check fails.
#include <cstdio>
#include <vector>

#include "bisect_dev.hpp"

int main() {
    using namespace mptg;
    int bad = 0;
    uint64_t rng = 0x9e3779b97f4a7c15ull;
    auto next = [&]() {
        rng ^= rng << 13;
        rng ^= rng >> 7;
        rng ^= rng << 17;
        return rng;
    };
    for (int h = 1; h <= 12; ++h) {
        const uint32_t W = (1u << h) - 1u;
        for (uint32_t o = 0; o < W; ++o) {
            uint32_t p;
            int d;
            preorder_path(h, o, &p, &d);
            if (path_preorder(h, p, d) != o) ++bad;
        }
        const int tl = h < kPickLevels ? h : kPickLevels;
        const uint32_t T = (1u << tl) - 1u;
        std::vector<int> seen(W, 0);
        int last_d = 0;
        long last_o = -1;
        for (uint32_t i = 0; i < T; ++i) {
            uint32_t p;
            int d;
            bfs_path(i, &p, &d);
            const uint32_t o = path_preorder(h, p, d);
            if (d >= tl || o >= W || d < last_d || (d == last_d && (long)o <= last_o)) {
                ++bad;
                continue;
            }
            last_d = d;
            last_o = (long)o;
            seen[o]++;
        }
        for (uint32_t o = 0; o < W; ++o) {
            uint32_t p;
            int d;
            preorder_path(h, o, &p, &d);
            if ((d < tl) != (seen[o] == 1)) ++bad;
        }
        if (W < (uint32_t)kPicks + 1) continue;
        for (int trial = 0; trial < 200; ++trial) {
            // kPicks distinct ordinals, sorted ascending
            uint32_t pk[kPicks];
            int have = 0;
            while (have < kPicks) {
                const uint32_t c = (uint32_t)(next() % W);
                bool dup = false;
                for (int i = 0; i < have; ++i) dup = dup || pk[i] == c;
                if (!dup) pk[have++] = c;
            }
            for (int i = 0; i < kPicks; ++i)
                for (int j = i + 1; j < kPicks; ++j)
                    if (pk[j] < pk[i]) {
                        const uint32_t x = pk[i];
                        pk[i] = pk[j];
                        pk[j] = x;
                    }
            uint32_t k = 0;
            for (uint32_t o = 0; o < W; ++o) {
                bool picked = false;
                for (int i = 0; i < kPicks; ++i) picked = picked || pk[i] == o;
                if (picked) continue;
                if (skip_picks(k, pk) != o) ++bad;
                ++k;
            }
            if (k != W - kPicks) ++bad;
        }
    }
    std::printf("bad=%d\n", bad);
    return bad != 0;
}
