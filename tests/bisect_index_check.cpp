// Host check of the bisection-tree index helpers used by the surrogate order
// (mpt_amd/csrc/bisect_dev.hpp), for trees of 1..12 levels:
//   * path_preorder inverts preorder_path;
//   * bfs_path enumerates the top min(h, kPickLevels) levels, each node once,
//     level by level and left to right (pre-order ordinals rise within a
//     level);
//   * skip_picks(k, picks) is the k-th reference ordinal outside the picks,
//     against a brute-force walk, for pick sets drawn over the whole tree;
//   * coord_mid equals the reference's midpoint recursion (lo + hi) / 2
//     along the path (demo/link_manipulator_scenario.hpp:118-138) bit for
//     bit, for random endpoints, paths and depths 0..20.
// Prints "bad=N" and exits non-zero when a check fails.
#include <cstdio>
#include <cstring>
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
    // coord_mid against the recursion as the reference writes it
    auto recursion = [](double lo, double hi, uint32_t path, int depth) {
        for (int k = 0; k < depth; ++k) {
            const double m = (lo + hi) / 2;
            if ((path >> k) & 1u) lo = m;
            else hi = m;
        }
        return (lo + hi) / 2;
    };
    for (int t = 0; t < 200000; ++t) {
        auto unit = [&]() { return (double)(next() >> 11) * 0x1.0p-53; };
        const double a = (unit() * 2.0 - 1.0) * 3.2;
        double b = (unit() * 2.0 - 1.0) * 3.2;
        if (t % 7 == 0) b = a + (unit() - 0.5) * 1e-9;  // nearly equal ends
        if (t % 11 == 0) b = -a;
        const int depth = (int)(next() % 21);
        const uint32_t path = depth ? (uint32_t)next() & ((1u << depth) - 1u) : 0u;
        const double x = coord_mid(a, b, path, depth), y = recursion(a, b, path, depth);
        uint64_t xb, yb;
        std::memcpy(&xb, &x, 8);
        std::memcpy(&yb, &y, 8);
        if (xb != yb) ++bad;
    }
    std::printf("bad=%d\n", bad);
    return bad != 0;
}
