// SE(3) rigid body among triangle meshes: the reference's valid(q) =
// !fcl::collide(robot at pose(q), env) (demo/se3_rigid_body_scenario.hpp:288-312)
// and its DiscreteMotionValidator link (:330-334), on the device.
//
// The predicate is mesh_hd.hpp's: posed robot triangle a and environment
// triangle b collide iff their boxes overlap and the separating-axis test
// finds no separating axis.  Culling below only skips pairs whose boxes are
// apart, so the verdict is the predicate's exactly.
//
// Layout in HBM (built on the host at upload):
//   env BVH   binned-SAH binary tree (<= 4 triangles per leaf, depth <= 38)
//             collapsed to MeshNode4[nnodes]: 128 B per node holding the
//             boxes (by axis) and references of up to four nodes two levels
//             below it, so one uniform load serves two levels of the walk
//   env tris  16 floats (64 B) per triangle in leaf order: v0 v1 v2, box lo, hi
//   robot     9 floats per triangle, Morton-sorted by centroid so that the 4
//             consecutive triangles a lane owns are close together
// Kernels (one 64-lane wave per block, one item per wave at a time):
//   mesh_valid  one state per wave
//   mesh_link   one edge per wave; waypoints in DMV order one at a time
//               (reference checks_done counts), early exit on the first hit
// Per state, lane l poses robot triangles [4l, 4l + 4) of each 256-triangle
// chunk and takes their box; the wave walks the env BVH depth-first once with
// one LDS stack, visiting a node when any lane's box overlaps it, and stops
// at the first leaf where some lane finds a hit.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "mesh_hd.hpp"
#include "space_dev.hpp"

namespace mptg {

namespace {

constexpr int kLaneTris = 4;                    // robot triangles per lane
constexpr int kChunkTris = 64 * kLaneTris;      // robot triangles per wave pass
constexpr int kBuildDepth = 38;                 // BVH depth bound
// LDS stack of a wave's walk: up to three (a, b) entries per four-wide level
constexpr int kStackSlots = 2 * 3 * (kBuildDepth / 2 + 1);
constexpr int kMaxDmvWaypointsMesh = 1 << 24;   // same limit as the other DMV kernels
constexpr int kTableMaxMesh = 255;

// ---------------------------------------------------------------------------
// host: BVH build
// ---------------------------------------------------------------------------
struct BuildRef {
    float lo[3], hi[3], c[3];
    int32_t id;
};

struct BvhBuilder {
    std::vector<BuildRef> refs;
    std::vector<MeshNode> nodes;

    static void grow(float* lo, float* hi, const float* blo, const float* bhi) {
        for (int i = 0; i < 3; ++i) {
            lo[i] = std::min(lo[i], blo[i]);
            hi[i] = std::max(hi[i], bhi[i]);
        }
    }
    static float area(const float* lo, const float* hi) {
        const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (!(dx >= 0)) return 0.0f;
        return dx * dy + dy * dz + dz * dx;
    }

    int build(int begin, int end, int depth) {
        const int id = (int)nodes.size();
        nodes.push_back(MeshNode{});
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int i = begin; i < end; ++i) {
            grow(lo, hi, refs[i].lo, refs[i].hi);
            grow(clo, chi, refs[i].c, refs[i].c);
        }
        auto set_box = [&](MeshNode& n) {
            for (int i = 0; i < 3; ++i) {
                n.lo[i] = lo[i];
                n.hi[i] = hi[i];
            }
        };
        const int count = end - begin;
        if (count <= kLeafTris) {
            MeshNode& n = nodes[id];
            set_box(n);
            n.a = begin;
            n.b = -count;
            return id;
        }
        int axis = 0;
        for (int i = 1; i < 3; ++i)
            if (chi[i] - clo[i] > chi[axis] - clo[axis]) axis = i;
        int mid = -1;
        const float ext = chi[axis] - clo[axis];
        if (depth < 20 && ext > 0) {
            constexpr int kBins = 16;
            int bn[kBins] = {};
            float blo[kBins][3], bhi[kBins][3];
            for (int b = 0; b < kBins; ++b)
                for (int i = 0; i < 3; ++i) {
                    blo[b][i] = INFINITY;
                    bhi[b][i] = -INFINITY;
                }
            auto bin_of = [&](const BuildRef& r) {
                int b = (int)((r.c[axis] - clo[axis]) / ext * kBins);
                return std::min(std::max(b, 0), kBins - 1);
            };
            for (int i = begin; i < end; ++i) {
                int b = bin_of(refs[i]);
                bn[b]++;
                grow(blo[b], bhi[b], refs[i].lo, refs[i].hi);
            }
            float best = INFINITY;
            int best_split = -1;
            for (int s = 1; s < kBins; ++s) {
                float llo[3] = {INFINITY, INFINITY, INFINITY}, lhi[3] = {-INFINITY, -INFINITY, -INFINITY};
                float rlo[3] = {INFINITY, INFINITY, INFINITY}, rhi[3] = {-INFINITY, -INFINITY, -INFINITY};
                int nl = 0, nr = 0;
                for (int b = 0; b < s; ++b) {
                    nl += bn[b];
                    if (bn[b]) grow(llo, lhi, blo[b], bhi[b]);
                }
                for (int b = s; b < kBins; ++b) {
                    nr += bn[b];
                    if (bn[b]) grow(rlo, rhi, blo[b], bhi[b]);
                }
                if (nl == 0 || nr == 0) continue;
                const float cost = area(llo, lhi) * nl + area(rlo, rhi) * nr;
                if (cost < best) {
                    best = cost;
                    best_split = s;
                }
            }
            if (best_split > 0) {
                auto it = std::partition(refs.begin() + begin, refs.begin() + end,
                                         [&](const BuildRef& r) { return bin_of(r) < best_split; });
                mid = (int)(it - refs.begin());
            }
        }
        if (mid <= begin || mid >= end) {  // median split: depth stays logarithmic
            mid = begin + count / 2;
            std::nth_element(refs.begin() + begin, refs.begin() + mid, refs.begin() + end,
                             [&](const BuildRef& x, const BuildRef& y) {
                                 return x.c[axis] < y.c[axis] || (x.c[axis] == y.c[axis] && x.id < y.id);
                             });
        }
        const int left = build(begin, mid, depth + 1);
        const int right = build(mid, end, depth + 1);
        MeshNode& n = nodes[id];
        set_box(n);
        n.a = left;
        n.b = right;
        return id;
    }
};

int bvh_depth(const std::vector<MeshNode>& nodes, int n, int d) {
    if (nodes[n].b <= 0) return d;
    return std::max(bvh_depth(nodes, nodes[n].a, d + 1), bvh_depth(nodes, nodes[n].b, d + 1));
}

// Four-wide node: slot k holds the box and reference of a binary node one or
// two levels below (the children, with an internal child replaced by its two
// children).  b[k] > 0: a[k] is a MeshNode4 index; b[k] <= 0: triangles
// [a[k], a[k] - b[k]); an unused slot has an empty box (lo > hi).
struct MeshNode4 {
    float lo[3][4], hi[3][4];
    int32_t a[4], b[4];
};
static_assert(sizeof(MeshNode4) == 128, "four-wide node is 128 bytes");

// Collapse the binary subtree under internal node id into out (pre-order).
int collapse4(const std::vector<MeshNode>& n2, int id, std::vector<MeshNode4>& out) {
    const int me = (int)out.size();
    out.push_back(MeshNode4{});
    int slot[4], ns = 0;
    for (const int c : {n2[id].a, n2[id].b}) {
        if (n2[c].b > 0) {
            slot[ns++] = n2[c].a;
            slot[ns++] = n2[c].b;
        } else {
            slot[ns++] = c;
        }
    }
    MeshNode4 node;
    for (int k = 0; k < 4; ++k) {
        for (int i = 0; i < 3; ++i) {
            node.lo[i][k] = INFINITY;
            node.hi[i][k] = -INFINITY;
        }
        node.a[k] = 0;
        node.b[k] = 0;
    }
    for (int k = 0; k < ns; ++k) {
        const MeshNode& g = n2[slot[k]];
        for (int i = 0; i < 3; ++i) {
            node.lo[i][k] = g.lo[i];
            node.hi[i][k] = g.hi[i];
        }
        if (g.b > 0) {
            node.a[k] = collapse4(n2, slot[k], out);
            node.b[k] = 1;
        } else {
            node.a[k] = g.a;
            node.b[k] = g.b;
        }
    }
    out[(size_t)me] = node;
    return me;
}

// The four-wide tree of a binary BVH (root first); a leaf root becomes one
// node with a single slot.
std::vector<MeshNode4> collapse_bvh(const std::vector<MeshNode>& n2) {
    std::vector<MeshNode4> out;
    if (n2.empty()) return out;
    if (n2[0].b > 0) {
        collapse4(n2, 0, out);
        return out;
    }
    MeshNode4 node;
    for (int k = 0; k < 4; ++k) {
        for (int i = 0; i < 3; ++i) {
            node.lo[i][k] = k == 0 ? n2[0].lo[i] : INFINITY;
            node.hi[i][k] = k == 0 ? n2[0].hi[i] : -INFINITY;
        }
        node.a[k] = k == 0 ? n2[0].a : 0;
        node.b[k] = k == 0 ? n2[0].b : 0;
    }
    out.push_back(node);
    return out;
}

uint32_t spread10(uint32_t v) {
    v &= 1023u;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// ---------------------------------------------------------------------------
// device
// ---------------------------------------------------------------------------
struct MeshArgs {
    const MeshNode4* nodes;
    const float4* tris;   // 4 float4 per triangle
    const float* robot;   // 9 floats per triangle
    int nnodes, nrobot;
};

// The scene's BVH nodes are written only at upload, so the walk reads them
// through the constant address space: its wave-uniform node loads then become
// scalar loads also in kernels that store elsewhere (mesh_link).
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(4))) T* scene_const(const T* p) {
    return (const __attribute__((address_space(4))) T*)p;
}

// Four-wide node i, word by word (lo[3][4], hi[3][4], a[4], b[4]).
__device__ __forceinline__ MeshNode4 load_node4(const MeshNode4* nodes, int i) {
    const __attribute__((address_space(4))) uint32_t* w = scene_const((const uint32_t*)nodes) + 32 * (size_t)i;
    MeshNode4 n;
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            n.lo[k][j] = __uint_as_float(w[4 * k + j]);
            n.hi[k][j] = __uint_as_float(w[12 + 4 * k + j]);
        }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        n.a[j] = (int32_t)w[24 + j];
        n.b[j] = (int32_t)w[28 + j];
    }
    return n;
}

// Triangle record i (16 floats: v0 v1 v2, box lo, box hi, one unused),
// vector loads: as scalar loads the leaf tests ran slower in mesh_filter.
__device__ __forceinline__ void load_env_tri(const float4* __restrict__ t, int i, float* v, float* lo, float* hi) {
    const float4 a = t[4 * i], b = t[4 * i + 1], c = t[4 * i + 2], d = t[4 * i + 3];
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    v[8] = c.x;
    lo[0] = c.y; lo[1] = c.z; lo[2] = c.w;
    hi[0] = d.x; hi[1] = d.y; hi[2] = d.z;
}

// Does the robot at state q collide with the environment?  Whole wave; q is
// the same on every lane.  stack: this wave's kStackSlots LDS slots.
//
// Per 256-triangle chunk, lane l poses robot triangles [4l, 4l + 4) and takes
// their box; the wave walks the environment BVH depth-first once, visiting a
// node when any lane's box overlaps it (a ballot), so node and triangle loads
// are wave-uniform and no lane diverges in the walk.  At a leaf each lane
// whose box overlaps it tests its triangles against the leaf's: a pair is
// tested only when the boxes overlap, as the predicate requires, and a lane
// whose box overlaps a leaf overlaps every ancestor of it, so the pairs
// tested are those a per-lane walk would test.
__device__ bool mesh_collides(const MeshArgs& m, const double* q, int32_t* stack) {
    const int lane = threadIdx.x & 63;
    if (m.nnodes == 0 || m.nrobot == 0) return false;
    const Pose32 pose = pose_from_state(q);
    for (int c0 = 0; c0 < m.nrobot; c0 += kChunkTris) {
        const int first = c0 + lane * kLaneTris;
        const int cnt = max(0, min(kLaneTris, m.nrobot - first));
        // box of this lane's posed triangles; the triangles themselves are
        // posed again only for environment triangles inside this box
        float llo[3] = {INFINITY, INFINITY, INFINITY}, lhi[3] = {-INFINITY, -INFINITY, -INFINITY};
#pragma unroll 1
        for (int t = 0; t < cnt; ++t) {
            const float* src = m.robot + (size_t)(first + t) * 9;
            float tri[9];
#pragma unroll
            for (int v = 0; v < 3; ++v) xform(pose, src + 3 * v, tri + 3 * v);
            float tlo[3], thi[3];
            tri_box(tri, tlo, thi);
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                llo[i] = tlo[i] < llo[i] ? tlo[i] : llo[i];
                lhi[i] = thi[i] > lhi[i] ? thi[i] : lhi[i];
            }
        }
        const bool live = cnt > 0;
        // Current entry (ca, cb): a four-wide node (cb > 0) or a leaf's
        // triangles [ca, ca - cb); an entry is entered only when some lane's
        // box overlaps its box.  The four slot boxes of a node arrive in one
        // uniform load; the wave descends into the first slot that some lane
        // overlaps and keeps the other such slots on the stack.
        int ca = 0, cb = 1;
        int sp = 0;  // wave-uniform; at most 3 entries per four-wide level
        for (;;) {
            if (cb > 0) {
                const MeshNode4 N = load_node4(m.nodes, ca);
                uint32_t take = 0;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const float blo[3] = {N.lo[0][k], N.lo[1][k], N.lo[2][k]};
                    const float bhi[3] = {N.hi[0][k], N.hi[1][k], N.hi[2][k]};
                    take |= __ballot(live && boxes_overlap(blo, bhi, llo, lhi)) != 0 ? 1u << k : 0u;
                }
                if (take != 0) {
#pragma unroll
                    for (int k = 3; k >= 1; --k) {
                        // every slot taken but the first goes on the stack
                        if (((take >> k) & 1u) && (take & ((1u << k) - 1u))) {
                            stack[2 * sp] = N.a[k];
                            stack[2 * sp + 1] = N.b[k];
                            ++sp;
                        }
                    }
                    const int f = __builtin_ctz(take);
                    ca = f == 0 ? N.a[0] : f == 1 ? N.a[1] : f == 2 ? N.a[2] : N.a[3];
                    cb = f == 0 ? N.b[0] : f == 1 ? N.b[1] : f == 2 ? N.b[2] : N.b[3];
                    continue;
                }
            } else {
                // a lane meets a triangle's box only if it meets the leaf's
                // and every ancestor's, so these pairs are the per-lane walk's
                bool hit = false;
#pragma unroll 1
                for (int i = ca; i < ca - cb; ++i) {  // wave-uniform bounds
                    float ev[9], elo[3], ehi[3];
                    load_env_tri(m.tris, i, ev, elo, ehi);
                    if (!live || hit || !boxes_overlap(elo, ehi, llo, lhi)) continue;
#pragma unroll 1
                    for (int t = 0; t < cnt && !hit; ++t) {
                        const float* src = m.robot + (size_t)(first + t) * 9;
                        float tri[9], tlo[3], thi[3];
#pragma unroll
                        for (int v = 0; v < 3; ++v) xform(pose, src + 3 * v, tri + 3 * v);
                        tri_box(tri, tlo, thi);
                        hit = pair_collides(tri, tlo, thi, ev, elo, ehi);
                    }
                }
                if (__ballot(hit)) return true;
            }
            if (sp == 0) break;
            --sp;
            ca = __builtin_amdgcn_readfirstlane(stack[2 * sp]);
            cb = __builtin_amdgcn_readfirstlane(stack[2 * sp + 1]);
        }
    }
    return false;
}

__global__ __launch_bounds__(64) void mesh_valid(MeshArgs m, const double* __restrict__ q, int64_t n,
                                                 uint8_t* __restrict__ ok) {
    __shared__ int32_t stack[kStackSlots];
    for (int64_t i = blockIdx.x; i < n; i += gridDim.x) {
        double s[7];
#pragma unroll
        for (int j = 0; j < 7; ++j) s[j] = q[i * 7 + j];
        const bool hit = mesh_collides(m, s, stack);
        if (threadIdx.x == 0) ok[i] = hit ? 0 : 1;
    }
}

__global__ __launch_bounds__(64) void mesh_filter(MeshArgs m, const double* __restrict__ q, int64_t n,
                                                  uint8_t* __restrict__ mask, uint8_t* __restrict__ status,
                                                  int invalid_code) {
    __shared__ int32_t stack[kStackSlots];
    for (int64_t i = blockIdx.x; i < n; i += gridDim.x) {
        if (!mask[i]) continue;
        double s[7];
#pragma unroll
        for (int j = 0; j < 7; ++j) s[j] = q[i * 7 + j];
        const bool hit = mesh_collides(m, s, stack);
        if (threadIdx.x == 0 && hit) {
            mask[i] = 0;
            status[i] = (uint8_t)invalid_code;
        }
    }
}

// DMV link (src/mpt/discrete_motion_validator.hpp:71-130), one edge per wave:
// valid(to) unless the caller vouches for it, then the interior waypoints in
// the reference's ring order (table for up to 255, ring replay beyond).
// link() of n edges, P waves per edge (P = waves_per_edge; each wave is one
// workgroup, and the whole wave checks one waypoint at a time).  The check
// sequence of an edge is valid(to) at position 0 (counted even when vouched
// for, as the reference's validator counts it), then the DMV waypoints in
// table order at positions 1..mw.  Wave p takes valid(to) (p = 0) and the
// waypoints of rank p, p + P, ...; a failure lowers the edge's slot in
// `checks` (set to ~0 beforehand) to its position with atomicMin, and a wave
// stops when the slot is already below its next position.  The first
// failing position lies in one wave's share, which that wave reaches, so the
// slot ends at exactly the first failing position; no wave waits for
// another.  Edges beyond 255 waypoints replay the 256-slot ring on wave 0
// alone.  mesh_link_finish turns the slots into verdicts and check counts.
__global__ __launch_bounds__(64) void mesh_link(SpaceArgs sp, MeshArgs m, double inv,
                                                const uint8_t* __restrict__ table, const double* __restrict__ from,
                                                const double* __restrict__ to, int64_t n, int known_valid,
                                                const uint8_t* __restrict__ mask, int waves_per_edge,
                                                uint32_t* __restrict__ slot) {
    __shared__ int32_t stack[kStackSlots];
    __shared__ uint64_t rmin[256], rmax[256];
    __shared__ uint64_t emitted[64];
    __shared__ int cnt_s;
    // edge endpoints live in LDS (read back per waypoint) to keep them out of
    // the registers the collision walk needs
    __shared__ double ab_s[14];
    volatile double* ab = ab_s;
    const int lane = threadIdx.x;
    const int64_t P = waves_per_edge;
    auto current = [&](int64_t e) {  // the edge's slot, wave-uniform
        return (uint32_t)__builtin_amdgcn_readfirstlane(
            (int)__hip_atomic_load(&slot[e], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    };
    for (int64_t item = blockIdx.x; item < n * P; item += gridDim.x) {
        const int64_t e = item / P;
        const int64_t p = item % P;
        if (mask && !mask[e]) continue;
        if (lane < 14) ab[lane] = lane < 7 ? from[e * 7 + lane] : to[e * 7 + lane - 7];
        __syncthreads();
        double a[7], b[7], q[7];
        auto load_ab = [&]() {
#pragma unroll
            for (int j = 0; j < 7; ++j) {
                a[j] = ab[j];
                b[j] = ab[7 + j];
            }
        };
        load_ab();
        if (p == 0 && !known_valid && mesh_collides(m, b, stack)) {
            if (lane == 0) atomicMin(&slot[e], 0u);
            __syncthreads();
            continue;
        }
        load_ab();
        double x = __builtin_ceil(space_dist(sp, a, b) * inv);
        int64_t steps = 0;
        if (x >= 2.0) steps = x < 0x1.0p62 ? (int64_t)x : ((int64_t)1 << 62);
        const int64_t mw = steps > 0 ? steps - 1 : 0;
        if (mw > kMaxDmvWaypointsMesh) {  // not checked; the finish kernel reports it
            __syncthreads();
            continue;
        }
        const double delta = steps > 0 ? 1.0 / (double)steps : 0.0;
        if (mw <= kTableMaxMesh) {
            const uint8_t* ord = table + (mw * (mw - 1)) / 2;
            for (int64_t o = p; o < mw; o += P) {
                const uint32_t pos = (uint32_t)(o + 1);
                if (current(e) < pos) break;  // an earlier position already failed
                load_ab();
                space_interp(sp, a, b, (double)ord[o] * delta, q);
                if (mesh_collides(m, q, stack)) {
                    if (lane == 0) atomicMin(&slot[e], pos);
                    break;
                }
            }
        } else if (p == 0) {
            // ring replay: lane 0 emits up to 64 indices at a time
            bool failed = false;
            int64_t done = 0;
            uint64_t qs = 0, qe = 1, seq_cur = 0, seq_end = 0;
            bool seq = false;
            if (lane == 0) {
                rmin[0] = 1;
                rmax[0] = (uint64_t)mw;
            }
            while (!failed) {
                if (lane == 0) {
                    int c = 0;
                    while (c < 64 && (seq || qs != qe)) {
                        if (seq) {
                            emitted[c++] = seq_cur;
                            if (seq_cur == seq_end) seq = false;
                            else ++seq_cur;
                            continue;
                        }
                        const uint64_t mn = rmin[qs % 256], mx = rmax[qs % 256];
                        ++qs;
                        if (mn == mx) {
                            emitted[c++] = mn;
                        } else if (qe + 2 < qs + 256) {
                            const uint64_t mid = (mn + mx) / 2;
                            emitted[c++] = mid;
                            if (mn < mid) {
                                rmin[qe % 256] = mn;
                                rmax[qe % 256] = mid - 1;
                                ++qe;
                            }
                            if (mid < mx) {
                                rmin[qe % 256] = mid + 1;
                                rmax[qe % 256] = mx;
                                ++qe;
                            }
                        } else {
                            seq = true;
                            seq_cur = mn;
                            seq_end = mx;
                        }
                    }
                    cnt_s = c;
                }
                __syncthreads();
                const int c = cnt_s;
                if (c == 0) break;
                for (int o = 0; o < c; ++o) {
                    load_ab();
                    space_interp(sp, a, b, (double)emitted[o] * delta, q);
                    ++done;
                    if (mesh_collides(m, q, stack)) {
                        failed = true;
                        break;
                    }
                }
                __syncthreads();
            }
            if (failed && lane == 0) atomicMin(&slot[e], (uint32_t)done);
        }
        __syncthreads();
    }
}

// Verdicts and check counts of the edges from their slots.
__global__ __launch_bounds__(256) void mesh_link_finish(SpaceArgs sp, double inv, const double* __restrict__ from,
                                                        const double* __restrict__ to, int64_t n,
                                                        const uint8_t* __restrict__ mask,
                                                        uint8_t* __restrict__ verdict, uint32_t* __restrict__ checks) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n || (mask && !mask[e])) return;
    const uint32_t s = checks[e];
    if (s != 0xffffffffu) {
        verdict[e] = 0;
        checks[e] = s + 1u;
        return;
    }
    double a[7], b[7];
#pragma unroll
    for (int j = 0; j < 7; ++j) {
        a[j] = from[e * 7 + j];
        b[j] = to[e * 7 + j];
    }
    const double x = __builtin_ceil(space_dist(sp, a, b) * inv);
    int64_t steps = 0;
    if (x >= 2.0) steps = x < 0x1.0p62 ? (int64_t)x : ((int64_t)1 << 62);
    const int64_t mw = steps > 0 ? steps - 1 : 0;
    if (mw > kMaxDmvWaypointsMesh) {
        verdict[e] = 0;
        checks[e] = MPTG_CHECKS_OVER_LIMIT;
    } else {
        verdict[e] = 1;
        checks[e] = (uint32_t)(1 + mw);
    }
}

}  // namespace

MeshArgs mesh_args(const mptg_ctx* ctx) {
    const SceneDev& s = ctx->scene;
    return MeshArgs{(const MeshNode4*)s.mesh_nodes.p, (const float4*)s.mesh_tris.p, (const float*)s.mesh_robot.p,
                    s.mesh_nnodes, s.mesh_nrobot};
}

int launch_mesh_valid(mptg_ctx* ctx, const double* d_q, int64_t n, uint8_t* d_ok) {
    if (n <= 0) return MPTG_OK;
    const unsigned g = (unsigned)std::min<int64_t>(n, (int64_t)ctx->cus * ctx->tune.mesh_waves_per_cu);
    hipLaunchKernelGGL(mesh_valid, dim3(g), dim3(64), 0, ctx->stream, mesh_args(ctx), d_q, n, d_ok);
    MPTG_HIP(ctx, hipGetLastError());
    return MPTG_OK;
}

int launch_mesh_filter(mptg_ctx* ctx, const double* d_states, int64_t n, uint8_t* d_mask, uint8_t* d_status,
                       int invalid_code) {
    if (n <= 0) return MPTG_OK;
    const unsigned g = (unsigned)std::min<int64_t>(n, (int64_t)ctx->cus * ctx->tune.mesh_waves_per_cu);
    hipLaunchKernelGGL(mesh_filter, dim3(g), dim3(64), 0, ctx->stream, mesh_args(ctx), d_states, n, d_mask, d_status,
                       invalid_code);
    MPTG_HIP(ctx, hipGetLastError());
    return MPTG_OK;
}

int launch_mesh_link(mptg_ctx* ctx, const uint8_t* d_table, const double* d_from, const double* d_to, int64_t n,
                     int known_valid, const uint8_t* d_mask, uint8_t* d_verdict, uint32_t* d_checks) {
    if (n <= 0) return MPTG_OK;
    int rc;
    if ((rc = launch_fill_u32(ctx, d_checks, n, 0xffffffffu))) return rc;  // first-failure slots
    // waves per edge: the tuned count, or more for a batch too small to fill
    // the device (the rewire pass's late batches hold a few edges), up to 256
    // (the verdicts and counts do not depend on it)
    const int64_t fill = (int64_t)ctx->cus * ctx->tune.mesh_waves_per_cu;
    const int P = (int)std::max<int64_t>(ctx->tune.mesh_edge_waves, std::min<int64_t>(256, fill / n));
    const unsigned gp = (unsigned)std::min<int64_t>(n * P, fill);
    hipLaunchKernelGGL(mesh_link, dim3(gp), dim3(64), 0, ctx->stream, space_args(ctx), mesh_args(ctx),
                       1.0 / ctx->link_param, d_table, d_from, d_to, n, known_valid, d_mask, P, d_checks);
    hipLaunchKernelGGL(mesh_link_finish, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, space_args(ctx),
                       1.0 / ctx->link_param, d_from, d_to, n, d_mask, d_verdict, d_checks);
    MPTG_HIP(ctx, hipGetLastError());
    return MPTG_OK;
}

}  // namespace mptg

using namespace mptg;

extern "C" int mptg_scene_mesh(mptg_ctx* ctx, const float* h_robot, int nrobot, const float* h_env, int nenv) {
    int rc = check_device_ready(ctx);
    if (rc) return rc;
    if (ctx->metric != MPTG_METRIC_SE3) return fail(ctx, MPTG_E_STATE, "mesh scenes need the SE3 space");
    if (nrobot < 0 || nenv < 0 || (nrobot > 0 && !h_robot) || (nenv > 0 && !h_env))
        return fail(ctx, MPTG_E_ARG, "bad mesh arrays");
    if (nenv > (1 << 20)) return fail(ctx, MPTG_E_RANGE, "at most 2^20 environment triangles");
    if (nrobot > (1 << 16)) return fail(ctx, MPTG_E_RANGE, "at most 2^16 robot triangles");
    for (int64_t i = 0; i < (int64_t)nrobot * 9; ++i)
        if (!std::isfinite(h_robot[i])) return fail(ctx, MPTG_E_ARG, "non-finite robot vertex");
    for (int64_t i = 0; i < (int64_t)nenv * 9; ++i)
        if (!std::isfinite(h_env[i])) return fail(ctx, MPTG_E_ARG, "non-finite environment vertex");
    // environment BVH
    BvhBuilder bb;
    bb.refs.resize((size_t)nenv);
    for (int t = 0; t < nenv; ++t) {
        BuildRef& r = bb.refs[(size_t)t];
        tri_box(h_env + (size_t)t * 9, r.lo, r.hi);
        for (int i = 0; i < 3; ++i) r.c[i] = 0.5f * (r.lo[i] + r.hi[i]);
        r.id = t;
    }
    if (nenv > 0) bb.build(0, nenv, 0);
    if (nenv > 0 && bvh_depth(bb.nodes, 0, 0) > kBuildDepth)
        return fail(ctx, MPTG_E_STATE, "environment BVH deeper than the traversal stack");
    std::vector<float> tris((size_t)nenv * kTriFloats, 0.0f);
    for (int i = 0; i < nenv; ++i) {
        const BuildRef& r = bb.refs[(size_t)i];
        float* d = &tris[(size_t)i * kTriFloats];
        std::memcpy(d, h_env + (size_t)r.id * 9, 9 * sizeof(float));
        for (int k = 0; k < 3; ++k) {
            d[9 + k] = r.lo[k];
            d[12 + k] = r.hi[k];
        }
    }
    // robot: Morton order of triangle centroids
    std::vector<float> robot((size_t)nrobot * 9);
    {
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        std::vector<float> cen((size_t)nrobot * 3);
        for (int t = 0; t < nrobot; ++t)
            for (int i = 0; i < 3; ++i) {
                const float* v = h_robot + (size_t)t * 9;
                float c = (v[i] + v[3 + i] + v[6 + i]) / 3.0f;
                cen[(size_t)t * 3 + i] = c;
                lo[i] = std::min(lo[i], c);
                hi[i] = std::max(hi[i], c);
            }
        std::vector<std::pair<uint32_t, int>> key((size_t)nrobot);
        for (int t = 0; t < nrobot; ++t) {
            uint32_t code = 0;
            for (int i = 0; i < 3; ++i) {
                float ext = hi[i] - lo[i];
                uint32_t qv = ext > 0 ? (uint32_t)std::min(1023.0f, (cen[(size_t)t * 3 + i] - lo[i]) / ext * 1023.0f) : 0;
                code |= spread10(qv) << i;
            }
            key[(size_t)t] = {code, t};
        }
        std::sort(key.begin(), key.end());
        for (int t = 0; t < nrobot; ++t)
            std::memcpy(&robot[(size_t)t * 9], h_robot + (size_t)key[(size_t)t].second * 9, 9 * sizeof(float));
    }
    SceneDev& s = ctx->scene;
    MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const std::vector<MeshNode4> nodes4 = collapse_bvh(bb.nodes);
    if ((rc = ensure(ctx, s.mesh_nodes, sizeof(MeshNode4) * std::max<size_t>(1, nodes4.size())))) return rc;
    if ((rc = ensure(ctx, s.mesh_tris, sizeof(float) * std::max<size_t>(16, tris.size())))) return rc;
    if ((rc = ensure(ctx, s.mesh_robot, sizeof(float) * std::max<size_t>(9, robot.size())))) return rc;
    if (!nodes4.empty())
        MPTG_HIP(ctx, hipMemcpy(s.mesh_nodes.p, nodes4.data(), sizeof(MeshNode4) * nodes4.size(),
                                hipMemcpyHostToDevice));
    if (!tris.empty())
        MPTG_HIP(ctx, hipMemcpy(s.mesh_tris.p, tris.data(), sizeof(float) * tris.size(), hipMemcpyHostToDevice));
    if (!robot.empty())
        MPTG_HIP(ctx, hipMemcpy(s.mesh_robot.p, robot.data(), sizeof(float) * robot.size(), hipMemcpyHostToDevice));
    s.mesh_nnodes = (int)nodes4.size();
    s.mesh_nrobot = nrobot;
    s.kind = MPTG_SCENE_MESH;
    ctx->valid_upto = 0;  // valid() changed: every node is checked again
    ctx->some_invalid = false;
    return MPTG_OK;
}
