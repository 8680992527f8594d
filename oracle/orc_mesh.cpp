// CPU oracle: rigid-body mesh validity for the SE(3) scene — TEST
// INFRASTRUCTURE ONLY.
//
// Restates the reference's valid(q) = !collide(robot at pose(q), env)
// (demo/se3_rigid_body_scenario.hpp:288-312, pose = Translation * Quaternion
// :258-261) with the project's FP32 predicate from mesh_hd.hpp (the same
// code the device runs): pose, per-pair box overlap, separating-axis
// triangle test.  FCL's own verdicts are not available here (parity
// unpinned against FCL; the triangle test itself is pinned against exact
// rational arithmetic in tests/test_oracle_mesh.py).
//
// Culling is this oracle's own uniform grid over the environment triangle
// boxes, a different structure from the device BVH; both only skip pairs
// whose boxes are apart, which the predicate rejects anyway.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../mpt_amd/csrc/mesh_hd.hpp"
#include "oracle.h"

namespace {

struct MeshIndex {
    float lo[3], hi[3], cell;
    int n[3];
    std::vector<int32_t> start;  // cell -> range in tris
    std::vector<int32_t> tris;
    std::vector<float> boxes;    // nenv x 6
};

int cell_of(const MeshIndex& m, int axis, float v) {
    int c = (int)std::floor((v - m.lo[axis]) / m.cell);
    return std::min(std::max(c, 0), m.n[axis] - 1);
}

}  // namespace

extern "C" void* orc_mesh_index_build(const orc_scene* sc) {
    auto* m = new MeshIndex();
    const int ne = sc->nenv;
    m->boxes.resize((size_t)ne * 6);
    for (int i = 0; i < 3; ++i) {
        m->lo[i] = INFINITY;
        m->hi[i] = -INFINITY;
    }
    for (int t = 0; t < ne; ++t) {
        float* b = &m->boxes[(size_t)t * 6];
        mptg::tri_box(sc->env_tris + (size_t)t * 9, b, b + 3);
        for (int i = 0; i < 3; ++i) {
            m->lo[i] = std::min(m->lo[i], b[i]);
            m->hi[i] = std::max(m->hi[i], b[3 + i]);
        }
    }
    float ext = 0;
    for (int i = 0; i < 3; ++i) ext = std::max(ext, m->hi[i] - m->lo[i]);
    m->cell = std::max(ext / 64.0f, 1e-6f);
    for (int i = 0; i < 3; ++i) m->n[i] = std::max(1, (int)std::ceil((m->hi[i] - m->lo[i]) / m->cell));
    const size_t ncell = (size_t)m->n[0] * m->n[1] * m->n[2];
    std::vector<int32_t> count(ncell + 1, 0);
    auto each_cell = [&](int t, auto&& f) {
        const float* b = &m->boxes[(size_t)t * 6];
        int c0[3], c1[3];
        for (int i = 0; i < 3; ++i) {
            c0[i] = cell_of(*m, i, b[i]);
            c1[i] = cell_of(*m, i, b[3 + i]);
        }
        for (int z = c0[2]; z <= c1[2]; ++z)
            for (int y = c0[1]; y <= c1[1]; ++y)
                for (int x = c0[0]; x <= c1[0]; ++x) f(((size_t)z * m->n[1] + y) * m->n[0] + x);
    };
    for (int t = 0; t < ne; ++t) each_cell(t, [&](size_t c) { count[c + 1]++; });
    for (size_t c = 0; c < ncell; ++c) count[c + 1] += count[c];
    m->start = count;
    m->tris.resize((size_t)count[ncell]);
    std::vector<int32_t> fill(count.begin(), count.end() - 1);
    for (int t = 0; t < ne; ++t) each_cell(t, [&](size_t c) { m->tris[(size_t)fill[c]++] = t; });
    return m;
}

extern "C" void orc_mesh_index_free(void* index) { delete static_cast<MeshIndex*>(index); }

extern "C" int32_t orc_tri_tri(const float* p, const float* q) { return mptg::tri_tri_overlap(p, q) ? 1 : 0; }

extern "C" void orc_pose_robot(const orc_scene* sc, const double* q, float* out) {
    const mptg::Pose32 pose = mptg::pose_from_state(q);
    for (int t = 0; t < sc->nrobot; ++t)
        for (int v = 0; v < 3; ++v) mptg::xform(pose, sc->robot_tris + (size_t)t * 9 + v * 3, out + (size_t)t * 9 + v * 3);
}

namespace orc_internal {

bool valid_mesh(const orc_scene* sc, const double* q) {
    const auto* m = static_cast<const MeshIndex*>(sc->mesh_index);
    if (!m) return false;
    const mptg::Pose32 pose = mptg::pose_from_state(q);
    for (int t = 0; t < sc->nrobot; ++t) {
        float a[9], alo[3], ahi[3];
        for (int v = 0; v < 3; ++v) mptg::xform(pose, sc->robot_tris + (size_t)t * 9 + v * 3, a + v * 3);
        mptg::tri_box(a, alo, ahi);
        if (!mptg::boxes_overlap(alo, ahi, m->lo, m->hi)) continue;
        int c0[3], c1[3];
        for (int i = 0; i < 3; ++i) {
            c0[i] = cell_of(*m, i, alo[i]);
            c1[i] = cell_of(*m, i, ahi[i]);
        }
        for (int z = c0[2]; z <= c1[2]; ++z)
            for (int y = c0[1]; y <= c1[1]; ++y)
                for (int x = c0[0]; x <= c1[0]; ++x) {
                    const size_t c = ((size_t)z * m->n[1] + y) * m->n[0] + x;
                    for (int32_t k = m->start[c]; k < m->start[c + 1]; ++k) {
                        const int32_t e = m->tris[(size_t)k];
                        const float* b = &m->boxes[(size_t)e * 6];
                        if (mptg::pair_collides(a, alo, ahi, sc->env_tris + (size_t)e * 9, b, b + 3)) return false;
                    }
                }
    }
    return true;
}

}  // namespace orc_internal
