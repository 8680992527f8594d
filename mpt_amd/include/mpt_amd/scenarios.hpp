// Scenario classes for the GPU planner: the reference demos' scenario models
// with their valid()/link() described as data for the device.
//
//   LinkManipulator<N>  demo/link_manipulator_scenario.hpp:54-155
//                       (L1 space, bounds [-pi,pi]^N, goal radius 1e-6,
//                        link = recursive midpoints to L-inf < 0.02)
//   PointAmongBoxes     demo/holonomic_2d_point_scenario.hpp:52-130 with
//                       Rect obstacles (demo/shape_hierarchy.hpp:168-211) and
//                       a DiscreteMotionValidator link (step given)
//   OccupancyGrid2D     demo/png_2d_scenario.hpp:71-166 with a
//                       DiscreteMotionValidator link (step given)
//   RigidBodyMeshes     demo/se3_rigid_body_scenario.hpp:250-340: SE(3) with
//                       the rotation weighted 50, robot and environment
//                       triangle meshes, DMV step (|max - min| + 50 pi/2) * res
//
// Each provides State, kDimensions, gpu_scene(mptg_ctx*), bounds_to(lo, hi),
// goal_state(), goal_radius() and distance(a, b) (used for path costs).
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "mptg.h"

namespace mpt_amd {

namespace detail {
inline void scene_check(mptg_ctx* ctx, int rc) {
    if (rc != MPTG_OK) throw std::runtime_error(std::string("mptg: ") + mptg_last_error(ctx));
}
}  // namespace detail

struct Disk {
    double cx, cy, r;
};

template <int N>
class LinkManipulator {
   public:
    static constexpr int kDimensions = N;
    using State = std::array<double, N>;

    LinkManipulator(const State& goal, std::vector<Disk> disks, std::vector<double> lengths, double robotRadius)
        : goal_(goal), disks_(std::move(disks)), lengths_(std::move(lengths)), radius_(robotRadius) {
        if ((int)lengths_.size() != N) throw std::invalid_argument("one length per link");
    }

    void gpu_scene(mptg_ctx* ctx) const {
        detail::scene_check(ctx, mptg_space(ctx, MPTG_METRIC_L1, N, 1.0, 1.0));
        std::vector<double> dk;
        dk.reserve(3 * disks_.size());
        for (const Disk& d : disks_) {
            dk.push_back(d.cx);
            dk.push_back(d.cy);
            dk.push_back(d.r);
        }
        detail::scene_check(ctx, mptg_scene_arm(ctx, lengths_.data(), N, radius_, dk.data(), (int)disks_.size()));
        detail::scene_check(ctx, mptg_set_link(ctx, MPTG_LINK_BISECT, 0.02));
    }
    void bounds_to(double* lo, double* hi) const {
        for (int i = 0; i < N; ++i) {
            lo[i] = -M_PI;
            hi[i] = M_PI;
        }
    }
    const State& goal_state() const { return goal_; }
    double goal_radius() const { return 1e-6; }
    double distance(const double* a, const double* b) const {
        double s = 0.0;
        for (int i = 0; i < N; ++i) s = s + std::fabs(a[i] - b[i]);
        return s;
    }

   private:
    State goal_;
    std::vector<Disk> disks_;
    std::vector<double> lengths_;
    double radius_;
};

class PointAmongBoxes {
   public:
    static constexpr int kDimensions = 2;
    using State = std::array<double, 2>;
    struct Box {
        double x0, y0, x1, y1;
    };

    PointAmongBoxes(double width, double height, std::vector<Box> boxes, const State& goal, double step)
        : w_(width), h_(height), boxes_(std::move(boxes)), goal_(goal), step_(step) {}

    void gpu_scene(mptg_ctx* ctx) const {
        detail::scene_check(ctx, mptg_space(ctx, MPTG_METRIC_L2, 2, 1.0, 1.0));
        detail::scene_check(ctx, mptg_scene_boxes(ctx, reinterpret_cast<const double*>(boxes_.data()),
                                                  (int)boxes_.size()));
        detail::scene_check(ctx, mptg_set_link(ctx, MPTG_LINK_DMV, step_));
    }
    void bounds_to(double* lo, double* hi) const {
        lo[0] = 0;
        lo[1] = 0;
        hi[0] = w_;
        hi[1] = h_;
    }
    const State& goal_state() const { return goal_; }
    double goal_radius() const { return 1e-6; }
    double distance(const double* a, const double* b) const {
        double dx = a[0] - b[0], dy = a[1] - b[1];
        return std::sqrt(dx * dx + dy * dy);
    }

   private:
    double w_, h_;
    std::vector<Box> boxes_;
    State goal_;
    double step_;
};

class OccupancyGrid2D {
   public:
    static constexpr int kDimensions = 2;
    using State = std::array<double, 2>;

    // cells: height x width, nonzero = obstacle
    OccupancyGrid2D(int width, int height, std::vector<uint8_t> cells, const State& goal, double step)
        : w_(width), h_(height), cells_(std::move(cells)), goal_(goal), step_(step) {
        if ((int64_t)cells_.size() != (int64_t)w_ * h_) throw std::invalid_argument("grid size");
    }

    void gpu_scene(mptg_ctx* ctx) const {
        detail::scene_check(ctx, mptg_space(ctx, MPTG_METRIC_L2, 2, 1.0, 1.0));
        detail::scene_check(ctx, mptg_scene_grid(ctx, cells_.data(), w_, h_));
        detail::scene_check(ctx, mptg_set_link(ctx, MPTG_LINK_DMV, step_));
    }
    void bounds_to(double* lo, double* hi) const {
        lo[0] = 0;
        lo[1] = 0;
        hi[0] = w_;
        hi[1] = h_;
    }
    const State& goal_state() const { return goal_; }
    double goal_radius() const { return 1e-6; }
    double distance(const double* a, const double* b) const {
        double dx = a[0] - b[0], dy = a[1] - b[1];
        return std::sqrt(dx * dx + dy * dy);
    }

   private:
    int w_, h_;
    std::vector<uint8_t> cells_;
    State goal_;
    double step_;
};

class RigidBodyMeshes {
   public:
    static constexpr int kDimensions = 7;
    using State = std::array<double, 7>;  // qx qy qz qw tx ty tz
    static constexpr double kSO3Weight = 50.0;

    // triangles: 9 floats each; volume [min, max]^3 bounds the translation
    RigidBodyMeshes(std::vector<float> robot, std::vector<float> env, const State& goal, const std::array<double, 3>& min,
                    const std::array<double, 3>& max, double checkResolution = 0.01)
        : robot_(std::move(robot)), env_(std::move(env)), goal_(goal), min_(min), max_(max), res_(checkResolution) {
        if (robot_.size() % 9 || env_.size() % 9) throw std::invalid_argument("triangles are 9 floats");
    }

    void gpu_scene(mptg_ctx* ctx) const {
        detail::scene_check(ctx, mptg_space(ctx, MPTG_METRIC_SE3, 7, kSO3Weight, 1.0));
        detail::scene_check(ctx, mptg_scene_mesh(ctx, robot_.data(), (int)(robot_.size() / 9), env_.data(),
                                                 (int)(env_.size() / 9)));
        double ext = 0;
        for (int i = 0; i < 3; ++i) ext += (max_[i] - min_[i]) * (max_[i] - min_[i]);
        detail::scene_check(ctx, mptg_set_link(ctx, MPTG_LINK_DMV, (std::sqrt(ext) + kSO3Weight * M_PI / 2) * res_));
    }
    void bounds_to(double* lo, double* hi) const {
        for (int i = 0; i < 4; ++i) {
            lo[i] = -1;
            hi[i] = 1;
        }
        for (int i = 0; i < 3; ++i) {
            lo[4 + i] = min_[i];
            hi[4 + i] = max_[i];
        }
    }
    const State& goal_state() const { return goal_; }
    double goal_radius() const { return 1e-6; }
    double distance(const double* a, const double* b) const {
        double d = std::fabs(a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3]);
        double r = d >= 1.0 ? 0.0 : std::acos(d);
        double t = 0;
        for (int i = 4; i < 7; ++i) t += (a[i] - b[i]) * (a[i] - b[i]);
        return kSO3Weight * r + std::sqrt(t);
    }

   private:
    std::vector<float> robot_, env_;
    State goal_;
    std::array<double, 3> min_, max_;
    double res_;
};

}  // namespace mpt_amd
