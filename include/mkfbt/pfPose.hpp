// pfPose.hpp -- the non-ROS core of the reference's PFTracker (src/pfPose.h:30-78)
// on the gfx950 C ABI of pf_tracker.h: two arm filters loaded from the GMM yml
// files as the constructor does (pfPose.cpp:7-74), and per frame the work of
// callback (pfPose.cpp:332-349) -- likelihood blur, proposals, association
// weights, proposal resampling, measurements, both updates -- with the pose
// output of e1/e2 (pfPose.cpp:347-348), get3Dpose (pfPose.cpp:93-127) and the
// joint positions publishTFtree and publish2Dpos send (pfPose.cpp:129-208).
// ROS message handling, drawing and publishing stay with the caller.
#pragma once

#include <cstdint>
#include <memory>
#include <string>

#include "../pf_tracker.h"
#include "Mat.hpp"
#include "pf2DRao.hpp"

namespace mkfbt {

struct PoseFrame {
    double lml[2] = {0.0, 0.0};   // log mean weight of each arm's update
    Mat e1, e2;                   // 1 x npose: pca_proj^T getEstimator() + pca_mean
    Mat p3D1, p3D2;               // 3 x 5: get3Dpose (rows x, y, z; hand, elbow, shoulder, head, neck)
    Mat tf;                       // 8 x 3: Head, Left_Hand, Right_Hand, Left_Elbow, Right_Elbow,
                                  // Right_Shoulder, Left_Shoulder, Neck relative to Head (TF axes)
    Mat pose2d;                   // 8 x 2: hand 1, hand 2, head, neck, elbow 1, elbow 2, shoulder 1, shoulder 2
    int status = 0;               // 1: a resampling fell back to uniform weights
};

class PFTracker {
public:
    // PFTracker::PFTracker (pfPose.cpp:7-74): left_arm_training -> arm 1,
    // right_arm_training -> arm 2, each with its own gamma; numParticles
    // (pfPose.cpp:57); tracks from resample(gmm.weight) + resetTracker
    // (pfPose.cpp:68-71).  cfg (optional) overrides the tracker defaults.
    PFTracker(const std::string &left_yml, const std::string &right_yml, int numParticles = 500,
              uint64_t seed = 1, int device = 0, const pf_tracker_config *cfg = nullptr);
    ~PFTracker();
    PFTracker(const PFTracker &) = delete;
    PFTracker &operator=(const PFTracker &) = delete;

    // callback (pfPose.cpp:332-349) for a MONO8 likelihood image of the
    // configured size and the face ROI {x_offset, y_offset, width, height};
    // roi == nullptr is a frame without a face (pfPose.cpp:370-377).
    PoseFrame callback(const uint8_t *likelihood, const int32_t *roi);
    // The pose output of the current tracks, without an update.
    PoseFrame pose();

    pfrao::ParticleFilter &pf1() { return *pf1_; }
    pfrao::ParticleFilter &pf2() { return *pf2_; }
    pf_tracker *handle() { return tr_; }

private:
    PoseFrame run(const uint8_t *likelihood, const int32_t *roi, bool update);
    std::unique_ptr<pfrao::ParticleFilter> pf1_, pf2_;
    Mat h1_pca, m1_pca, h2_pca, m2_pca;
    pf_tracker *tr_ = nullptr;
    int npose_ = 0;
};

}  // namespace mkfbt
