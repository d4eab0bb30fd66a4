// pf2DRao.hpp -- mirror of the reference Rao-Blackwellised ParticleFilter
// (src/pf2DRao.h:13-31) on the gfx950 C ABI of pf_rbpf.h.
#pragma once

#include <cstdint>
#include <vector>

#include "../pf_rbpf.h"
#include "Mat.hpp"
#include "my_gmm.hpp"

namespace mkfbt {
namespace pfrao {

class ParticleFilter {
public:
    // ParticleFilter(int nParticles) (pf2DRao.cpp:13-16).  seed selects the
    // Philox4x32-10 stream that replaces the reference's cv::RNG(getTickCount()).
    explicit ParticleFilter(int nParticles, uint64_t seed = 0, int device = 0);
    ~ParticleFilter();
    ParticleFilter(const ParticleFilter &) = delete;
    ParticleFilter &operator=(const ParticleFilter &) = delete;

    // update (pf2DRao.cpp:125-158): measurement is m x N, column j for particle j.
    void update(Mat measurement);
    // Candidate (PDAF) form of update: detections shared by all particles.
    void update_candidates(const std::vector<int> &count, const std::vector<double> &cand);
    // getEstimator (pf2DRao.cpp:23-31): d x 1.
    Mat getEstimator();
    // getSamples (pf2DRao.cpp:85-103): 2 x N proposals around H xbar + M.
    Mat getSamples(Mat H, Mat M, int N, double scale);
    // getSampleProb (pf2DRao.cpp:105-122).
    void getSampleProb(Mat H, Mat M, Mat input1, Mat input2, std::vector<double> &weight1,
                       std::vector<double> &weight2, double scale);
    // resample (pf2DRao.cpp:175-210): N systematic indices from the weights.
    std::vector<int> resample(std::vector<double> weights, int N);
    // PDAF joint kinds for update_candidates (1 = cluttered joint).
    void setJointClutter(const std::vector<int> &clutter);

    my_gmm gmm;

    pf_handle *handle() { return h_; }
    uint32_t draw_u32();
    // Position of the host random stream: the next draw_u32() returns
    // Philox4x32-10 word draws() of (seed, stream 0).
    uint64_t draws() const { return draws_; }
    int lastStatus() const { return last_status_; }

private:
    friend class my_gmm;
    pf_handle *h_ = nullptr;
    uint64_t seed_ = 0, draws_ = 0;
    int last_status_ = 0;
};

}  // namespace pfrao
}  // namespace mkfbt
