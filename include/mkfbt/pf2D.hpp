// pf2D.hpp -- mirror of the reference plain ParticleFilter and its my_gmm
// (src/pf2D.h:12-51) on the gfx950 C ABI of pf2d.h.
#pragma once

#include <cstdint>
#include <vector>

#include "../pf2d.h"
#include "Mat.hpp"

namespace mkfbt {
namespace pf2d {

// my_gmm (pf2D.h:12-23): the GMM prior in the members the update reads.
class my_gmm {
public:
    // loadGaussian(mean 1 x d, sigma d x d, weight) (pf2D.cpp:28-37): appends
    // mean, sigma_i = sigma^-1 (from its Cholesky factor; PfException
    // PF_ERR_NOT_SPD if sigma is not positive definite), det_s =
    // 1 / ((2 pi)^(d/2) sqrt(det sigma)) and weight, and counts N.
    void loadGaussian(Mat mean, Mat sigma, double weight);
    std::vector<Mat> mean;
    std::vector<Mat> sigma_i;
    std::vector<double> det_s;
    std::vector<double> weight;
    int N = 0;
};

class ParticleFilter {
public:
    // ParticleFilter() (pf2D.h:29): no particles; update, predict and
    // getEstimator throw PF_ERR_STATE until a sized filter is move-assigned.
    ParticleFilter();
    // ParticleFilter(numParticles, numDims, side1) (pf2D.cpp:44-71).
    ParticleFilter(int numParticles, int numDims, bool side1, uint64_t seed = 0, int device = 0);
    ~ParticleFilter();
    ParticleFilter(const ParticleFilter &) = delete;
    ParticleFilter &operator=(const ParticleFilter &) = delete;
    ParticleFilter(ParticleFilter &&o) noexcept;
    ParticleFilter &operator=(ParticleFilter &&o) noexcept;
    // predict() (pf2D.cpp:90-102): the random walk x += 5 n, normals drawn
    // per particle for columns 2..5, 6, 7, 0, 1.
    void predict();
    // update(measurement 2 x 2) (pf2D.cpp:148-210): the gmm members as they
    // are at the call (mean, sigma_i, det_s, weight of the first N
    // components), likelihood, resample, predict.
    void update(Mat measurement);
    // getEstimator() -> 1 x d (pf2D.cpp:79-87)
    Mat getEstimator();
    my_gmm gmm;

    pf2d_handle *handle() { return h_; }

private:
    void need_handle(const char *what) const;
    void sync_gmm();
    std::vector<double> uniforms(size_t n);
    std::vector<double> normals(size_t n);
    pf2d_handle *h_ = nullptr;
    int N_ = 0, d_ = 0;
    uint64_t seed_ = 0, draws_ = 0;
    std::vector<double> gmm_sent_;   // the gmm members of the last upload
};

}  // namespace pf2d
}  // namespace mkfbt
