// pf2D.hpp -- mirror of the reference plain ParticleFilter and its my_gmm
// (src/pf2D.h:12-51) on the gfx950 C ABI of pf2d.h.
#pragma once

#include <cstdint>
#include <vector>

#include "../pf2d.h"
#include "Mat.hpp"

namespace mkfbt {
namespace pf2d {

class ParticleFilter;

class my_gmm {
public:
    // loadGaussian(mean 1 x d, sigma d x d, weight) (pf2D.cpp:28-37)
    void loadGaussian(Mat mean, Mat sigma, double weight);
    std::vector<Mat> mean;
    std::vector<Mat> sigma;
    std::vector<double> weight;
    int N = 0;

private:
    friend class ParticleFilter;
    ParticleFilter *pf_ = nullptr;
};

class ParticleFilter {
public:
    // ParticleFilter(numParticles, numDims, side1) (pf2D.cpp:44-71).
    ParticleFilter(int numParticles, int numDims, bool side1, uint64_t seed = 0, int device = 0);
    ~ParticleFilter();
    ParticleFilter(const ParticleFilter &) = delete;
    ParticleFilter &operator=(const ParticleFilter &) = delete;
    // update(measurement 2 x 2) (pf2D.cpp:148-210)
    void update(Mat measurement);
    // getEstimator() -> 1 x d (pf2D.cpp:79-87)
    Mat getEstimator();
    my_gmm gmm;

    pf2d_handle *handle() { return h_; }

private:
    friend class my_gmm;
    std::vector<double> uniforms(size_t n);
    std::vector<double> normals(size_t n);
    pf2d_handle *h_ = nullptr;
    int N_ = 0, d_ = 0;
    uint64_t seed_ = 0, draws_ = 0;
};

}  // namespace pf2d
}  // namespace mkfbt
