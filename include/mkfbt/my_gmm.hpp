// my_gmm.hpp -- mirror of the reference state_params / my_gmm (src/my_gmm.h:9-33).
#pragma once

#include <vector>

#include "KF_model.hpp"
#include "Mat.hpp"

namespace mkfbt {
namespace pfrao {

class ParticleFilter;

// One track {state d x 1, cov d x d, weight} (my_gmm.h:9-18).
class state_params {
public:
    Mat state;
    Mat cov;
    double weight = 0.0;
};

// GMM of mean-reverting random walks (my_gmm.h:20-33).  Components are kept
// on the host and uploaded to the filter's device tables on first use; the
// tracks live in HBM and are copied out by tracks_snapshot().
class my_gmm {
public:
    my_gmm();
    ~my_gmm();
    // my_gmm::loadGaussian (my_gmm.cpp:45-75): u 1 x d mean, s d x d covariance,
    // H = pca_proj (d x npose), m = pca_mean (1 x npose), weight w, gamma g.
    void loadGaussian(Mat u, Mat s, Mat &H, Mat &m, double w, double g);
    // my_gmm::resetTracker (my_gmm.cpp:30-42).
    void resetTracker(std::vector<int> bins);
    // Copy of the current tracks (gmm.tracks of the reference).
    std::vector<state_params> tracks_snapshot() const;

    std::vector<Mat> mean;
    std::vector<Mat> cov;
    std::vector<double> weight;
    std::vector<double> gamma;
    std::vector<KF_model> KFtracker;
    int nParticles = 0;
    // measurement selection (rows of H1, my_gmm.cpp:61-67) and noise R
    std::vector<int> sel{9, 10, 0, 1, 12, 13};
    Mat R;

private:
    friend class ParticleFilter;
    void upload();
    ParticleFilter *pf_ = nullptr;
    Mat proj_, pmean_;
    bool dirty_ = true;
};

}  // namespace pfrao
}  // namespace mkfbt
