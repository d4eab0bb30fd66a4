// my_gmm.hpp -- mirror of the reference state_params / my_gmm (src/my_gmm.h:9-33).
#pragma once

#include <cstddef>
#include <vector>

#include "KF_model.hpp"
#include "Mat.hpp"

namespace mkfbt {

class PFTracker;

namespace pfrao {

class ParticleFilter;
class my_gmm;

// One track {state d x 1, cov d x d, weight} (my_gmm.h:9-18).
class state_params {
public:
    Mat state;
    Mat cov;
    double weight = 0.0;
};

// gmm.tracks (my_gmm.h:31): the filter's tracks as a list of state_params.
// The tracks live in HBM.  The list copies them to the host on the first
// access after a device operation changed them (update, update_candidates,
// resetTracker), and copies the host list back before the next device
// operation if it was accessed for writing since (any non-const access).
// Elements and iterators stay valid until the next device operation.  As in
// the reference, weight is 1/nParticles after every update; the filter
// never reads it.
class TrackList {
public:
    typedef std::vector<state_params>::iterator iterator;
    typedef std::vector<state_params>::const_iterator const_iterator;

    size_t size() const { pull(); return v_.size(); }
    bool empty() const { return size() == 0; }
    state_params &operator[](size_t i) { return vec()[i]; }
    const state_params &operator[](size_t i) const { pull(); return v_[i]; }
    state_params &at(size_t i) { return vec().at(i); }
    const state_params &at(size_t i) const { pull(); return v_.at(i); }
    iterator begin() { return vec().begin(); }
    iterator end() { return vec().end(); }
    const_iterator begin() const { pull(); return v_.begin(); }
    const_iterator end() const { pull(); return v_.end(); }
    void clear() { vec().clear(); }
    void push_back(const state_params &s) { vec().push_back(s); }
    // The whole list, writable (marks it for upload).
    std::vector<state_params> &vec() { pull(); dirty_ = true; return v_; }
    operator const std::vector<state_params> &() const { pull(); return v_; }

private:
    friend class my_gmm;
    friend class ParticleFilter;
    void pull() const;
    mutable std::vector<state_params> v_;
    mutable bool stale_ = false;   // the device holds newer tracks than v_
    bool dirty_ = false;           // v_ may differ from the device tracks
    const my_gmm *gmm_ = nullptr;
};

// GMM of mean-reverting random walks (my_gmm.h:20-33).  Components are kept
// on the host and uploaded to the filter's device tables on first use.
class my_gmm {
public:
    my_gmm();
    ~my_gmm();
    my_gmm(const my_gmm &) = delete;
    my_gmm &operator=(const my_gmm &) = delete;
    // my_gmm::loadGaussian (my_gmm.cpp:45-75): u 1 x d mean, s d x d covariance,
    // H = pca_proj (d x npose), m = pca_mean (1 x npose), weight w, gamma g.
    void loadGaussian(Mat u, Mat s, Mat &H, Mat &m, double w, double g);
    // my_gmm::resetTracker (my_gmm.cpp:30-42).
    void resetTracker(std::vector<int> bins);
    // A copy of the current tracks.
    std::vector<state_params> tracks_snapshot() const { return tracks; }

    std::vector<Mat> mean;
    std::vector<Mat> cov;
    std::vector<double> weight;
    std::vector<double> gamma;
    std::vector<KF_model> KFtracker;
    TrackList tracks;
    int nParticles = 0;
    // measurement selection (rows of H1, my_gmm.cpp:61-67) and noise R
    std::vector<int> sel{9, 10, 0, 1, 12, 13};
    Mat R;

private:
    friend class ParticleFilter;
    friend class TrackList;
    friend class ::mkfbt::PFTracker;
    void upload();
    // Before a device operation: the model, then the host tracks if written.
    void push();
    // After a device operation that changed the tracks.
    void device_changed() { tracks.stale_ = true; tracks.dirty_ = false; }
    int dims() const { return cov.empty() ? 0 : cov[0].rows; }
    ParticleFilter *pf_ = nullptr;
    Mat proj_, pmean_;
    bool dirty_ = true;
};

}  // namespace pfrao
}  // namespace mkfbt
