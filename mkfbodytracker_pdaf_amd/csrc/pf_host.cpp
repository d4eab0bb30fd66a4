// pf_host.cpp -- C++ mirror of the reference classes (src/KF_model.h,
// src/my_gmm.h, src/pf2DRao.h, src/pf2D.h) on top of the C ABI.  Every
// filter computation is a C-ABI call into the gfx950 kernels; this file only
// keeps the reference's object model (public members, method names, argument
// meaning), draws the host-supplied random numbers and turns negative status
// codes into exceptions.
#include <algorithm>
#include <cmath>
#include <memory>
#include <mutex>
#include <string>

#include "../../include/mkfbt/KF_model.hpp"
#include "../../include/mkfbt/my_gmm.hpp"
#include "../../include/mkfbt/pf2D.hpp"
#include "../../include/mkfbt/pf2DRao.hpp"
#include "../../include/mkfbt/pfPose.hpp"
#include "../../include/pf2d.h"
#include "../../include/pf_rbpf.h"
#include "../../include/pf_tracker.h"
#include "pf_internal.hpp"

namespace mkfbt {

namespace {

int ok(int st)
{
    if (st < 0) throw PfException(st, pf_last_error());
    return st;
}

// Box-Muller normals from pairs of 32-bit Philox words (host randomness, the
// counterpart of cv::randn).
void normals_from(uint64_t seed, uint32_t stream, uint64_t first, size_t n, double *out)
{
    std::vector<uint32_t> u(2 * ((n + 1) / 2));
    philox_u32(seed, stream, first, (int64_t)u.size(), u.data());
    const double two_pi = 6.283185307179586;
    for (size_t i = 0; i < u.size() / 2; ++i) {
        const double a = ((double)u[2 * i] + 1.0) * 2.3283064365386963e-10;
        const double b = (double)u[2 * i + 1] * 2.3283064365386963e-10;
        const double r = std::sqrt(-2.0 * std::log(a));
        if (2 * i < n) out[2 * i] = r * std::cos(two_pi * b);
        if (2 * i + 1 < n) out[2 * i + 1] = r * std::sin(two_pi * b);
    }
}

// Single-state handle shared by the KF_model objects (KF_model has no
// filter of its own in the reference).
pf_handle *kf_handle()
{
    static std::mutex mu;
    static pf_handle *h = nullptr;
    std::lock_guard<std::mutex> lock(mu);
    if (!h) {
        pf_config c;
        pf_config_default(&c);
        c.n_particles = 1;
        ok(pf_create(&h, &c));
    }
    return h;
}

}  // namespace

namespace pfrao {

// ------------------------------------------------------------- KF_model
KF_model::KF_model() {}
KF_model::~KF_model() {}

void KF_model::predict(Mat &state, Mat &cov)
{
    const int d = state.rows * state.cols;
    if (F.rows != d || cov.rows != d || cov.cols != d) throw PfException(PF_ERR_ARG, "KF_model::predict: shape mismatch");
    ok(pf_kf_predict(kf_handle(), 1, d, F.ptr(), B.ptr(), Q.ptr(), state.ptr(), cov.ptr()));
}

void KF_model::update(Mat measurement, Mat &state, Mat &cov)
{
    const int d = state.rows * state.cols, m = measurement.rows * measurement.cols;
    if (H.rows != m || H.cols != d || cov.rows != d) throw PfException(PF_ERR_ARG, "KF_model::update: shape mismatch");
    ok(pf_kf_update(kf_handle(), 1, d, m, H.ptr(), BH.ptr(), R.ptr(), measurement.ptr(), state.ptr(), cov.ptr()));
}

// --------------------------------------------------------------- my_gmm
my_gmm::my_gmm() { tracks.gmm_ = this; }
my_gmm::~my_gmm() {}

void my_gmm::loadGaussian(Mat u, Mat s, Mat &H, Mat &m, double w, double g)
{
    const int d = s.rows, m6 = (int)sel.size(), npose = H.cols;
    if (s.cols != d || u.rows * u.cols != d || H.rows != d || m.rows * m.cols != npose)
        throw PfException(PF_ERR_ARG, "loadGaussian: u 1 x d, s d x d, H d x npose, m 1 x npose");
    for (int r = 0; r < m6; ++r)
        if (sel[r] < 0 || sel[r] >= npose) throw PfException(PF_ERR_ARG, "loadGaussian: sel outside the pose rows");
    mean.push_back(u);
    cov.push_back(s);
    weight.push_back(w);
    gamma.push_back(g);
    if (R.empty()) {
        R = Mat::zeros(m6, m6);
        for (int i = 0; i < m6; ++i) R(i, i) = 100.0;             // my_gmm.cpp:54
    }
    KF_model tracker;                                          // my_gmm.cpp:51-72
    tracker.Q = Mat(d, d);
    tracker.F = Mat::zeros(d, d);
    tracker.B = Mat(d, 1);
    for (int i = 0; i < d; ++i) {
        for (int j = 0; j < d; ++j) tracker.Q(i, j) = (1.0 - g * g) * s(i, j);
        tracker.F(i, i) = g;
        tracker.B(i, 0) = (1.0 - g) * u.data[i];
    }
    tracker.R = R;
    tracker.H = Mat(m6, d);
    tracker.BH = Mat(m6, 1);
    for (int r = 0; r < m6; ++r) {
        for (int j = 0; j < d; ++j) tracker.H(r, j) = H(j, sel[r]);
        tracker.BH(r, 0) = m.data[sel[r]];
    }
    KFtracker.push_back(tracker);
    proj_ = H;
    pmean_ = m;
    (void)npose;
    dirty_ = true;
}

void my_gmm::upload()
{
    if (!dirty_) return;
    if (!pf_ || mean.empty()) throw PfException(PF_ERR_STATE, "my_gmm: no components loaded");
    const int K = (int)mean.size(), d = cov[0].rows, npose = proj_.cols;
    std::vector<double> means((size_t)K * d), covs((size_t)K * d * d);
    for (int k = 0; k < K; ++k) {
        for (int i = 0; i < d; ++i) means[(size_t)k * d + i] = mean[k].data[i];
        for (int i = 0; i < d * d; ++i) covs[(size_t)k * d * d + i] = cov[k].data[i];
    }
    std::vector<int32_t> s32(sel.begin(), sel.end());
    ok(pf_load_model(pf_->h_, K, d, npose, (int)sel.size(), s32.data(), means.data(), covs.data(), weight.data(),
                     gamma.data(), proj_.ptr(), pmean_.ptr(), R.ptr()));
    dirty_ = false;
}

void my_gmm::resetTracker(std::vector<int> bins)
{
    upload();
    if ((int)bins.size() != nParticles) throw PfException(PF_ERR_ARG, "resetTracker: need one bin per particle");
    std::vector<int32_t> b(bins.begin(), bins.end());
    ok(pf_reset(pf_->h_, b.data()));
    device_changed();
}

void my_gmm::push()
{
    upload();
    if (!tracks.dirty_) return;
    const int d = dims(), N = nParticles;
    const std::vector<state_params> &v = tracks.v_;
    if ((int)v.size() != N) throw PfException(PF_ERR_STATE, "gmm.tracks must hold nParticles tracks");
    std::vector<double> x((size_t)N * d), P((size_t)N * d * d);
    for (int j = 0; j < N; ++j) {
        const state_params &t = v[j];
        if (t.state.rows * t.state.cols != d || t.cov.rows != d || t.cov.cols != d)
            throw PfException(PF_ERR_ARG, "gmm.tracks[" + std::to_string(j) + "] has the wrong shape");
        std::copy(t.state.data.begin(), t.state.data.end(), x.begin() + (size_t)j * d);
        std::copy(t.cov.data.begin(), t.cov.data.end(), P.begin() + (size_t)j * d * d);
    }
    ok(pf_set_tracks(pf_->h_, x.data(), P.data()));
    tracks.dirty_ = false;
}

void TrackList::pull() const
{
    if (!stale_) return;
    const int d = gmm_->dims(), N = gmm_->nParticles;
    std::vector<double> x((size_t)N * d), P((size_t)N * d * d);
    ok(pf_get_tracks(gmm_->pf_->handle(), x.data(), P.data()));
    v_.assign((size_t)N, state_params());
    for (int j = 0; j < N; ++j) {
        v_[j].state = Mat(d, 1, x.data() + (size_t)j * d);
        v_[j].cov = Mat(d, d, P.data() + (size_t)j * d * d);
        v_[j].weight = 1.0 / N;
    }
    stale_ = false;
}

// -------------------------------------------------------- ParticleFilter
ParticleFilter::ParticleFilter(int nParticles, uint64_t seed, int device) : seed_(seed)
{
    pf_config c;
    pf_config_default(&c);
    c.n_particles = nParticles;
    c.device = device;
    ok(pf_create(&h_, &c));
    gmm.nParticles = nParticles;
    gmm.pf_ = this;
}

ParticleFilter::~ParticleFilter()
{
    if (h_) pf_destroy(h_);
}

uint32_t ParticleFilter::draw_u32()
{
    uint32_t u;
    philox_u32(seed_, 0, draws_++, 1, &u);
    return u;
}

void ParticleFilter::setJointClutter(const std::vector<int> &clutter)
{
    gmm.push();
    std::vector<int32_t> c(clutter.begin(), clutter.end());
    ok(pf_set_joint_clutter(h_, c.data()));
}

void ParticleFilter::update(Mat measurement)
{
    gmm.push();
    if (measurement.cols != gmm.nParticles) throw PfException(PF_ERR_ARG, "update: measurement must be m x N");
    const uint32_t u1 = draw_u32(), u2 = draw_u32();
    last_status_ = ok(pf_step(h_, measurement.ptr(), u1, u2, nullptr, nullptr, nullptr, nullptr));
    gmm.device_changed();
}

void ParticleFilter::update_candidates(const std::vector<int> &count, const std::vector<double> &cand)
{
    gmm.push();
    std::vector<int32_t> c(count.begin(), count.end());
    const uint32_t u1 = draw_u32(), u2 = draw_u32();
    last_status_ = ok(pf_step_candidates(h_, c.data(), cand.data(), u1, u2, nullptr, nullptr, nullptr, nullptr));
    gmm.device_changed();
}

Mat ParticleFilter::getEstimator()
{
    gmm.push();
    Mat e(gmm.dims(), 1);
    ok(pf_estimate(h_, e.ptr()));
    return e;
}

Mat ParticleFilter::getSamples(Mat H, Mat M, int N, double scale)
{
    gmm.push();
    std::vector<double> n((size_t)2 * N);
    normals_from(seed_, 1, 2 * draws_, n.size(), n.data());
    draws_ += (uint64_t)N;
    Mat out(2, N);
    ok(pf_get_samples(h_, H.ptr(), M.ptr(), H.rows, N, scale, n.data(), out.ptr()));
    return out;
}

void ParticleFilter::getSampleProb(Mat H, Mat M, Mat input1, Mat input2, std::vector<double> &weight1,
                                   std::vector<double> &weight2, double scale)
{
    if (input1.cols != input2.cols) throw PfException(PF_ERR_ARG, "getSampleProb: inputs differ in size");
    gmm.push();
    weight1.assign((size_t)input1.cols, 0.0);
    weight2.assign((size_t)input2.cols, 0.0);
    ok(pf_get_sample_prob(h_, H.ptr(), M.ptr(), H.rows, input1.ptr(), input2.ptr(), input1.cols, scale,
                          weight1.data(), weight2.data()));
}

std::vector<int> ParticleFilter::resample(std::vector<double> weights, int N)
{
    std::vector<int32_t> out((size_t)N);
    last_status_ = ok(pf_resample(h_, weights.data(), (int)weights.size(), N, draw_u32(), out.data()));
    return std::vector<int>(out.begin(), out.end());
}

}  // namespace pfrao

namespace pf2d {

void my_gmm::loadGaussian(Mat mean_, Mat sigma_, double weight_)
{
    const int d = sigma_.rows;
    if (sigma_.cols != d || mean_.rows * mean_.cols != d)
        throw PfException(PF_ERR_ARG, "pf2D loadGaussian: mean 1 x d and sigma d x d");
    // invert(s, sigma_i, DECOMP_CHOLESKY) and det(s) from the same factor
    std::vector<double> L((size_t)d * d, 0.0);
    for (int j = 0; j < d; ++j) {
        double s = sigma_(j, j);
        for (int q = 0; q < j; ++q) s -= L[(size_t)j * d + q] * L[(size_t)j * d + q];
        if (!(s > 0.0)) throw PfException(PF_ERR_NOT_SPD, "pf2D loadGaussian: sigma is not positive definite");
        L[(size_t)j * d + j] = std::sqrt(s);
        for (int i = j + 1; i < d; ++i) {
            double t = 0.5 * (sigma_(i, j) + sigma_(j, i));
            for (int q = 0; q < j; ++q) t -= L[(size_t)i * d + q] * L[(size_t)j * d + q];
            L[(size_t)i * d + j] = t / L[(size_t)j * d + j];
        }
    }
    Mat inv(d, d);
    std::vector<double> y((size_t)d), xx((size_t)d);
    for (int c = 0; c < d; ++c) {
        for (int i = 0; i < d; ++i) {
            double t = (i == c) ? 1.0 : 0.0;
            for (int q = 0; q < i; ++q) t -= L[(size_t)i * d + q] * y[q];
            y[i] = t / L[(size_t)i * d + i];
        }
        for (int i = d - 1; i >= 0; --i) {
            double t = y[i];
            for (int q = i + 1; q < d; ++q) t -= L[(size_t)q * d + i] * xx[q];
            xx[i] = t / L[(size_t)i * d + i];
        }
        for (int i = 0; i < d; ++i) inv(i, c) = xx[i];
    }
    double logdet = 0.0;
    for (int i = 0; i < d; ++i) logdet += 2.0 * std::log(L[(size_t)i * d + i]);
    ++N;
    mean.push_back(mean_);
    sigma_i.push_back(inv);
    det_s.push_back(std::exp(-0.5 * d * std::log(6.283185307179586) - 0.5 * logdet));
    weight.push_back(weight_);
}

ParticleFilter::ParticleFilter() {}

ParticleFilter::ParticleFilter(int numParticles, int numDims, bool side1, uint64_t seed, int device)
    : N_(numParticles), d_(numDims), seed_(seed)
{
    ok(pf2d_create(&h_, numParticles, numDims, side1 ? 1 : 0, device));
    std::vector<double> u = uniforms((size_t)N_ * d_);
    ok(pf2d_init(h_, u.data()));
}

ParticleFilter::ParticleFilter(ParticleFilter &&o) noexcept
    : gmm(std::move(o.gmm)), h_(o.h_), N_(o.N_), d_(o.d_), seed_(o.seed_), draws_(o.draws_),
      gmm_sent_(std::move(o.gmm_sent_))
{
    o.h_ = nullptr;
}

ParticleFilter &ParticleFilter::operator=(ParticleFilter &&o) noexcept
{
    if (this != &o) {
        if (h_) pf2d_destroy(h_);
        gmm = std::move(o.gmm);
        h_ = o.h_;
        N_ = o.N_;
        d_ = o.d_;
        seed_ = o.seed_;
        draws_ = o.draws_;
        gmm_sent_ = std::move(o.gmm_sent_);
        o.h_ = nullptr;
    }
    return *this;
}

ParticleFilter::~ParticleFilter()
{
    if (h_) pf2d_destroy(h_);
}

void ParticleFilter::need_handle(const char *what) const
{
    if (!h_) throw PfException(PF_ERR_STATE, std::string("pf2D ") + what + ": the filter has no particles");
}

// Uploads the gmm members when they differ from the last upload.
void ParticleFilter::sync_gmm()
{
    const int K = gmm.N;
    if (K <= 0) throw PfException(PF_ERR_STATE, "pf2D update: no GMM components loaded");
    if ((int)gmm.mean.size() < K || (int)gmm.sigma_i.size() < K || (int)gmm.det_s.size() < K ||
        (int)gmm.weight.size() < K)
        throw PfException(PF_ERR_ARG, "pf2D update: gmm.N exceeds its members");
    std::vector<double> mu((size_t)K * d_), Si((size_t)K * d_ * d_), ds(gmm.det_s.begin(), gmm.det_s.begin() + K),
        w(gmm.weight.begin(), gmm.weight.begin() + K);
    for (int k = 0; k < K; ++k) {
        const Mat &m = gmm.mean[k], &S = gmm.sigma_i[k];
        if (m.rows * m.cols != d_ || S.rows != d_ || S.cols != d_)
            throw PfException(PF_ERR_ARG, "pf2D update: gmm component " + std::to_string(k) + " has the wrong shape");
        std::copy(m.data.begin(), m.data.end(), mu.begin() + (size_t)k * d_);
        std::copy(S.data.begin(), S.data.end(), Si.begin() + (size_t)k * d_ * d_);
    }
    std::vector<double> key;
    key.reserve(1 + mu.size() + Si.size() + 2 * (size_t)K);
    key.push_back((double)K);
    key.insert(key.end(), mu.begin(), mu.end());
    key.insert(key.end(), Si.begin(), Si.end());
    key.insert(key.end(), ds.begin(), ds.end());
    key.insert(key.end(), w.begin(), w.end());
    if (key == gmm_sent_) return;
    ok(pf2d_load_gmm_precision(h_, K, mu.data(), Si.data(), ds.data(), w.data()));
    gmm_sent_.swap(key);
}

std::vector<double> ParticleFilter::uniforms(size_t n)
{
    std::vector<uint32_t> u(n);
    philox_u32(seed_, 0, draws_, (int64_t)n, u.data());
    draws_ += n;
    std::vector<double> out(n);
    for (size_t i = 0; i < n; ++i) out[i] = (double)u[i] * 2.3283064365386963e-10;
    return out;
}

std::vector<double> ParticleFilter::normals(size_t n)
{
    std::vector<double> out(n);
    normals_from(seed_, 1, draws_, n, out.data());
    draws_ += n;
    return out;
}

void ParticleFilter::predict()
{
    need_handle("predict");
    std::vector<double> n = normals((size_t)N_ * 8);
    ok(pf2d_predict(h_, n.data()));
}

void ParticleFilter::update(Mat measurement)
{
    need_handle("update");
    if (measurement.rows != 2 || measurement.cols != 2) throw PfException(PF_ERR_ARG, "pf2D update: measurement is 2 x 2");
    sync_gmm();
    uint32_t u;
    philox_u32(seed_, 2, draws_++, 1, &u);
    std::vector<double> n = normals((size_t)N_ * 8), r = uniforms((size_t)N_ * d_);
    ok(pf2d_update(h_, measurement.ptr(), u, n.data(), r.data(), nullptr, nullptr));
}

Mat ParticleFilter::getEstimator()
{
    need_handle("getEstimator");
    Mat e(1, d_);
    ok(pf2d_estimate(h_, e.ptr()));
    return e;
}

}  // namespace pf2d

// ------------------------------------------------------------- PFTracker
namespace {

Mat read_yml(const std::string &path, const char *key)
{
    int32_t r = 0, c = 0;
    ok(pf_yml_read_matrix(path.c_str(), key, &r, &c, nullptr, 0));
    Mat m(r, c);
    ok(pf_yml_read_matrix(path.c_str(), key, &r, &c, m.ptr(), (int64_t)m.data.size()));
    return m;
}

}  // namespace

PFTracker::PFTracker(const std::string &left_yml, const std::string &right_yml, int numParticles, uint64_t seed,
                     int device, const pf_tracker_config *cfg)
{
    const std::string paths[2] = {left_yml, right_yml};
    for (int a = 0; a < 2; ++a) {
        Mat means = read_yml(paths[a], "means"), covs = read_yml(paths[a], "covs");
        Mat weights = read_yml(paths[a], "weights"), gam = read_yml(paths[a], "gamma");
        Mat proj = read_yml(paths[a], "pca_proj"), pmean = read_yml(paths[a], "pca_mean");
        const int K = means.rows, d = means.cols;
        if (covs.rows != K * d || covs.cols != d || (int)gam.data.size() != K || proj.rows != d)
            throw PfException(PF_ERR_ARG, "PFTracker: inconsistent model in " + paths[a]);
        auto pf = std::make_unique<pfrao::ParticleFilter>(numParticles, seed + (uint64_t)a, device);
        for (int i = 0; i < K; ++i) {                                 // pfPose.cpp:61-65
            Mat s(d, d, covs.ptr() + (size_t)d * d * i);
            pf->gmm.loadGaussian(means.row(i), s, proj, pmean, weights.data[i], gam.data[i]);
        }
        pf->gmm.resetTracker(pf->resample(pf->gmm.weight, numParticles));   // pfPose.cpp:68-71
        (a ? h2_pca : h1_pca) = proj;
        (a ? m2_pca : m1_pca) = pmean;
        (a ? pf2_ : pf1_) = std::move(pf);
    }
    npose_ = h1_pca.cols;
    if (h2_pca.cols != npose_) throw PfException(PF_ERR_ARG, "PFTracker: the arms' pose sizes differ");
    pf_tracker_config c;
    pf_tracker_config_default(&c);
    if (cfg) c = *cfg;
    else c.seed = seed;
    ok(pf_tracker_create(&tr_, &c, pf1_->handle(), pf2_->handle(), npose_, h1_pca.ptr(), m1_pca.ptr(),
                         h2_pca.ptr(), m2_pca.ptr()));
}

PFTracker::~PFTracker()
{
    if (tr_) pf_tracker_destroy(tr_);
}

PoseFrame PFTracker::run(const uint8_t *likelihood, const int32_t *roi, bool update)
{
    PoseFrame f;
    std::vector<double> pose((size_t)2 * npose_, 0.0), p3(30, 0.0);
    f.tf = Mat::zeros(8, 3);
    f.pose2d = Mat::zeros(8, 2);
    pf_tracker_outputs o = {f.lml, pose.data(), p3.data(), f.tf.ptr(), f.pose2d.ptr()};
    pf1_->gmm.push();
    pf2_->gmm.push();
    f.status = update ? ok(pf_tracker_frame(tr_, likelihood, roi, &o)) : ok(pf_tracker_pose(tr_, &o));
    if (update) {
        pf1_->gmm.device_changed();
        pf2_->gmm.device_changed();
    }
    f.e1 = Mat(1, npose_, pose.data());
    f.e2 = Mat(1, npose_, pose.data() + npose_);
    f.p3D1 = Mat(3, 5, p3.data());
    f.p3D2 = Mat(3, 5, p3.data() + 15);
    return f;
}

PoseFrame PFTracker::callback(const uint8_t *likelihood, const int32_t *roi) { return run(likelihood, roi, true); }

PoseFrame PFTracker::pose() { return run(nullptr, nullptr, false); }

}  // namespace mkfbt
