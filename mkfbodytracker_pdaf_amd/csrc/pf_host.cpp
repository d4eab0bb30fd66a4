// pf_host.cpp -- C++ mirror of the reference classes (src/KF_model.h,
// src/my_gmm.h, src/pf2DRao.h, src/pf2D.h) on top of the C ABI.  Every
// filter computation is a C-ABI call into the gfx950 kernels; this file only
// keeps the reference's object model (public members, method names, argument
// meaning), draws the host-supplied random numbers and turns negative status
// codes into exceptions.
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
my_gmm::my_gmm() {}
my_gmm::~my_gmm() {}

void my_gmm::loadGaussian(Mat u, Mat s, Mat &H, Mat &m, double w, double g)
{
    const int d = s.rows, m6 = (int)sel.size(), npose = H.cols;
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
}

std::vector<state_params> my_gmm::tracks_snapshot() const
{
    const int d = cov.empty() ? 0 : cov[0].rows, N = nParticles;
    std::vector<double> x((size_t)N * d), P((size_t)N * d * d);
    ok(pf_get_tracks(pf_->h_, x.data(), P.data()));
    std::vector<state_params> out((size_t)N);
    for (int j = 0; j < N; ++j) {
        out[j].state = Mat(d, 1, x.data() + (size_t)j * d);
        out[j].cov = Mat(d, d, P.data() + (size_t)j * d * d);
        out[j].weight = 1.0 / N;
    }
    return out;
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
    gmm.upload();
    std::vector<int32_t> c(clutter.begin(), clutter.end());
    ok(pf_set_joint_clutter(h_, c.data()));
}

void ParticleFilter::update(Mat measurement)
{
    gmm.upload();
    if (measurement.cols != gmm.nParticles) throw PfException(PF_ERR_ARG, "update: measurement must be m x N");
    const uint32_t u1 = draw_u32(), u2 = draw_u32();
    last_status_ = ok(pf_step(h_, measurement.ptr(), u1, u2, nullptr, nullptr, nullptr, nullptr));
}

void ParticleFilter::update_candidates(const std::vector<int> &count, const std::vector<double> &cand)
{
    gmm.upload();
    std::vector<int32_t> c(count.begin(), count.end());
    const uint32_t u1 = draw_u32(), u2 = draw_u32();
    last_status_ = ok(pf_step_candidates(h_, c.data(), cand.data(), u1, u2, nullptr, nullptr, nullptr, nullptr));
}

Mat ParticleFilter::getEstimator()
{
    const int d = gmm.cov.empty() ? 0 : gmm.cov[0].rows;
    Mat e(d, 1);
    ok(pf_estimate(h_, e.ptr()));
    return e;
}

Mat ParticleFilter::getSamples(Mat H, Mat M, int N, double scale)
{
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
    mean.push_back(mean_);
    sigma.push_back(sigma_);
    weight.push_back(weight_);
    ++N;
    const int d = sigma_.rows;
    std::vector<double> mu((size_t)N * d), S((size_t)N * d * d);
    for (int k = 0; k < N; ++k) {
        for (int i = 0; i < d; ++i) mu[(size_t)k * d + i] = mean[k].data[i];
        for (int i = 0; i < d * d; ++i) S[(size_t)k * d * d + i] = sigma[k].data[i];
    }
    ok(pf2d_load_gmm(pf_->h_, N, mu.data(), S.data(), weight.data()));
}

ParticleFilter::ParticleFilter(int numParticles, int numDims, bool side1, uint64_t seed, int device)
    : N_(numParticles), d_(numDims), seed_(seed)
{
    ok(pf2d_create(&h_, numParticles, numDims, side1 ? 1 : 0, device));
    gmm.pf_ = this;
    std::vector<double> u = uniforms((size_t)N_ * d_);
    ok(pf2d_init(h_, u.data()));
}

ParticleFilter::~ParticleFilter()
{
    if (h_) pf2d_destroy(h_);
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

void ParticleFilter::update(Mat measurement)
{
    if (measurement.rows != 2 || measurement.cols != 2) throw PfException(PF_ERR_ARG, "pf2D update: measurement is 2 x 2");
    uint32_t u;
    philox_u32(seed_, 2, draws_++, 1, &u);
    std::vector<double> n = normals((size_t)N_ * 8), r = uniforms((size_t)N_ * d_);
    ok(pf2d_update(h_, measurement.ptr(), u, n.data(), r.data(), nullptr, nullptr));
}

Mat ParticleFilter::getEstimator()
{
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
    f.status = update ? ok(pf_tracker_frame(tr_, likelihood, roi, &o)) : ok(pf_tracker_pose(tr_, &o));
    f.e1 = Mat(1, npose_, pose.data());
    f.e2 = Mat(1, npose_, pose.data() + npose_);
    f.p3D1 = Mat(3, 5, p3.data());
    f.p3D2 = Mat(3, 5, p3.data() + 15);
    return f;
}

PoseFrame PFTracker::callback(const uint8_t *likelihood, const int32_t *roi) { return run(likelihood, roi, true); }

PoseFrame PFTracker::pose() { return run(nullptr, nullptr, false); }

}  // namespace mkfbt
