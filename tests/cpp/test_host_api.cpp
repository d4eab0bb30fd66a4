// Drives the C++ mirror of the reference classes the way PFTracker does
// (src/pfPose.cpp:34-71 and 300-348): load the shipped left-arm GMM/PCA model,
// loadGaussian x K, resample(gmm.weight) + resetTracker, then per-frame update
// with a 6 x N measurement matrix, getEstimator, getSamples, getSampleProb;
// plus KF_model, the pf2D filter, and the PFTracker core (both arms, the whole
// per-frame pipeline).  Exits non-zero on any failed check.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "mkfbt/pf2D.hpp"
#include "mkfbt/pf2DRao.hpp"
#include "mkfbt/pfPose.hpp"
#include "pf_rbpf.h"

using mkfbt::Mat;

static int failures = 0;
#define CHECK(c)                                                              \
    do {                                                                      \
        if (!(c)) {                                                           \
            std::printf("CHECK failed: %s (line %d)\n", #c, __LINE__);       \
            ++failures;                                                       \
        }                                                                     \
    } while (0)

static Mat read(const std::string &path, const char *key)
{
    int32_t r = 0, c = 0;
    pf_yml_read_matrix(path.c_str(), key, &r, &c, nullptr, 0);
    Mat m(r, c);
    if (pf_yml_read_matrix(path.c_str(), key, &r, &c, m.ptr(), (int64_t)m.data.size()) != 0)
        std::printf("read %s failed: %s\n", key, pf_last_error());
    return m;
}

int main(int argc, char **argv)
{
    const std::string dir = argc > 1 ? argv[1] : "tests/golden";
    const std::string f = dir + "/data13D_PCA_100000_15_12.yml";
    Mat means = read(f, "means"), covs = read(f, "covs"), weights = read(f, "weights");
    Mat proj = read(f, "pca_proj"), pmean = read(f, "pca_mean"), gam = read(f, "gamma");
    CHECK(means.rows == 15 && means.cols == 12 && proj.rows == 12 && proj.cols == 22);

    const int N = 500;                                   // pfPose.cpp:57
    mkfbt::pfrao::ParticleFilter pf(N, 1234);
    for (int i = 0; i < means.rows; ++i) {              // pfPose.cpp:61-65
        Mat s(12, 12, covs.ptr() + (size_t)12 * 12 * i);
        pf.gmm.loadGaussian(means.row(i), s, proj, pmean, weights.at(0, i), gam.data[i]);
    }
    CHECK(pf.gmm.KFtracker.size() == 15 && pf.gmm.KFtracker[0].H.rows == 6);
    CHECK(std::fabs(pf.gmm.KFtracker[3].F(0, 0) - gam.data[3]) < 1e-15);
    std::vector<int> bins = pf.resample(pf.gmm.weight, N);   // pfPose.cpp:68-71
    CHECK((int)bins.size() == N);
    for (int i = 1; i < N; ++i) CHECK(bins[i] >= bins[i - 1]);
    pf.gmm.resetTracker(bins);
    std::vector<mkfbt::pfrao::state_params> tr = pf.gmm.tracks_snapshot();
    CHECK(std::fabs(tr[0].state(3, 0) - means(bins[0], 3)) < 1e-3 * (1 + std::fabs(means(bins[0], 3))));

    // measurement rows: head u,v; hand u,v; neck u,v (pfPose.cpp:303-323)
    Mat e0 = pf.getEstimator();
    Mat Hp = proj.t(), M = pmean.t();
    for (int t = 0; t < 5; ++t) {
        Mat z(6, N);
        for (int j = 0; j < N; ++j) {
            z(0, j) = 324 + t;  z(1, j) = 75 + t;
            z(2, j) = 388 + 0.1 * (j % 17); z(3, j) = 281 - t;
            z(4, j) = 326; z(5, j) = 129;
        }
        pf.update(z);
        CHECK(pf.lastStatus() == 0);
    }
    Mat e = pf.getEstimator();
    CHECK(e.rows == 12 && e.cols == 1);
    double mean0 = 0.0;
    tr = pf.gmm.tracks_snapshot();
    for (int j = 0; j < N; ++j) mean0 += tr[j].state(0, 0) / N;
    CHECK(std::fabs(mean0 - e(0, 0)) < 1e-6 * (1 + std::fabs(mean0)));
    Mat samples = pf.getSamples(Hp, M, 10 * N, 60.0);   // pfPose.cpp:222
    CHECK(samples.rows == 2 && samples.cols == 10 * N);
    std::vector<double> w1, w2;
    pf.getSampleProb(Hp, M, samples, samples, w1, w2, 60.0);
    CHECK(w1.size() == (size_t)(10 * N) && w1 == w2);
    for (double w : w1) CHECK(std::isfinite(w) && w >= 0.0);

    // KF_model on one state (KF_model.cpp:11-25)
    mkfbt::pfrao::KF_model kf = pf.gmm.KFtracker[0];
    Mat x = means.row(0).t(), P(12, 12, covs.ptr());
    kf.predict(x, P);
    Mat zz(6, 1);
    for (int r = 0; r < 6; ++r) zz(r, 0) = pmean.data[pf.gmm.sel[r]];
    const double tr0 = [&] { double s = 0; for (int i = 0; i < 12; ++i) s += P(i, i); return s; }();
    kf.update(zz, x, P);
    double tr1 = 0;
    for (int i = 0; i < 12; ++i) tr1 += P(i, i);
    CHECK(tr1 < tr0);                                   // an update never increases the trace

    // pf2D (pf2D.cpp:44-210)
    mkfbt::pf2d::ParticleFilter p2(2000, 8, true, 7);
    for (int k = 0; k < 3; ++k) {
        Mat mu(1, 8), S = Mat::eye(8);
        for (int c = 0; c < 8; ++c) { mu(0, c) = 200 + 40 * k + 10 * c; S(c, c) = 2500.0; }
        p2.gmm.loadGaussian(mu, S, 1.0 / 3);
    }
    Mat meas(2, 2);
    meas(0, 0) = 420; meas(0, 1) = 300; meas(1, 0) = 220; meas(1, 1) = 260;
    for (int t = 0; t < 3; ++t) p2.update(meas);
    Mat est = p2.getEstimator();
    CHECK(est.rows == 1 && est.cols == 8);
    for (double v : est.data) CHECK(std::isfinite(v));

    // errors surface as exceptions, like cv::Exception in the reference
    bool threw = false;
    try { pf.update(Mat(6, N + 1)); } catch (const mkfbt::PfException &) { threw = true; }
    CHECK(threw);

    // PFTracker core (pfPose.cpp:7-74, 332-377): a likelihood image with blobs
    // at the two mean hand positions, the face ROI on the mean head
    {
        mkfbt::PFTracker trk(dir + "/data13D_PCA_100000_15_12.yml", dir + "/data23D_PCA_100000_15_12.yml", 500, 3);
        std::vector<uint8_t> img((size_t)480 * 640, 0);
        const double hx[2] = {388.0, 264.0}, hy[2] = {281.0, 268.0};
        for (int y = 0; y < 480; ++y)
            for (int x = 0; x < 640; ++x) {
                double v = 0.0;
                for (int a = 0; a < 2; ++a)
                    v += 230.0 * std::exp(-((x - hx[a]) * (x - hx[a]) + (y - hy[a]) * (y - hy[a])) / 98.0);
                img[(size_t)y * 640 + x] = (uint8_t)std::lround(v > 255.0 ? 255.0 : v);
            }
        const int32_t roi[4] = {304, 50, 40, 48};
        mkfbt::PoseFrame f;
        for (int t = 0; t < 4; ++t) {
            f = trk.callback(img.data(), roi);
            CHECK(f.status == 0 && std::isfinite(f.lml[0]) && std::isfinite(f.lml[1]));
        }
        CHECK(f.e1.cols == 22 && f.p3D1.rows == 3 && f.p3D1.cols == 5 && f.tf.rows == 8);
        CHECK(f.tf(0, 0) == 0.0 && f.tf(0, 1) == 0.0 && f.tf(0, 2) == 0.0);   // Head relative to itself
        for (double v : f.e1.data) CHECK(std::isfinite(v));
        for (double v : f.p3D2.data) CHECK(std::isfinite(v));
        // the tracked hands stay near the blobs (image coordinates of e1, e2)
        CHECK(std::fabs(f.pose2d(0, 0) - hx[0]) < 40.0 && std::fabs(f.pose2d(0, 1) - hy[0]) < 40.0);
        CHECK(std::fabs(f.pose2d(1, 0) - hx[1]) < 40.0 && std::fabs(f.pose2d(1, 1) - hy[1]) < 40.0);
        mkfbt::PoseFrame g = trk.pose();
        CHECK(g.e1.data == f.e1.data && g.tf.data == f.tf.data);
        mkfbt::PoseFrame none = trk.callback(img.data(), nullptr);           // no face this frame
        for (double v : none.pose2d.data) CHECK(v == 0.0);
        bool bad = false;
        const int32_t empty_roi[4] = {304, 50, 0, 48};
        try { trk.callback(img.data(), empty_roi); } catch (const mkfbt::PfException &) { bad = true; }
        CHECK(bad);
    }

    std::printf(failures ? "FAILED (%d)\n" : "OK\n", failures);
    return failures ? 1 : 0;
}
