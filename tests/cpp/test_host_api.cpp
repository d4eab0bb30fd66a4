// Drives the C++ mirror of the reference classes the way PFTracker does
// (src/pfPose.cpp:34-71 and 300-348): load the shipped left-arm GMM/PCA model,
// loadGaussian x K, resample(gmm.weight) + resetTracker, gmm.tracks read and
// written (my_gmm.h:31), then per-frame update with a 6 x N measurement
// matrix, getEstimator, getSamples, getSampleProb; plus KF_model, the pf2D
// filter (ParticleFilter(), predict(), my_gmm::sigma_i and det_s), and the
// PFTracker core (both arms, the whole per-frame pipeline).  Tracks, the
// update from written tracks and the pf2D GMM members are checked against the
// FP64 oracle (oracle/liboracle.so).  Exits non-zero on any failed check.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "mkfbt/pf2D.hpp"
#include "mkfbt/pf2DRao.hpp"
#include "mkfbt/pfPose.hpp"
#include "pf_oracle.h"
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

// Normwise relative error of n values.
static double relerr(const double *a, const double *b, size_t n)
{
    double num = 0.0, den = 0.0;
    for (size_t i = 0; i < n; ++i) {
        num += (a[i] - b[i]) * (a[i] - b[i]);
        den += b[i] * b[i];
    }
    return std::sqrt(num) / std::max(std::sqrt(den), 1e-300);
}

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
    const int d = 12;
    // the oracle's model: weights, gamma, means, covariances, H, BH, R of the mirror
    const mkfbt::pfrao::KF_model &k0 = pf.gmm.KFtracker[0];
    orc_model om{15, d, 6, pf.gmm.weight.data(), pf.gmm.gamma.data(), means.ptr(), covs.ptr(), k0.H.ptr(),
                 k0.BH.ptr(), k0.R.ptr()};
    {
        // gmm.tracks after resetTracker (my_gmm.cpp:30-42), read through a const view
        std::vector<double> xr((size_t)N * d), Pr((size_t)N * d * d);
        std::vector<int> b32(bins.begin(), bins.end());
        orc_reset_tracker(&om, N, b32.data(), xr.data(), Pr.data());
        const mkfbt::pfrao::TrackList &tl = pf.gmm.tracks;
        CHECK(tl.size() == (size_t)N);
        double e = 0.0;
        for (int j = 0; j < N; ++j) {
            CHECK(tl[j].state.rows == d && tl[j].state.cols == 1 && tl[j].cov.rows == d && tl[j].weight == 1.0 / N);
            e = std::max(e, relerr(tl[j].state.ptr(), &xr[(size_t)j * d], d));
            e = std::max(e, relerr(tl[j].cov.ptr(), &Pr[(size_t)j * d * d], (size_t)d * d));
        }
        std::printf("reset tracks: max rel err %.2e\n", e);
        CHECK(e < 1e-6);
    }
    {
        // write gmm.tracks: shift every state and scale every covariance; the
        // next device operation runs from the written tracks
        for (int j = 0; j < N; ++j) {
            mkfbt::pfrao::state_params &s = pf.gmm.tracks[j];
            for (int i = 0; i < d; ++i) s.state(i, 0) += 0.5 * std::sin(0.1 * j + i);
            for (double &v : s.cov.data) v *= 1.5;
        }
        std::vector<double> xw((size_t)N * d), Pw((size_t)N * d * d);
        for (int j = 0; j < N; ++j) {
            std::copy(pf.gmm.tracks[j].state.data.begin(), pf.gmm.tracks[j].state.data.end(), &xw[(size_t)j * d]);
            std::copy(pf.gmm.tracks[j].cov.data.begin(), pf.gmm.tracks[j].cov.data.end(), &Pw[(size_t)j * d * d]);
        }
        Mat ew = pf.getEstimator();                       // uploads the written tracks
        std::vector<double> xd((size_t)N * d), Pd((size_t)N * d * d), xbar(d);
        CHECK(pf_get_tracks(pf.handle(), xd.data(), Pd.data()) == 0);   // their FP32 device copy
        CHECK(relerr(xd.data(), xw.data(), xd.size()) < 1e-6 && relerr(Pd.data(), Pw.data(), Pd.size()) < 1e-6);
        orc_rbpf_estimate(N, d, xd.data(), xbar.data());
        CHECK(relerr(ew.ptr(), xbar.data(), d) < 1e-9);
        // one update from the written tracks (pf2DRao.cpp:125-158) against the oracle
        Mat z(6, N);
        for (int j = 0; j < N; ++j) {
            const double *xj = &xd[(size_t)j * d];
            for (int r = 0; r < 6; ++r) {
                double y = k0.BH(r, 0);
                for (int c = 0; c < d; ++c) y += k0.H(r, c) * xj[c];
                z(r, j) = (double)(float)(y + 3.0 * std::cos(0.37 * j + r));
            }
        }
        uint32_t uu[2];
        pf_philox_u32(1234, 0, pf.draws(), 2, uu);       // the two draws update makes next
        pf.update(z);
        CHECK(pf.lastStatus() == 0);
        std::vector<double> lw(N), xu((size_t)N * d), Pu((size_t)N * d * d);
        std::vector<int> anc(N), comp(N);
        double lml = 0.0;
        orc_rbpf_update_ref(&om, N, xd.data(), Pd.data(), z.ptr(), uu[0], uu[1], lw.data(), anc.data(), comp.data(),
                            &lml, xu.data(), Pu.data());
        const mkfbt::pfrao::TrackList &tl = pf.gmm.tracks;   // copied out after the update
        double ex = 0.0, eP = 0.0;
        for (int j = 0; j < N; ++j) {
            ex = std::max(ex, relerr(tl[j].state.ptr(), &xd[(size_t)j * d], d));
            eP = std::max(eP, relerr(tl[j].cov.ptr(), &Pd[(size_t)j * d * d], (size_t)d * d));
        }
        std::printf("update from written tracks: max rel err x %.2e P %.2e\n", ex, eP);
        CHECK(ex < 1e-4 && eP < 1e-4);
        // a wrongly shaped written track is refused at the next device operation
        pf.gmm.tracks[7].cov = Mat(3, 3);
        bool threw = false;
        try { pf.getEstimator(); } catch (const mkfbt::PfException &) { threw = true; }
        CHECK(threw);
        pf.gmm.resetTracker(bins);
    }
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
    const int N2 = 2000;
    mkfbt::pf2d::ParticleFilter p2(N2, 8, true, 7);
    for (int k = 0; k < 3; ++k) {
        Mat mu(1, 8), S(8, 8);
        for (int c = 0; c < 8; ++c) {
            mu(0, c) = 200 + 40 * k + 10 * c;
            for (int q = 0; q < 8; ++q) S(c, q) = 2500.0 * std::pow(0.3 + 0.1 * k, std::abs(c - q));
        }
        p2.gmm.loadGaussian(mu, S, 1.0 / 3);
        // sigma_i and det_s against the oracle's loadGaussian (pf2D.cpp:28-37)
        double Si[64], cn = 0.0;
        CHECK(orc_pf2d_load_gaussian(8, S.ptr(), Si, &cn) == 0);
        CHECK(p2.gmm.N == k + 1 && p2.gmm.sigma_i[k].rows == 8 && p2.gmm.sigma_i[k].cols == 8);
        CHECK(relerr(p2.gmm.sigma_i[k].ptr(), Si, 64) < 1e-12);
        CHECK(std::fabs(p2.gmm.det_s[k] - cn) <= 1e-12 * cn);
    }
    Mat meas(2, 2);
    meas(0, 0) = 420; meas(0, 1) = 300; meas(1, 0) = 220; meas(1, 1) = 260;
    for (int t = 0; t < 3; ++t) p2.update(meas);
    Mat est = p2.getEstimator();
    CHECK(est.rows == 1 && est.cols == 8);
    for (double v : est.data) CHECK(std::isfinite(v));
    {
        // predict() (pf2D.cpp:90-102): every particle moves by 5 n, n standard
        // normal; the weights are kept
        std::vector<double> a((size_t)N2 * 8), wa(N2), b((size_t)N2 * 8), wb(N2);
        CHECK(pf2d_get_particles(p2.handle(), a.data(), wa.data()) == 0);
        p2.predict();
        CHECK(pf2d_get_particles(p2.handle(), b.data(), wb.data()) == 0);
        CHECK(wa == wb);
        for (int c = 0; c < 8; ++c) {
            double s1 = 0.0, s2 = 0.0;
            for (int i = 0; i < N2; ++i) {
                const double dlt = b[(size_t)i * 8 + c] - a[(size_t)i * 8 + c];
                s1 += dlt;
                s2 += dlt * dlt;
            }
            const double mean = s1 / N2, sd = std::sqrt(s2 / N2 - mean * mean);
            CHECK(std::fabs(mean) < 0.6 && sd > 4.6 && sd < 5.4);   // 5 sigma bands of the sample moments
        }
    }
    {
        // ParticleFilter() (pf2D.h:29), then a sized filter moved in; the gmm
        // members are read as they are at each update
        mkfbt::pf2d::ParticleFilter p0;
        bool threw = false;
        try { p0.update(meas); } catch (const mkfbt::PfException &ex) { threw = ex.code == PF_ERR_STATE; }
        CHECK(threw);
        p0 = mkfbt::pf2d::ParticleFilter(500, 8, false, 9);
        Mat mu(1, 8), S = Mat::eye(8);
        for (int c = 0; c < 8; ++c) { mu(0, c) = 250 + 10 * c; S(c, c) = 4000.0; }
        p0.gmm.loadGaussian(mu, S, 1.0);
        p0.update(meas);
        p0.gmm.det_s[0] *= 2.0;                          // a member edited between updates
        p0.update(meas);
        p0.predict();
        for (double v : p0.getEstimator().data) CHECK(std::isfinite(v));
    }

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
