// Host sanitizer check (AddressSanitizer + UndefinedBehaviorSanitizer, built
// with g++ by tests/cpp/build.py) of the library's host-side code and of the
// FP64 oracle:
//   * pf_io.cpp: the yml / camera_info readers behind pf_yml_read_matrix,
//     pf_read_camera_info and pf_load_model_yml, on the shipped model files
//     and on malformed files written here;
//   * pf_model.cpp: model validation, the selection basis and the model
//     image (the host half of pf_load_model / pf_batch_load_model), on the
//     shipped models, synthetic d = 8, 12, 16, 24 models and invalid inputs;
//   * pf_host.cpp: the host-only members of the C++ mirror classes
//     (my_gmm::loadGaussian of pf2DRao and pf2D);
//   * the oracle: resampling with zero, NaN and infinite weights, the RBPF
//     and pf2D updates, the blur and the 3-D pose.
// This build has no GPU: pf_load_model runs its host half here, and the
// device entry points the mirror classes would reach report
// PF_ERR_UNSUPPORTED.  Exits non-zero on a failed check; a sanitizer report
// aborts the process.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <random>
#include <string>
#include <vector>

#include "../../mkfbodytracker_pdaf_amd/csrc/pf_internal.hpp"
#include "../../mkfbodytracker_pdaf_amd/csrc/pf_model.hpp"
#include "mkfbt/pf2D.hpp"
#include "mkfbt/pf2DRao.hpp"
#include "pf2d.h"
#include "pf_oracle.h"
#include "pf_rbpf.h"
#include "pf_tracker.h"

static int failures = 0;
#define CHECK(c)                                                              \
    do {                                                                      \
        if (!(c)) {                                                           \
            std::printf("CHECK failed: %s (line %d)\n", #c, __LINE__);       \
            ++failures;                                                       \
        }                                                                     \
    } while (0)

// ---------------------------------------------------------------- harness
static char g_msg[1024];

namespace mkfbt {
int set_error(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_msg, sizeof g_msg, fmt, ap);
    va_end(ap);
    return code;
}
int status_code(unsigned word) { return (int)(word & 3u); }
}  // namespace mkfbt

extern "C" {
const char *pf_last_error(void) { return g_msg; }
void pf_config_default(pf_config *c)
{
    std::memset(c, 0, sizeof *c);
    c->n_particles = 500;
}
// The host half of pf_load_model: validation, basis and model image.
int pf_load_model(pf_handle *, int32_t K, int32_t d, int32_t npose, int32_t m, const int32_t *sel,
                  const double *means, const double *covs, const double *weights, const double *gamma,
                  const double *pca_proj, const double *pca_mean, const double *R)
{
    mkfbt::HostModel hm;
    int st = mkfbt::host_model_from_args(hm, K, d, npose, m, sel, means, covs, weights, gamma, pca_proj, pca_mean, R);
    if (st) return st;
    mkfbt::ModelImage img;
    return mkfbt::make_model_image(hm, mkfbt::joint_order(0, hm.J), 1000, img);
}
#define NODEV(name, ...) int name(__VA_ARGS__) { return mkfbt::set_error(PF_ERR_UNSUPPORTED, "no device in this build"); }
NODEV(pf_create, pf_handle **, const pf_config *)
NODEV(pf_destroy, pf_handle *)
NODEV(pf_set_joint_clutter, pf_handle *, const int32_t *)
NODEV(pf_reset, pf_handle *, const int32_t *)
NODEV(pf_set_tracks, pf_handle *, const double *, const double *)
NODEV(pf_get_tracks, pf_handle *, double *, double *)
NODEV(pf_step, pf_handle *, const double *, uint32_t, uint32_t, float *, int32_t *, int32_t *, double *)
NODEV(pf_step_candidates, pf_handle *, const int32_t *, const double *, uint32_t, uint32_t, float *, int32_t *,
      int32_t *, double *)
NODEV(pf_estimate, pf_handle *, double *)
NODEV(pf_get_samples, pf_handle *, const double *, const double *, int32_t, int32_t, double, const double *, double *)
NODEV(pf_get_sample_prob, pf_handle *, const double *, const double *, int32_t, const double *, const double *,
      int32_t, double, double *, double *)
NODEV(pf_resample, pf_handle *, const double *, int32_t, int32_t, uint32_t, int32_t *)
NODEV(pf_kf_predict, pf_handle *, int32_t, int32_t, const double *, const double *, const double *, double *, double *)
NODEV(pf_kf_update, pf_handle *, int32_t, int32_t, int32_t, const double *, const double *, const double *,
      const double *, double *, double *)
NODEV(pf2d_create, pf2d_handle **, int32_t, int32_t, int32_t, int32_t)
NODEV(pf2d_destroy, pf2d_handle *)
NODEV(pf2d_init, pf2d_handle *, const double *)
NODEV(pf2d_load_gmm_precision, pf2d_handle *, int32_t, const double *, const double *, const double *, const double *)
NODEV(pf2d_update, pf2d_handle *, const double *, uint32_t, const double *, const double *, float *, int32_t *)
NODEV(pf2d_predict, pf2d_handle *, const double *)
NODEV(pf2d_estimate, pf2d_handle *, double *)
NODEV(pf_tracker_create, pf_tracker **, const pf_tracker_config *, pf_handle *, pf_handle *, int32_t, const double *,
      const double *, const double *, const double *)
NODEV(pf_tracker_destroy, pf_tracker *)
NODEV(pf_tracker_frame, pf_tracker *, const uint8_t *, const int32_t *, const pf_tracker_outputs *)
NODEV(pf_tracker_pose, pf_tracker *, const pf_tracker_outputs *)
void pf_tracker_config_default(pf_tracker_config *c) { std::memset(c, 0, sizeof *c); }
}

// ------------------------------------------------------------------ helpers
static std::string write_file(const std::string &dir, const char *name, const std::string &text)
{
    const std::string p = dir + "/" + name;
    std::ofstream(p, std::ios::binary) << text;
    return p;
}

static std::vector<double> spd(std::mt19937 &g, int d, double scale)
{
    std::normal_distribution<double> n(0.0, 1.0);
    std::vector<double> A((size_t)d * d), S((size_t)d * d, 0.0);
    for (double &v : A) v = n(g);
    for (int i = 0; i < d; ++i)
        for (int j = 0; j < d; ++j) {
            double s = 0.0;
            for (int k = 0; k < d; ++k) s += A[(size_t)i * d + k] * A[(size_t)j * d + k];
            S[(size_t)i * d + j] = scale * (s + (i == j ? d : 0.0));
        }
    return S;
}

// A synthetic my_gmm argument set: K components in d dimensions, npose = 2 m
// pose rows of which sel picks m.
struct Args {
    int K, d, npose, m;
    std::vector<int32_t> sel;
    std::vector<double> means, covs, weights, gamma, proj, pmean, R;
};

static Args synthetic(std::mt19937 &g, int K, int d, int m)
{
    Args a{K, d, 2 * m, m, {}, {}, {}, {}, {}, {}, {}, {}};
    std::normal_distribution<double> n(0.0, 1.0);
    std::uniform_real_distribution<double> u(0.1, 1.0);
    for (int r = 0; r < m; ++r) a.sel.push_back(2 * r);
    for (int k = 0; k < K; ++k) {
        for (int i = 0; i < d; ++i) a.means.push_back(n(g));
        const std::vector<double> S = spd(g, d, 0.5);
        a.covs.insert(a.covs.end(), S.begin(), S.end());
        a.weights.push_back(u(g));
        a.gamma.push_back(0.5 + 0.45 * u(g));
    }
    for (int i = 0; i < d * a.npose; ++i) a.proj.push_back(n(g));
    for (int i = 0; i < a.npose; ++i) a.pmean.push_back(300.0 * u(g));
    a.R.assign((size_t)m * m, 0.0);
    for (int i = 0; i < m; ++i) a.R[(size_t)i * m + i] = 100.0;
    return a;
}

static int load(const Args &a)
{
    return pf_load_model(nullptr, a.K, a.d, a.npose, a.m, a.sel.data(), a.means.data(), a.covs.data(),
                         a.weights.data(), a.gamma.data(), a.proj.data(), a.pmean.data(), a.R.data());
}

int main(int argc, char **argv)
{
    const std::string dir = argc > 1 ? argv[1] : "tests/golden";
    const std::string tmp = argc > 2 ? argv[2] : "/tmp";
    std::mt19937 g(12345);

    // ---- pf_io.cpp on the shipped files
    for (const char *f : {"data13D_PCA_100000_15_12.yml", "data23D_PCA_100000_15_12.yml"}) {
        const std::string p = dir + "/" + f;
        for (const char *key : {"means", "covs", "weights", "gamma", "pca_proj", "pca_mean"}) {
            int32_t r = 0, c = 0;
            CHECK(pf_yml_read_matrix(p.c_str(), key, &r, &c, nullptr, 0) == PF_OK && r > 0 && c > 0);
            std::vector<double> v((size_t)r * c);
            CHECK(pf_yml_read_matrix(p.c_str(), key, &r, &c, v.data(), (int64_t)v.size()) == PF_OK);
            for (double x : v) CHECK(std::isfinite(x));
            if (v.size() > 1)
                CHECK(pf_yml_read_matrix(p.c_str(), key, &r, &c, v.data(), (int64_t)v.size() - 1) == PF_ERR_ARG);
        }
        CHECK(pf_load_model_yml((pf_handle *)1, p.c_str(), nullptr, 6, nullptr, nullptr) == PF_OK);
    }
    {
        int32_t w = 0, h = 0;
        double K[9], D[5], P[12];
        CHECK(pf_read_camera_info((dir + "/cal.yml").c_str(), &w, &h, K, D, P) == PF_OK && w > 0 && h > 0);
    }
    // ---- malformed files: every one is an error code, none a crash
    const char *bad[] = {
        "",
        "%YAML:1.0\nmeans: !!opencv-matrix\n   rows: 2\n   cols: 2\n   dt: d\n   data: [ 1., 2., 3. ]\n",
        "%YAML:1.0\nmeans: !!opencv-matrix\n   rows: -3\n   cols: 2\n   dt: d\n   data: [ 1., 2. ]\n",
        "%YAML:1.0\nmeans: !!opencv-matrix\n   rows: 2000000000\n   cols: 2000000000\n   dt: d\n   data: [ 1. ]\n",
        "%YAML:1.0\nmeans: !!opencv-matrix\n   rows: 1\n   cols: 2\n   dt: d\n   data: [ 1., abc ]\n",
        "%YAML:1.0\nmeans: !!opencv-matrix\n   rows: 1\n   cols: 2\n   dt: d\n   data: [ 1., 2.",
        "%YAML:1.0\nmeans: !!opencv-matrix\n   rows: 1\n",
        "%YAML:1.0\nmeans: !!opencv-matrix\n   rows: 1\n   cols: 1\n   dt: d\n   data: [ 1e999 ]\n",
        "\x01\x02\x03\xff\xfe garbage \n\n\n::::[[[[",
    };
    int n_bad = 0;
    for (const char *text : bad) {
        const std::string p = write_file(tmp, "sanitize_bad.yml", text);
        int32_t r = 0, c = 0;
        std::vector<double> v(16);
        const int st = pf_yml_read_matrix(p.c_str(), "means", &r, &c, v.data(), (int64_t)v.size());
        if (st != PF_OK) ++n_bad;
        CHECK(pf_load_model_yml((pf_handle *)1, p.c_str(), nullptr, 6, nullptr, nullptr) != PF_OK);
    }
    CHECK(n_bad >= 7);
    CHECK(pf_yml_read_matrix((tmp + "/does_not_exist.yml").c_str(), "means", nullptr, nullptr, nullptr, 0) != PF_OK);

    // ---- pf_model.cpp: valid models of every supported size, then invalid ones
    for (int d : {8, 12, 16, 24}) {
        for (int K : {1, 15, 64}) {
            const int m = d == 8 ? 4 : (d == 24 ? 10 : 6);
            Args a = synthetic(g, K, d, m);
            CHECK(load(a) == PF_OK);
        }
    }
    {
        Args a = synthetic(g, 15, 12, 6);
        Args b = a;
        b.sel[2] = 99;
        CHECK(load(b) == PF_ERR_ARG);
        b = a;
        for (int i = 0; i < 12; ++i) b.covs[(size_t)3 * 144 + i * 12 + i] = -1.0;
        CHECK(load(b) == PF_ERR_NOT_SPD);
        b = a;
        b.gamma[4] = 1.5;
        CHECK(load(b) == PF_ERR_ARG);
        b = a;
        b.R[1] = 5.0;                       // not block diagonal in joints: rows 0/1 are joint 0
        b.R[2] = 5.0;
        CHECK(load(b) != PF_OK);
        b = a;
        for (double &w : b.weights) w = 0.0;
        CHECK(load(b) == PF_ERR_ARG);
        b = a;
        std::fill(b.proj.begin(), b.proj.end(), 0.0);   // H rank-deficient: generic (non-selection) image
        CHECK(load(b) == PF_OK);
        CHECK(pf_load_model(nullptr, 0, 12, 12, 6, a.sel.data(), a.means.data(), a.covs.data(), a.weights.data(),
                            a.gamma.data(), a.proj.data(), a.pmean.data(), a.R.data()) == PF_ERR_ARG);
        CHECK(pf_load_model(nullptr, 15, 13, 12, 6, a.sel.data(), a.means.data(), a.covs.data(), a.weights.data(),
                            a.gamma.data(), a.proj.data(), a.pmean.data(), a.R.data()) != PF_OK);
    }

    // ---- pf_host.cpp: host-only members of the mirror classes
    {
        mkfbt::pf2d::my_gmm gm;
        for (int k = 0; k < 5; ++k) {
            const std::vector<double> S = spd(g, 8, 300.0);
            mkfbt::Mat mu(1, 8, 0.0), Sm(8, 8, S.data());
            gm.loadGaussian(mu, Sm, 0.2);
        }
        CHECK(gm.N == 5 && gm.sigma_i.size() == 5 && gm.det_s.size() == 5);
        for (double v : gm.det_s) CHECK(std::isfinite(v) && v > 0.0);
        bool threw = false;
        try {
            gm.loadGaussian(mkfbt::Mat(1, 8, 0.0), mkfbt::Mat(8, 8, -1.0), 1.0);
        } catch (const mkfbt::PfException &e) {
            threw = e.code == PF_ERR_NOT_SPD;
        }
        CHECK(threw);
        mkfbt::pfrao::my_gmm rg;
        Args a = synthetic(g, 15, 12, 6);
        mkfbt::Mat proj(12, 12, a.proj.data()), pm(1, 12, a.pmean.data());
        bool bad_sel = false;                   // the default sel {9,10,0,1,12,13} needs 14 pose rows
        try {
            rg.loadGaussian(mkfbt::Mat(1, 12, a.means.data()), mkfbt::Mat(12, 12, a.covs.data()), proj, pm, 0.5, 0.9);
        } catch (const mkfbt::PfException &e) {
            bad_sel = e.code == PF_ERR_ARG;
        }
        CHECK(bad_sel && rg.KFtracker.empty());
        rg.sel.assign(a.sel.begin(), a.sel.end());
        for (int k = 0; k < 15; ++k)
            rg.loadGaussian(mkfbt::Mat(1, 12, a.means.data() + 12 * k), mkfbt::Mat(12, 12, a.covs.data() + 144 * k), proj,
                            pm, a.weights[k], a.gamma[k]);
        CHECK(rg.KFtracker.size() == 15 && rg.KFtracker[7].H.rows == 6 && rg.tracks.empty());
    }

    // ---- oracle
    {
        std::vector<double> w(1000), lw(1000);
        std::vector<int> out(777);
        std::uniform_real_distribution<double> u(0.0, 1.0);
        for (int i = 0; i < 1000; ++i) {
            w[i] = (i % 7 == 0) ? 0.0 : u(g);
            lw[i] = (i % 11 == 0) ? -INFINITY : std::log(u(g) + 1e-300) * 50.0;
        }
        lw[5] = NAN;
        for (uint32_t U : {0u, 1u, 0x7fffffffu, 0xffffffffu}) {
            CHECK(orc_resample_linear(w.data(), 1000, 777, U, out.data()) >= 0);
            for (int a : out) CHECK(a >= 0 && a < 1000);
            CHECK(orc_resample_log(lw.data(), 1000, 777, U, out.data()) >= 0);
            for (int a : out) CHECK(a >= 0 && a < 1000);
        }
        std::vector<double> z(1000, 0.0);
        CHECK(orc_resample_linear(z.data(), 1000, 777, 5, out.data()) == 1);      // no positive weight
        std::vector<double> ninf(1000, -INFINITY);
        CHECK(orc_resample_log(ninf.data(), 1000, 777, 5, out.data()) == 1);
    }
    {
        // RBPF update in candidate mode on a synthetic model (FP64 reference)
        const int K = 4, d = 12, m = 6, N = 64, cap = 4;
        Args a = synthetic(g, K, d, m);
        std::vector<double> H((size_t)m * d), BH(m);
        for (int r = 0; r < m; ++r) {
            for (int j = 0; j < d; ++j) H[(size_t)r * d + j] = a.proj[(size_t)j * a.npose + a.sel[r]];
            BH[r] = a.pmean[a.sel[r]];
        }
        orc_model om{K, d, m, a.weights.data(), a.gamma.data(), a.means.data(), a.covs.data(), H.data(), BH.data(),
                     a.R.data()};
        std::vector<int> bins(N);
        for (int j = 0; j < N; ++j) bins[j] = j % K;
        std::vector<double> x((size_t)N * d), P((size_t)N * d * d), lw(N), lml(1);
        orc_reset_tracker(&om, N, bins.data(), x.data(), P.data());
        const int clutter[3] = {0, 1, 1}, count[3] = {1, 3, 0};
        std::vector<double> cand((size_t)3 * cap * 2, 0.0);
        for (int b = 0; b < 3; ++b)
            for (int c = 0; c < cap; ++c) {
                cand[((size_t)b * cap + c) * 2] = BH[2 * b] + 3.0 * c;
                cand[((size_t)b * cap + c) * 2 + 1] = BH[2 * b + 1] - 2.0 * c;
            }
        orc_frame fr{3, cap, clutter, count, cand.data(), 0.95, 9.21, 1e-5};
        std::vector<int> anc(N), comp(N);
        for (int t = 0; t < 3; ++t)
            CHECK(orc_rbpf_update_pdaf(&om, N, x.data(), P.data(), &fr, 11 + t, 77 + t, lw.data(), anc.data(),
                                       comp.data(), lml.data(), nullptr, nullptr) >= 0);
        for (double v : x) CHECK(std::isfinite(v));
    }
    {
        // pf2D update, blur, 3-D pose
        const int N = 200, K = 3;
        std::vector<double> mu((size_t)K * 8), Sinv((size_t)K * 64), cn(K), pi(K, 1.0 / K);
        std::uniform_real_distribution<double> u(0.0, 1.0);
        for (int k = 0; k < K; ++k) {
            const std::vector<double> S = spd(g, 8, 400.0);
            CHECK(orc_pf2d_load_gaussian(8, S.data(), Sinv.data() + 64 * k, &cn[k]) == 0);
            for (int i = 0; i < 8; ++i) mu[(size_t)k * 8 + i] = 300.0 * u(g);
        }
        orc_pf2d_gmm gm{K, 8, pi.data(), mu.data(), Sinv.data(), cn.data()};
        std::vector<double> unif((size_t)N * 8), parts((size_t)N * 8), wts(N), lw(N), nrm((size_t)N * 8);
        for (double &v : unif) v = u(g);
        for (double &v : nrm) v = u(g) - 0.5;
        orc_pf2d_init(N, 8, 1, unif.data(), parts.data());
        std::vector<int> anc(N);
        const double meas[4] = {420.0, 300.0, 220.0, 260.0};
        for (int t = 0; t < 3; ++t)
            CHECK(orc_pf2d_update(&gm, N, parts.data(), wts.data(), lw.data(), anc.data(), meas, 99u + t, nrm.data(),
                                  unif.data(), 1, 0) >= 0);
        float taps[15];
        orc_gauss_taps(15, 3.0, taps);
        const int rows = 48, cols = 64;
        std::vector<uint8_t> img((size_t)rows * cols), dst(img.size());
        std::vector<float> tmpb(img.size());
        for (auto &p : img) p = (uint8_t)(g() & 255);
        CHECK(orc_blur_u8(img.data(), rows, cols, 15, taps, tmpb.data(), dst.data()) == 0);
        double e[22], p3[15];
        for (double &v : e) v = 0.3 * (u(g) - 0.5);
        (void)orc_get3Dpose(e, 525.0, 525.0, 320.0, 240.0, p3);
    }

    std::printf(failures ? "FAILED (%d)\n" : "OK\n", failures);
    return failures ? 1 : 0;
}
