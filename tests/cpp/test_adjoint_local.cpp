// Compares the hand-derived reverse of the joint-local function
// (local_joint_rev) with its forward-mode dual-number derivatives
// (local_joint_dual) on random joints, certain and PDAF.  Host only.
// Prints the largest relative error and exits 1 if it exceeds 2e-3.
#include <cmath>
#include <cstdio>
#include <random>

#include "../../mkfbodytracker_pdaf_amd/csrc/pf_adjoint_local.hpp"

using namespace mkfbt::adj;

int main()
{
    std::mt19937 rng(12345);
    std::normal_distribution<float> nd(0.f, 1.f);
    std::uniform_real_distribution<float> ud(0.f, 1.f);
    double worst = 0.0;
    int n_checked = 0;
    for (int trial = 0; trial < 4000; ++trial) {
        const bool clutter = trial & 1;
        const PdafConst k{9.21034f, 0.95f / (3.0f / (640.f * 480.f)), 1.f - 0.95f * 0.99f};
        // S = P_rr + R with P_rr random SPD, R = 100 I
        const float a = 30.f * ud(rng), b = 30.f * ud(rng), c = 20.f * nd(rng);
        const float s00 = a * a + c * c * 0.5f + 100.f, s11 = b * b + c * c * 0.5f + 100.f;
        const float s01 = 0.5f * c * (a + b) * 0.3f;
        const float y0 = 300.f + 50.f * nd(rng), y1 = 200.f + 50.f * nd(rng);
        const float sd = sqrtf(0.5f * (s00 + s11));
        float cz[32];
        const int cnt = clutter ? 1 + (int)(ud(rng) * 15.f) : 1;
        for (int i = 0; i < cnt; ++i) {
            cz[2 * i] = y0 + 1.4f * sd * nd(rng);
            cz[2 * i + 1] = y1 + 1.4f * sd * nd(rng);
        }
        const float gb0 = nd(rng), gb1 = nd(rng), nb00 = 10.f * nd(rng), nb01 = 10.f * nd(rng),
                    nb11 = 10.f * nd(rng), ellb = nd(rng);
        const LocalDual D = local_joint_dual(k, y0, y1, s00, s01, s11, clutter, cnt, cz);
        const LocalRev R = local_joint_rev(k, y0, y1, s00, s01, s11, clutter, cnt, cz, gb0, gb1, nb00, nb01, nb11,
                                           ellb);
        double ref[5];
        for (int q = 0; q < 5; ++q)
            ref[q] = (double)gb0 * D.g0.d[q] + (double)gb1 * D.g1.d[q] + (double)nb00 * D.n00.d[q] +
                     2.0 * nb01 * D.n01.d[q] + (double)nb11 * D.n11.d[q] + (double)ellb * D.ell.d[q];
        const double got[5] = {R.yb0, R.yb1, R.sb00, R.sb01, R.sb11};
        double nr = 0.0, ne = 0.0;
        for (int q = 0; q < 5; ++q) {
            nr += ref[q] * ref[q];
            ne += (got[q] - ref[q]) * (got[q] - ref[q]);
        }
        const double vals[5] = {R.g0 - D.g0.v, R.g1 - D.g1.v, R.n00 - D.n00.v, R.n01 - D.n01.v, R.n11 - D.n11.v};
        double vr = 0.0;
        for (double v : vals) vr = std::fmax(vr, std::fabs(v));
        const double rel = std::sqrt(ne) / std::fmax(std::sqrt(nr), 1e-30);
        if (!(rel <= worst)) worst = rel;
        if (vr > 1e-6 * (1.0 + std::fabs(D.n00.v) + std::fabs(D.g0.v))) {
            std::printf("value mismatch at trial %d: %g\n", trial, vr);
            return 1;
        }
        ++n_checked;
    }
    std::printf("checked %d joints, max relative adjoint error %.3g\n", n_checked, worst);
    return worst <= 2e-3 ? 0 : 1;
}
