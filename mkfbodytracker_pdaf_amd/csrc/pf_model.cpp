// pf_model.cpp -- host-side model preparation (see pf_model.hpp).
#include "pf_model.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "../../include/pf_rbpf.h"
#include "pf_common.hpp"
#include "pf_internal.hpp"

namespace mkfbt {

namespace {

constexpr int kMaxJointsHost = 8;

// Gauss-Jordan inverse with partial pivoting (FP64, n <= 32).
bool invert(int n, const std::vector<double> &A, std::vector<double> &Ai)
{
    std::vector<double> a = A;
    Ai.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) Ai[(size_t)i * n + i] = 1.0;
    for (int c = 0; c < n; ++c) {
        int p = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(a[(size_t)r * n + c]) > std::fabs(a[(size_t)p * n + c])) p = r;
        if (a[(size_t)p * n + c] == 0.0) return false;
        if (p != c)
            for (int j = 0; j < n; ++j) {
                std::swap(a[(size_t)c * n + j], a[(size_t)p * n + j]);
                std::swap(Ai[(size_t)c * n + j], Ai[(size_t)p * n + j]);
            }
        const double piv = a[(size_t)c * n + c];
        for (int j = 0; j < n; ++j) {
            a[(size_t)c * n + j] /= piv;
            Ai[(size_t)c * n + j] /= piv;
        }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            const double f = a[(size_t)r * n + c];
            if (f == 0.0) continue;
            for (int j = 0; j < n; ++j) {
                a[(size_t)r * n + j] -= f * a[(size_t)c * n + j];
                Ai[(size_t)r * n + j] -= f * Ai[(size_t)c * n + j];
            }
        }
    }
    return true;
}

// State basis in which the measurement matrix is a row selection.  Because
// every particle shares H and F = g I commutes with any change of basis
// (my_gmm.cpp:56,71), the filter can run on x' = T x, P' = T P T^T with
//   T = [H rows in processing order ; s * orthonormal complement of row(H)],
// so joint slot s observes rows 2s, 2s+1 of x' directly and H P costs
// nothing.  Likelihoods and weights are basis independent.  Returns false
// (T = I, dense H kept) if H is rank deficient or T is badly conditioned.
bool make_basis(const HostModel &hm, unsigned order, std::vector<double> &T, std::vector<double> &Ti)
{
    const int d = hm.D, m = hm.m, J = hm.J;
    T.assign((size_t)d * d, 0.0);
    if (m > d) return false;
    std::vector<double> rows((size_t)m * d);
    for (int sl = 0; sl < J; ++sl) {
        const int b = (int)((order >> (4 * sl)) & 15u);
        for (int r = 0; r < 2; ++r)
            for (int c = 0; c < d; ++c) rows[(size_t)(2 * sl + r) * d + c] = hm.H[(size_t)(2 * b + r) * d + c];
    }
    std::vector<std::vector<double>> q;
    double fro = 0.0;
    for (double v : rows) fro += v * v;
    const double scale = std::sqrt(fro / m);
    auto orth = [&](std::vector<double> v, double ref) -> bool {
        for (int pass = 0; pass < 2; ++pass)
            for (const auto &u : q) {
                double dot = 0.0;
                for (int c = 0; c < d; ++c) dot += u[c] * v[c];
                for (int c = 0; c < d; ++c) v[c] -= dot * u[c];
            }
        double n2 = 0.0;
        for (int c = 0; c < d; ++c) n2 += v[c] * v[c];
        if (!(std::sqrt(n2) > 1e-8 * ref)) return false;
        const double nv = std::sqrt(n2);
        for (int c = 0; c < d; ++c) v[c] /= nv;
        q.push_back(v);
        return true;
    };
    for (int r = 0; r < m; ++r) {
        std::vector<double> v(rows.begin() + (size_t)r * d, rows.begin() + (size_t)(r + 1) * d);
        double n2 = 0.0;
        for (double x : v) n2 += x * x;
        if (!orth(v, std::sqrt(n2))) return false;
    }
    for (int i = 0; i < d && (int)q.size() < d; ++i) {
        std::vector<double> e((size_t)d, 0.0);
        e[i] = 1.0;
        orth(e, 1e-2);
    }
    if ((int)q.size() != d) return false;
    for (int r = 0; r < m; ++r)
        for (int c = 0; c < d; ++c) T[(size_t)r * d + c] = rows[(size_t)r * d + c];
    for (int r = m; r < d; ++r)
        for (int c = 0; c < d; ++c) T[(size_t)r * d + c] = scale * q[r][c];
    if (!invert(d, T, Ti)) return false;
    double nt = 0.0, ni = 0.0;
    for (double v : T) nt += v * v;
    for (double v : Ti) ni += v * v;
    return std::sqrt(nt * ni) < 1e6;
}

}  // namespace

int host_model_from_args(HostModel &hm, int K, int d, int npose, int m, const int32_t *sel,
                         const double *means, const double *covs, const double *weights,
                         const double *gamma, const double *pca_proj, const double *pca_mean,
                         const double *R)
{
    if (K <= 0 || K > 256) return set_error(PF_ERR_ARG, "K=%d out of range (1..256)", K);
    if (!(d == 8 || d == 12 || d == 16 || d == 24))
        return set_error(PF_ERR_UNSUPPORTED, "state dimension %d not supported (8, 12, 16, 24)", d);
    static const int32_t ref_sel[6] = {9, 10, 0, 1, 12, 13};  // my_gmm.cpp:61-67
    if (!sel) {
        if (m != 6) return set_error(PF_ERR_ARG, "default selection needs m == 6");
        sel = ref_sel;
    }
    if (m <= 0 || (m & 1) || m > 2 * kMaxJointsHost)
        return set_error(PF_ERR_ARG, "m=%d must be even and <= %d", m, 2 * kMaxJointsHost);
    if (!means || !covs || !weights || !gamma || !pca_proj || !pca_mean)
        return set_error(PF_ERR_ARG, "null model array");
    for (int a = 0; a < m; ++a)
        if (sel[a] < 0 || sel[a] >= npose) return set_error(PF_ERR_ARG, "sel[%d]=%d out of range", a, sel[a]);
    std::vector<double> Rm((size_t)m * m, 0.0);
    if (R) std::copy(R, R + (size_t)m * m, Rm.begin());
    else for (int i = 0; i < m; ++i) Rm[(size_t)i * m + i] = 100.0;  // my_gmm.cpp:54
    for (int r = 0; r < m; ++r)
        for (int c = 0; c < m; ++c)
            if ((r / 2) != (c / 2) && Rm[(size_t)r * m + c] != 0.0)
                return set_error(PF_ERR_UNSUPPORTED, "R must be block diagonal in 2x2 joints");
    for (int b = 0; b < m / 2; ++b) {
        const double r00 = Rm[(2 * b) * m + 2 * b], r01 = Rm[(2 * b) * m + 2 * b + 1],
                     r10 = Rm[(2 * b + 1) * m + 2 * b], r11 = Rm[(2 * b + 1) * m + 2 * b + 1];
        if (!(r00 > 0.0) || !(r00 * r11 - r01 * r10 > 0.0) || r01 != r10)
            return set_error(PF_ERR_NOT_SPD, "R joint block %d not symmetric positive definite", b);
    }
    double wmax = 0.0;
    for (int k = 0; k < K; ++k) {
        if (!(weights[k] >= 0.0) || !std::isfinite(weights[k]))
            return set_error(PF_ERR_ARG, "weight %d invalid", k);
        wmax = std::max(wmax, weights[k]);
        if (!(gamma[k] >= -1.0 && gamma[k] <= 1.0)) return set_error(PF_ERR_ARG, "gamma %d outside [-1,1]", k);
    }
    if (!(wmax > 0.0)) return set_error(PF_ERR_ARG, "all component weights are zero");
    // Sigma_k must be SPD (the tracks start from them, my_gmm.cpp:36).
    for (int k = 0; k < K; ++k) {
        std::vector<double> L(covs + (size_t)k * d * d, covs + (size_t)(k + 1) * d * d);
        for (int jj = 0; jj < d; ++jj) {
            double s = L[jj * d + jj];
            for (int q = 0; q < jj; ++q) s -= L[jj * d + q] * L[jj * d + q];
            if (!(s > 0.0)) return set_error(PF_ERR_NOT_SPD, "covariance of component %d is not SPD", k);
            const double l = std::sqrt(s);
            L[jj * d + jj] = l;
            for (int i = jj + 1; i < d; ++i) {
                double t = L[i * d + jj];
                for (int q = 0; q < jj; ++q) t -= L[i * d + q] * L[jj * d + q];
                L[i * d + jj] = t / l;
            }
        }
    }
    hm.D = d; hm.K = K; hm.m = m; hm.J = m / 2; hm.npose = npose;
    hm.mu.assign(means, means + (size_t)K * d);
    hm.sig.assign(covs, covs + (size_t)K * d * d);
    hm.pi.assign(weights, weights + K);
    hm.gamma.assign(gamma, gamma + K);
    hm.H.assign((size_t)m * d, 0.0);
    hm.BH.assign(m, 0.0);
    for (int a = 0; a < m; ++a) {                       // H = H1 pca_proj^T, BH = H1 pca_mean^T
        for (int j = 0; j < d; ++j) hm.H[(size_t)a * d + j] = pca_proj[(size_t)j * npose + sel[a]];
        hm.BH[a] = pca_mean[sel[a]];
    }
    hm.R = Rm;
    hm.wmax = wmax;
    return PF_OK;
}

unsigned joint_order(unsigned mask, int J)
{
    unsigned order = 0;
    int slot = 0;
    for (int b = 0; b < J; ++b)
        if (!((mask >> b) & 1u)) order |= (unsigned)b << (4 * slot++);
    for (int b = 0; b < J; ++b)
        if ((mask >> b) & 1u) order |= (unsigned)b << (4 * slot++);
    return order;
}

int make_model_image(const HostModel &hm, unsigned order, int N, ModelImage &img)
{
    const int d = hm.D, K = hm.K, m = hm.m, J = hm.J, NP = npacked(d), CS = 2 + d + NP;
    std::vector<double> T, Ti;
    img.sel_mode = make_basis(hm, order, T, Ti);
    if (!img.sel_mode) {
        T.assign((size_t)d * d, 0.0);
        for (int i = 0; i < d; ++i) T[(size_t)i * d + i] = 1.0;
        Ti = T;
    }
    img.T = T;
    img.Ti = Ti;
    auto tmul = [&](const double *v, double *out) {               // out = T v
        for (int i = 0; i < d; ++i) {
            double acc = 0.0;
            for (int k = 0; k < d; ++k) acc += T[(size_t)i * d + k] * v[k];
            out[i] = acc;
        }
    };
    auto tsandwich = [&](const double *S, double *out) {          // out = T S T^T
        std::vector<double> TS((size_t)d * d, 0.0);
        for (int i = 0; i < d; ++i)
            for (int j = 0; j < d; ++j) {
                double acc = 0.0;
                for (int k = 0; k < d; ++k) acc += T[(size_t)i * d + k] * S[(size_t)k * d + j];
                TS[(size_t)i * d + j] = acc;
            }
        for (int i = 0; i < d; ++i)
            for (int j = 0; j < d; ++j) {
                double acc = 0.0;
                for (int k = 0; k < d; ++k) acc += TS[(size_t)i * d + k] * T[(size_t)j * d + k];
                out[(size_t)i * d + j] = acc;
            }
    };
    const int head = image_head_floats(K, m, d, J);
    std::vector<float> tab((size_t)K * CS);                     // component rows
    std::vector<float> hbr((size_t)head - 4 * K, 0.f);         // H, BH, R blocks
    img.mu.assign((size_t)K * d, 0.f);
    img.sg.assign((size_t)K * NP, 0.f);
    std::vector<double> mut(d), sgt((size_t)d * d), sym((size_t)d * d);
    for (int k = 0; k < K; ++k) {                       // my_gmm.cpp:51-58, in the state basis
        const double g = hm.gamma[k];
        const double *Sk = &hm.sig[(size_t)k * d * d];
        for (int i = 0; i < d; ++i)
            for (int j = 0; j < d; ++j) sym[(size_t)i * d + j] = 0.5 * (Sk[(size_t)i * d + j] + Sk[(size_t)j * d + i]);
        tmul(&hm.mu[(size_t)k * d], mut.data());
        tsandwich(sym.data(), sgt.data());
        float *c = &tab[(size_t)k * CS];
        c[0] = (float)g;
        c[1] = (float)(g * g);
        for (int i = 0; i < d; ++i) {
            c[2 + i] = (float)((1.0 - g) * mut[i]);
            img.mu[(size_t)k * d + i] = (float)mut[i];
        }
        for (int i = 0; i < d; ++i)
            for (int q = i; q < d; ++q) {
                const double v = 0.5 * (sgt[(size_t)i * d + q] + sgt[(size_t)q * d + i]);
                c[2 + d + pk(d, i, q)] = (float)((1.0 - g * g) * v);
                img.sg[(size_t)k * NP + pk(d, i, q)] = (float)v;
            }
    }
    float *Hs = hbr.data();
    for (int i = 0; i < m * d; ++i) Hs[i] = (float)hm.H[i];
    for (int a = 0; a < m; ++a) Hs[m * d + a] = (float)hm.BH[a];
    const std::vector<double> &Rm = hm.R;
    for (int b = 0; b < J; ++b) {
        Hs[m * d + m + 3 * b] = (float)Rm[(2 * b) * m + 2 * b];
        Hs[m * d + m + 3 * b + 1] = (float)Rm[(2 * b) * m + 2 * b + 1];
        Hs[m * d + m + 3 * b + 2] = (float)Rm[(2 * b + 1) * m + 2 * b + 1];
    }
    // FP64 innovation blocks (the joints whose S is factored in FP64)
    std::vector<double> j64((size_t)5 * J);
    for (int b = 0; b < J; ++b) {
        j64[5 * b] = hm.BH[2 * b];
        j64[5 * b + 1] = hm.BH[2 * b + 1];
        j64[5 * b + 2] = Rm[(2 * b) * m + 2 * b];
        j64[5 * b + 3] = Rm[(2 * b) * m + 2 * b + 1];
        j64[5 * b + 4] = Rm[(2 * b + 1) * m + 2 * b + 1];
    }
    // Component thresholds for resample(gmm.weight, N) (pf2DRao.cpp:128).
    std::vector<unsigned long long> thr(2 * (size_t)K);
    unsigned long long acc = 0;
    for (int k = 0; k < K; ++k) {
        acc += fixed32(hm.pi[k] / hm.wmax);
        const u128 th = (u128)acc * ((u128)(unsigned long long)N << 32);
        thr[2 * k] = (unsigned long long)th;
        thr[2 * k + 1] = (unsigned long long)(th >> 64);
    }
    img.pi_total = acc;
    img.blob.assign((size_t)head + tab.size(), 0.f);
    std::memcpy(img.blob.data(), thr.data(), sizeof(unsigned long long) * thr.size());
    std::memcpy(img.blob.data() + 4 * K, hbr.data(), sizeof(float) * hbr.size());
    std::memcpy(img.blob.data() + image_j64_floats(K, m, d, J), j64.data(), sizeof(double) * j64.size());
    std::memcpy(img.blob.data() + head, tab.data(), sizeof(float) * tab.size());
    img.blob.resize((img.blob.size() + 3) & ~(size_t)3, 0.f);
    img.head_floats = head;
    img.comp_in_lds = sizeof(float) * img.blob.size() <= (size_t)kImageLdsBytes;
    if (sizeof(float) * (size_t)head > (size_t)kImageLdsBytes)
        return set_error(PF_ERR_UNSUPPORTED, "model image head of %zu bytes exceeds the %d-byte LDS budget",
                         sizeof(float) * (size_t)head, kImageLdsBytes);
    return PF_OK;
}

}  // namespace mkfbt
