// pf2d.hip -- the plain particle filter pf2D (src/pf2D.h, src/pf2D.cpp) on
// gfx950, C ABI in include/pf2d.h.
//
// State: N x 8 pixel particles, FP32 structure-of-arrays (x[c][ld]), double
// buffered for the resampling gather.  Per update: k_pf2d_weight (GMM prior
// + measurement likelihood, log domain, block maximum), the shared
// fixed-point systematic resampler (k_weights, k_ancestors), then
// k_pf2d_move (gather + random walk).  The normalised weights of the last
// update stay on the device as log-weights plus the log of their sum.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/pf2d.h"
#include "../../include/pf_rbpf.h"
#include "pf_internal.hpp"
#include "pf_kernels.hpp"

using namespace mkfbt;

#define HIP_TRY2(expr)                                                                         \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return set_error(PF_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));       \
    } while (0)

namespace {

constexpr int kD = 8;
constexpr float kImgW = 640.f, kImgH = 480.f;

// Column bounds of the uniform initialisation (pf2D.cpp:62-69).
__host__ __device__ inline void col_bounds(int c, int side, float &lo, float &hi)
{
    if (c == 6) {
        lo = kImgW / 2.f * side + 1.f;
        hi = kImgW / 2.f + kImgW / 2.f * side;
    } else {
        lo = 1.f;
        hi = ((c % 2) == 0) ? kImgW : kImgH;
    }
}

// Per-particle log weight (pf2D.cpp:155-172), log domain:
//   logsumexp_k [log(pi_k c_k) - (x - mu_k)^T Sigma_k^-1 (x - mu_k) / 2]
//   + log N2(x[6:8]; z_0, 15 I) + log N2(x[0:2]; z_1, 15 I).
// GMM table per component: [log(pi_k c_k), mu_k (8), Sigma_k^-1 (64)].
__global__ void __launch_bounds__(256) k_pf2d_weight(const float *xs, int N, int ld, const float *gmm, int K,
                                                     float z00, float z01, float z10, float z11, float *lw)
{
    constexpr int CS = 1 + kD + kD * kD;
    extern __shared__ float sg[];
    for (int i = threadIdx.x; i < K * CS; i += blockDim.x) sg[i] = gmm[i];
    __syncthreads();
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    float out = -INFINITY;
    if (j < N) {
        float x[kD];
#pragma unroll
        for (int c = 0; c < kD; ++c) x[c] = xs[(size_t)c * ld + j];
        float m = -INFINITY, s = 0.f;
        for (int k = 0; k < K; ++k) {
            const float *g = sg + k * CS;
            float dv[kD];
#pragma unroll
            for (int c = 0; c < kD; ++c) dv[c] = x[c] - g[1 + c];
            float q = 0.f;
#pragma unroll
            for (int a = 0; a < kD; ++a) {
                float r = 0.f;
#pragma unroll
                for (int b = 0; b < kD; ++b) r = fmaf(g[1 + kD + a * kD + b], dv[b], r);
                q = fmaf(dv[a], r, q);
            }
            const float t = g[0] - 0.5f * q;
            if (t > m) {                  // online logsumexp
                s = s * __expf(m - t) + 1.f;
                m = t;
            } else {
                s += __expf(t - m);
            }
        }
        const float lprior = m + logf(s);
        const float u0 = x[6] - z00, u1 = x[7] - z01;
        const float v0 = x[0] - z10, v1 = x[1] - z11;
        const float lnorm = -logf(6.283185307179586f * 15.f);
        out = lprior + 2.f * lnorm - (fmaf(u0, u0, u1 * u1) + fmaf(v0, v0, v1 * v1)) / 30.f;
        lw[j] = out;
    }
}

// Resampling gather (anc; the identity when null) or re-initialisation, then
// predict (pf2D.cpp:90-102): x += 5 n, normals of particle i consumed for
// columns 2, 3, 4, 5, 6, 7, 0, 1.
__global__ void __launch_bounds__(256) k_pf2d_move(const float *xin, float *xout, const int *anc,
                                                   const double *normals, const double *reinit, int side,
                                                   int N, int ld)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    float base[kD];
    if (reinit) {
#pragma unroll
        for (int c = 0; c < kD; ++c) {
            float lo, hi;
            col_bounds(c, side, lo, hi);
            base[c] = (float)(lo + (double)(hi - lo) * reinit[(size_t)j * kD + c]);
        }
    } else {
        const int s = anc ? anc[j] : j;
#pragma unroll
        for (int c = 0; c < kD; ++c) base[c] = xin[(size_t)c * ld + s];
    }
    const int col_of_draw[kD] = {2, 3, 4, 5, 6, 7, 0, 1};
#pragma unroll
    for (int t = 0; t < kD; ++t) {
        const int c = col_of_draw[t];
        base[c] = base[c] + (float)(5.0 * normals[(size_t)j * kD + t]);
    }
#pragma unroll
    for (int c = 0; c < kD; ++c) xout[(size_t)c * ld + j] = base[c];
}

// Constructor randomisation (pf2D.cpp:58-70).
__global__ void k_pf2d_init(float *xs, const double *unif, int side, int N, int ld)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
#pragma unroll
    for (int c = 0; c < kD; ++c) {
        float lo, hi;
        col_bounds(c, side, lo, hi);
        xs[(size_t)c * ld + j] = (float)(lo + (double)(hi - lo) * unif[(size_t)j * kD + c]);
    }
}

// Normalised weight of particle j from the kept log-weights.
__device__ __forceinline__ double pf2d_weight_of(const float *lw, const double *lsum, int uniform, int N, int j)
{
    return uniform ? 1.0 / (double)N : exp((double)lw[j] - *lsum);
}

// getEstimator (pf2D.cpp:79-87): sum_i w_i x_i, FP64 accumulation in a fixed
// order: block sums (block_sum256), then k_pf2d_estimate_final adds them.
__global__ void __launch_bounds__(256) k_pf2d_estimate(const float *xs, const float *lw, const double *lsum,
                                                       int uniform, int N, int ld, double *partial)
{
    __shared__ double s_v[kD][16][17];
    __shared__ double s_run[kD][16];
    __shared__ double s_sum[kD];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const double w = (j < N) ? pf2d_weight_of(lw, lsum, uniform, N, j) : 0.0;
    double v[kD];
#pragma unroll
    for (int c = 0; c < kD; ++c) v[c] = (j < N) ? w * (double)xs[(size_t)c * ld + j] : 0.0;
    block_sum256<kD>(v, s_v, s_run, s_sum);
    if (threadIdx.x < kD) partial[(size_t)blockIdx.x * kD + threadIdx.x] = s_sum[threadIdx.x];
}

__global__ void __launch_bounds__(256) k_pf2d_estimate_final(const double *partial, int nb, double *out)
{
    __shared__ double s_v[kD][16][17];
    __shared__ double s_run[kD][16];
    __shared__ double s_sum[kD];
    double v[kD];
#pragma unroll
    for (int c = 0; c < kD; ++c) v[c] = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
#pragma unroll
        for (int c = 0; c < kD; ++c) v[c] += partial[(size_t)b * kD + c];
    }
    block_sum256<kD>(v, s_v, s_run, s_sum);
    if (threadIdx.x < kD) out[threadIdx.x] = s_sum[threadIdx.x];
}

// Particles (row-major double) and weights out.
__global__ void k_pf2d_export(const float *xs, const float *lw, const double *lsum, int uniform, int N, int ld,
                              double *p, double *w)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
#pragma unroll
    for (int c = 0; c < kD; ++c) p[(size_t)j * kD + c] = (double)xs[(size_t)c * ld + j];
    if (w) w[j] = pf2d_weight_of(lw, lsum, uniform, N, j);
}

// Particles and weights in (weights become log-weights with lsum = 0).
__global__ void k_pf2d_import(const double *p, const double *w, int N, int ld, float *xs, float *lw, double *lsum)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j == 0) *lsum = 0.0;
    if (j >= N) return;
#pragma unroll
    for (int c = 0; c < kD; ++c) xs[(size_t)c * ld + j] = (float)p[(size_t)j * kD + c];
    if (w) lw[j] = (float)log(w[j]);
}

// log of the sum of the kept weights: lml + log N (lml = log mean weight).
__global__ void k_pf2d_lsum(double *lsum, const double *lml, int N)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *lsum = *lml + log((double)N);
}

template <typename T>
hipError_t dalloc2(T **p, size_t n)
{
    return hipMalloc((void **)p, sizeof(T) * std::max<size_t>(n, 1));
}

}  // namespace

struct pf2d_handle {
    int device = 0, N = 0, ld = 0, side = 0, K = 0;
    bool gmm = false, uniform = true;
    hipStream_t stream = nullptr;
    float *xA = nullptr, *xB = nullptr, *lw = nullptr, *table = nullptr;
    int *anc = nullptr, *frame = nullptr;
    uint32_t *uni = nullptr;
    FrameScratch *fs = nullptr;
    unsigned long long *cloc = nullptr, *bsum = nullptr;
    double *bexp = nullptr, *lml = nullptr, *lsum = nullptr, *partial = nullptr, *out = nullptr;
    double *nbuf = nullptr, *rbuf = nullptr, *pbuf = nullptr, *wbuf = nullptr;
    int *bpow = nullptr;
};

extern "C" {

int pf2d_destroy(pf2d_handle *h)
{
    if (!h) return PF_OK;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    void *ptrs[] = {h->xA, h->xB, h->lw, h->table, h->anc, h->frame, h->uni, h->fs, h->cloc, h->bsum,
                    h->bexp, h->bpow, h->lml, h->lsum, h->partial, h->out, h->nbuf, h->rbuf, h->pbuf, h->wbuf};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return PF_OK;
}

int pf2d_create(pf2d_handle **out, int32_t n_particles, int32_t n_dims, int32_t side, int32_t device)
{
    if (!out) return set_error(PF_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (n_dims != kD) return set_error(PF_ERR_UNSUPPORTED, "pf2D draws noise for 8 columns; numDims must be 8");
    if (n_particles <= 0 || n_particles > kMaxScanBlocks * kWeightsBlock)
        return set_error(PF_ERR_ARG, "n_particles %d out of range", n_particles);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return set_error(PF_ERR_HIP, "no HIP device available");
    if (device < 0 || device >= ndev) return set_error(PF_ERR_ARG, "device %d out of range", device);
    HIP_TRY2(hipSetDevice(device));
    pf2d_handle *h = new pf2d_handle();
    h->device = device;
    h->N = n_particles;
    h->ld = (n_particles + 63) & ~63;
    h->side = side ? 1 : 0;
    const int nb = (h->N + kWeightsBlock - 1) / kWeightsBlock;
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    e = e ? e : dalloc2(&h->xA, (size_t)kD * h->ld);
    e = e ? e : dalloc2(&h->xB, (size_t)kD * h->ld);
    e = e ? e : dalloc2(&h->lw, (size_t)h->ld);
    e = e ? e : dalloc2(&h->anc, (size_t)h->ld);
    e = e ? e : dalloc2(&h->frame, 2);
    e = e ? e : dalloc2(&h->uni, 2);
    e = e ? e : dalloc2(&h->fs, 1);
    e = e ? e : dalloc2(&h->cloc, (size_t)h->N);
    e = e ? e : dalloc2(&h->bsum, (size_t)nb);
    e = e ? e : dalloc2(&h->bexp, (size_t)nb);
    e = e ? e : dalloc2(&h->bpow, (size_t)nb);
    e = e ? e : dalloc2(&h->lml, 1);
    e = e ? e : dalloc2(&h->lsum, 1);
    e = e ? e : dalloc2(&h->partial, (size_t)((h->N + 255) / 256) * kD);
    e = e ? e : dalloc2(&h->out, kD);
    e = e ? e : dalloc2(&h->nbuf, (size_t)h->N * kD);
    e = e ? e : dalloc2(&h->rbuf, (size_t)h->N * kD);
    e = e ? e : dalloc2(&h->pbuf, (size_t)h->N * kD);
    e = e ? e : dalloc2(&h->wbuf, (size_t)h->N);
    if (e == hipSuccess) e = hipMemsetAsync(h->fs, 0, sizeof(FrameScratch), h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
        pf2d_destroy(h);
        return set_error(PF_ERR_HIP, "pf2d allocation failed: %s", hipGetErrorString(e));
    }
    *out = h;
    return PF_OK;
}

int pf2d_load_gmm_precision(pf2d_handle *h, int32_t K, const double *means, const double *sigma_i,
                            const double *det_s, const double *weights)
{
    if (!h || !means || !sigma_i || !det_s || !weights) return set_error(PF_ERR_ARG, "null argument");
    if (K <= 0 || K > 256) return set_error(PF_ERR_ARG, "K=%d out of range", K);
    constexpr int CS = 1 + kD + kD * kD;
    std::vector<float> tab((size_t)K * CS);
    for (int k = 0; k < K; ++k) {
        // weight_k det_s_k exp(-(x - mu_k) sigma_i_k (x - mu_k)^T / 2) (pf2D.cpp:162)
        const double c = weights[k] * det_s[k];
        if (!(c > 0.0) || !std::isfinite(c))
            return set_error(PF_ERR_ARG, "component %d: weight * det_s must be finite and > 0", k);
        const double *Si = sigma_i + (size_t)k * kD * kD;
        float *g = &tab[(size_t)k * CS];
        g[0] = (float)std::log(c);
        for (int i = 0; i < kD; ++i) g[1 + i] = (float)means[(size_t)k * kD + i];
        for (int i = 0; i < kD * kD; ++i) g[1 + kD + i] = (float)(0.5 * (Si[i] + Si[(i % kD) * kD + i / kD]));
    }
    HIP_TRY2(hipSetDevice(h->device));
    (void)hipFree(h->table);
    h->table = nullptr;
    HIP_TRY2(dalloc2(&h->table, tab.size()));
    HIP_TRY2(hipMemcpyAsync(h->table, tab.data(), sizeof(float) * tab.size(), hipMemcpyHostToDevice, h->stream));
    HIP_TRY2(hipStreamSynchronize(h->stream));
    h->K = K;
    h->gmm = true;
    return PF_OK;
}

int pf2d_load_gmm(pf2d_handle *h, int32_t K, const double *means, const double *covs, const double *weights)
{
    if (!h || !means || !covs || !weights) return set_error(PF_ERR_ARG, "null argument");
    if (K <= 0 || K > 256) return set_error(PF_ERR_ARG, "K=%d out of range", K);
    std::vector<double> Sinv((size_t)K * kD * kD), dets((size_t)K);
    for (int k = 0; k < K; ++k) {
        // Cholesky of Sigma_k, its inverse and determinant (my_gmm::loadGaussian, pf2D.cpp:33-35)
        double L[kD * kD];
        const double *S = covs + (size_t)k * kD * kD;
        for (int i = 0; i < kD * kD; ++i) L[i] = 0.0;
        for (int j = 0; j < kD; ++j) {
            double s = S[j * kD + j];
            for (int q = 0; q < j; ++q) s -= L[j * kD + q] * L[j * kD + q];
            if (!(s > 0.0)) return set_error(PF_ERR_NOT_SPD, "covariance of component %d is not SPD", k);
            L[j * kD + j] = std::sqrt(s);
            for (int i = j + 1; i < kD; ++i) {
                double t = 0.5 * (S[i * kD + j] + S[j * kD + i]);
                for (int q = 0; q < j; ++q) t -= L[i * kD + q] * L[j * kD + q];
                L[i * kD + j] = t / L[j * kD + j];
            }
        }
        double logdet = 0.0;
        for (int i = 0; i < kD; ++i) logdet += 2.0 * std::log(L[i * kD + i]);
        double *Si = &Sinv[(size_t)k * kD * kD];
        for (int c = 0; c < kD; ++c) {
            double y[kD], xx[kD];
            for (int i = 0; i < kD; ++i) {
                double t = (i == c) ? 1.0 : 0.0;
                for (int q = 0; q < i; ++q) t -= L[i * kD + q] * y[q];
                y[i] = t / L[i * kD + i];
            }
            for (int i = kD - 1; i >= 0; --i) {
                double t = y[i];
                for (int q = i + 1; q < kD; ++q) t -= L[q * kD + i] * xx[q];
                xx[i] = t / L[i * kD + i];
            }
            for (int i = 0; i < kD; ++i) Si[i * kD + c] = xx[i];
        }
        if (!(weights[k] > 0.0)) return set_error(PF_ERR_ARG, "component weight %d must be > 0", k);
        dets[k] = std::exp(-0.5 * kD * std::log(6.283185307179586) - 0.5 * logdet);
    }
    return pf2d_load_gmm_precision(h, K, means, Sinv.data(), dets.data(), weights);
}

int pf2d_init(pf2d_handle *h, const double *unif)
{
    if (!h || !unif) return set_error(PF_ERR_ARG, "null argument");
    HIP_TRY2(hipSetDevice(h->device));
    HIP_TRY2(hipMemcpyAsync(h->rbuf, unif, sizeof(double) * (size_t)h->N * kD, hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_pf2d_init, dim3((h->N + 255) / 256), dim3(256), 0, h->stream, h->xA, h->rbuf, h->side,
                       h->N, h->ld);
    HIP_TRY2(hipGetLastError());
    HIP_TRY2(hipStreamSynchronize(h->stream));
    h->uniform = true;
    return PF_OK;
}

int pf2d_set_particles(pf2d_handle *h, const double *particles, const double *weights)
{
    if (!h || !particles) return set_error(PF_ERR_ARG, "null argument");
    HIP_TRY2(hipSetDevice(h->device));
    HIP_TRY2(hipMemcpyAsync(h->pbuf, particles, sizeof(double) * (size_t)h->N * kD, hipMemcpyHostToDevice, h->stream));
    if (weights)
        HIP_TRY2(hipMemcpyAsync(h->wbuf, weights, sizeof(double) * (size_t)h->N, hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_pf2d_import, dim3((h->N + 255) / 256), dim3(256), 0, h->stream, h->pbuf,
                       weights ? h->wbuf : nullptr, h->N, h->ld, h->xA, h->lw, h->lsum);
    HIP_TRY2(hipGetLastError());
    HIP_TRY2(hipStreamSynchronize(h->stream));
    h->uniform = (weights == nullptr);
    return PF_OK;
}

int pf2d_get_particles(pf2d_handle *h, double *particles, double *weights)
{
    if (!h) return set_error(PF_ERR_ARG, "null handle");
    HIP_TRY2(hipSetDevice(h->device));
    hipLaunchKernelGGL(k_pf2d_export, dim3((h->N + 255) / 256), dim3(256), 0, h->stream, h->xA, h->lw, h->lsum,
                       (int)h->uniform, h->N, h->ld, h->pbuf, h->wbuf);
    HIP_TRY2(hipGetLastError());
    if (particles)
        HIP_TRY2(hipMemcpyAsync(particles, h->pbuf, sizeof(double) * (size_t)h->N * kD, hipMemcpyDeviceToHost, h->stream));
    if (weights)
        HIP_TRY2(hipMemcpyAsync(weights, h->wbuf, sizeof(double) * (size_t)h->N, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY2(hipStreamSynchronize(h->stream));
    return PF_OK;
}

int pf2d_update(pf2d_handle *h, const double *meas, uint32_t u, const double *normals, const double *reinit_unif,
                float *lw, int32_t *anc)
{
    if (!h || !meas || !normals || !reinit_unif) return set_error(PF_ERR_ARG, "null argument");
    if (!h->gmm) return set_error(PF_ERR_STATE, "load the GMM first");
    HIP_TRY2(hipSetDevice(h->device));
    const int N = h->N, nb = (N + kWeightsBlock - 1) / kWeightsBlock;
    const uint32_t uu[2] = {u, u};
    HIP_TRY2(hipMemcpyAsync(h->uni, uu, sizeof(uu), hipMemcpyHostToDevice, h->stream));
    HIP_TRY2(hipMemcpyAsync(h->nbuf, normals, sizeof(double) * (size_t)N * kD, hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, h->stream, h->frame, 0);
    const size_t lds = sizeof(float) * (size_t)h->K * (1 + kD + kD * kD);
    hipLaunchKernelGGL(k_pf2d_weight, dim3((N + 255) / 256), dim3(256), lds, h->stream, h->xA, N, h->ld, h->table,
                       h->K, (float)meas[0], (float)meas[1], (float)meas[2], (float)meas[3], h->lw);
    HIP_TRY2(hipGetLastError());
    WeightArgs w{};
    w.lw32 = h->lw;
    w.cloc = h->cloc;
    w.bsum = h->bsum;
    w.bexp = h->bexp;
    w.bpow = h->bpow;
    w.fs = h->fs;
    w.frame = h->frame;
    w.frame_next = h->frame + 1;
    w.uni = h->uni;
    w.L = N;
    w.parity = 0;
    w.ustride = 2;
    w.uoff = 1;
    hipLaunchKernelGGL((k_weights<kLogF32>), dim3(nb), dim3(kWeightsBlock), 0, h->stream, w);
    HIP_TRY2(hipGetLastError());
    AncArgs an{};
    an.pend.cloc = h->cloc;
    an.pend.bsum = h->bsum;
    an.pend.bexp = h->bexp;
    an.pend.bpow = h->bpow;
    an.pend.L = N;
    an.pend.nb = nb;
    an.fs = h->fs;
    an.anc = h->anc;
    an.lml = h->lml;
    an.N = N;
    an.parity = 0;
    an.kind_linear = 0;
    hipLaunchKernelGGL(k_ancestors, dim3((N + 255) / 256), dim3(256), 8 * (size_t)resolve_lds_words(nb), h->stream, an);
    HIP_TRY2(hipGetLastError());
    hipLaunchKernelGGL(k_pf2d_lsum, dim3(1), dim3(64), 0, h->stream, h->lsum, h->lml, N);
    FrameScratch fs;
    HIP_TRY2(hipMemcpyAsync(&fs, h->fs, sizeof(fs), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY2(hipStreamSynchronize(h->stream));
    const bool degenerate = (fs.status & 1u) != 0;
    if (degenerate) {
        const unsigned int zero = 0;
        HIP_TRY2(hipMemcpyAsync(&h->fs->status, &zero, sizeof(zero), hipMemcpyHostToDevice, h->stream));
        HIP_TRY2(hipMemcpyAsync(h->rbuf, reinit_unif, sizeof(double) * (size_t)N * kD, hipMemcpyHostToDevice, h->stream));
    }
    hipLaunchKernelGGL(k_pf2d_move, dim3((N + 255) / 256), dim3(256), 0, h->stream, h->xA, h->xB, h->anc,
                       h->nbuf, degenerate ? h->rbuf : nullptr, h->side, N, h->ld);
    HIP_TRY2(hipGetLastError());
    std::swap(h->xA, h->xB);
    h->uniform = degenerate;
    if (lw) HIP_TRY2(hipMemcpyAsync(lw, h->lw, sizeof(float) * N, hipMemcpyDeviceToHost, h->stream));
    if (anc) HIP_TRY2(hipMemcpyAsync(anc, h->anc, sizeof(int) * N, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY2(hipStreamSynchronize(h->stream));
    return degenerate ? PF_WARN_DEGENERATE : PF_OK;
}

int pf2d_predict(pf2d_handle *h, const double *normals)
{
    if (!h || !normals) return set_error(PF_ERR_ARG, "null argument");
    HIP_TRY2(hipSetDevice(h->device));
    const int N = h->N;
    HIP_TRY2(hipMemcpyAsync(h->nbuf, normals, sizeof(double) * (size_t)N * kD, hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_pf2d_move, dim3((N + 255) / 256), dim3(256), 0, h->stream, h->xA, h->xB,
                       (const int *)nullptr, h->nbuf, (const double *)nullptr, h->side, N, h->ld);
    HIP_TRY2(hipGetLastError());
    HIP_TRY2(hipStreamSynchronize(h->stream));
    std::swap(h->xA, h->xB);
    return PF_OK;
}

int pf2d_estimate(pf2d_handle *h, double *out)
{
    if (!h || !out) return set_error(PF_ERR_ARG, "null argument");
    HIP_TRY2(hipSetDevice(h->device));
    const int nb = (h->N + 255) / 256;
    hipLaunchKernelGGL(k_pf2d_estimate, dim3(nb), dim3(256), 0, h->stream, h->xA, h->lw, h->lsum, (int)h->uniform,
                       h->N, h->ld, h->partial);
    hipLaunchKernelGGL(k_pf2d_estimate_final, dim3(1), dim3(256), 0, h->stream, (const double *)h->partial, nb, h->out);
    HIP_TRY2(hipGetLastError());
    HIP_TRY2(hipMemcpyAsync(out, h->out, sizeof(double) * kD, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY2(hipStreamSynchronize(h->stream));
    return PF_OK;
}

}  // extern "C"
