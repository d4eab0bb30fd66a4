// pf_kernels.hpp -- gfx950 kernels of the Rao-Blackwellised particle filter
// (reference: src/pf2DRao.cpp, src/KF_model.cpp, src/my_gmm.cpp).
//
// Device layout (one filter):
//   state   x[D][ld], P[D(D+1)/2][ld]  FP32 structure-of-arrays, packed upper
//           triangle of the symmetric covariance, ld = particle capacity.
//           Two copies: the step reads copy A at the ancestor index and writes
//           copy B at the particle index (the resampling gather of
//           pf2DRao.cpp:153-156 is folded into the next frame's loads).
//   table   per component k: [g, g^2, (1-g) mu_k (D), (1-g^2) Sigma_k packed],
//           then the shared H (m x D), BH (m) and the 2x2 joint blocks of R
//           (r00, r01, r11 per joint).  Staged into LDS by every block.
//   thr     component thresholds C_pi[k] * N * 2^32 (128-bit, lo/hi words).
#pragma once

#include "pf_common.hpp"

namespace mkfbt {

constexpr int kMaxJoints = 8;
constexpr int kWeightsBlock = 256;

enum StepMode { kModeRef = 0, kModeCand = 1 };

struct FrameScratch {
    unsigned long long maxkey;   // ordered key of the largest finite weight (0 = none)
    unsigned int done_est;       // completion counter of k_estimate blocks
    unsigned int status;         // bit 0: a degenerate (all-zero) resampling happened
};

struct StepArgs {
    const float *__restrict__ xin;
    const float *__restrict__ Pin;
    float *__restrict__ xout;
    float *__restrict__ Pout;
    const int *__restrict__ anc;          // nullptr: identity
    float *__restrict__ lw;
    int *__restrict__ comp_out;           // optional
    FrameScratch *fs;
    const float *__restrict__ table;
    const int *__restrict__ frame;
    const uint32_t *__restrict__ uni;     // 2 per frame: U_comp, U_res
    // reference mode: Z[t][r][zld]
    const float *__restrict__ Z;
    long long z_frame_stride;
    // candidate mode: count[t][J], cand[t][J][cap][2]
    const int *__restrict__ count;
    const float *__restrict__ cand;
    unsigned long long pi_total;
    int table_floats;   // LDS image: 2K threshold words, then the float table (multiple of 4)
    int K, N, ld, m, J, cap, zld;
    unsigned int order_packed;   // joint processing order, 4 bits per slot
    unsigned int clutter_mask;
    float pdl;    // P_D / lambda
    float b0;     // 1 - P_D * P_G
    float gate;   // chi-square gate
};

__device__ __forceinline__ unsigned long long order_key(double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__device__ __forceinline__ double key_value(unsigned long long k)
{
    const unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double((long long)b);
}

template <typename T>
__device__ __forceinline__ T wave_max(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// --------------------------------------------------------------- step kernel
// One lane = one particle.  Fuses, for particle j of frame t:
//   ancestor gather (pf2DRao.cpp:155), component draw resample(gmm.weight, N)
//   (:128, systematic, thresholds precomputed), KF_model::predict with
//   F = g I (KF_model.cpp:11-15, my_gmm.cpp:53-58), the innovation likelihood
//   (:138) and KF_model::update (KF_model.cpp:17-25), processed as sequential
//   2-row joint blocks (exact for block-diagonal R), then the log-weight write
//   and the block maximum for the resampler.
template <int D>
struct Joint {
    static constexpr int NP = npacked(D);

    // W = L^-1 H_b P (2 x D), y = H_b x + BH_b, S = L L^T.
    __device__ static __forceinline__ void prepare(const float (&x)[D], const float (&P)[NP],
                                                   const float *h0, const float *h1, float bh0,
                                                   float bh1, float r00, float r01, float r11,
                                                   float (&w0)[D], float (&w1)[D], float &y0,
                                                   float &y1, float &il00, float &l10, float &il11,
                                                   float &logl)
    {
        float hp0[D], hp1[D];
#pragma unroll
        for (int c = 0; c < D; ++c) {
            float s0 = 0.f, s1 = 0.f;
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const float p = P[pk(D, k, c)];
                s0 = fmaf(h0[k], p, s0);
                s1 = fmaf(h1[k], p, s1);
            }
            hp0[c] = s0;
            hp1[c] = s1;
        }
        float a0 = bh0, a1 = bh1, s00 = r00, s01 = r01, s11 = r11;
#pragma unroll
        for (int c = 0; c < D; ++c) {
            a0 = fmaf(h0[c], x[c], a0);
            a1 = fmaf(h1[c], x[c], a1);
            s00 = fmaf(h0[c], hp0[c], s00);
            s01 = fmaf(h1[c], hp0[c], s01);
            s11 = fmaf(h1[c], hp1[c], s11);
        }
        y0 = a0;
        y1 = a1;
        const float l00 = sqrtf(s00);
        il00 = 1.f / l00;
        l10 = s01 * il00;
        const float l11 = sqrtf(s11 - l10 * l10);
        il11 = 1.f / l11;
        logl = logf(l00 * l11);
#pragma unroll
        for (int c = 0; c < D; ++c) {
            w0[c] = hp0[c] * il00;
            w1[c] = (hp1[c] - l10 * w0[c]) * il11;
        }
    }

    // x += W^T v ; P -= W^T M W with M = [[m00, m01], [m01, m11]].
    __device__ static __forceinline__ void apply(float (&x)[D], float (&P)[NP], const float (&w0)[D],
                                                 const float (&w1)[D], float v0, float v1, float m00,
                                                 float m01, float m11)
    {
        float u0[D], u1[D];
#pragma unroll
        for (int c = 0; c < D; ++c) {
            x[c] = fmaf(w0[c], v0, fmaf(w1[c], v1, x[c]));
            u0[c] = fmaf(m00, w0[c], m01 * w1[c]);
            u1[c] = fmaf(m01, w0[c], m11 * w1[c]);
        }
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
            for (int k = i; k < D; ++k) P[pk(D, i, k)] -= fmaf(w0[i], u0[k], w1[i] * u1[k]);
    }
};

// Copies n4 float4 of the model table into LDS: all loads of a chunk are
// issued before the first LDS write so the copy costs one memory latency.
__device__ __forceinline__ void stage_table(float4 *dst, const float4 *__restrict__ src, int n4)
{
    constexpr int kChunk = 8;
    for (int base = 0; base < n4; base += kChunk * (int)blockDim.x) {
        float4 r[kChunk];
#pragma unroll
        for (int it = 0; it < kChunk; ++it) {
            const int i = base + it * (int)blockDim.x + (int)threadIdx.x;
            if (i < n4) r[it] = src[i];
        }
#pragma unroll
        for (int it = 0; it < kChunk; ++it) {
            const int i = base + it * (int)blockDim.x + (int)threadIdx.x;
            if (i < n4) dst[i] = r[it];
        }
    }
    __syncthreads();
}

// Structure-of-arrays rows through buffer instructions: one resource per
// array, the row offset in an SGPR (soffset) and the particle offset in a
// 32-bit VGPR (voffset), so no per-load 64-bit address arithmetic.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t soa_rsrc(const float *p, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(p), (short)0, (int)bytes, 0x00020000);
}

__device__ __forceinline__ float soa_ld(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    return __int_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}

__device__ __forceinline__ void soa_st(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_int(v), r, (int)voff, (int)soff, 0);
}

__device__ __forceinline__ int draw_component(const unsigned long long *thr, int K, uint32_t j,
                                              uint32_t U, unsigned long long total)
{
    const u128 t = sys_target(j, U, total);
    int k = 0;
    while (k < K - 1) {
        const u128 th = ((u128)thr[2 * k + 1] << 64) | (u128)thr[2 * k];
        if (th > t) break;
        ++k;
    }
    return k;
}

template <int D, int MODE>
__global__ void __launch_bounds__(256) k_step(StepArgs a)
{
    constexpr int NP = npacked(D);
    constexpr int CS = 2 + D + NP;
    extern __shared__ float4 sm4[];
    stage_table(sm4, reinterpret_cast<const float4 *>(a.table), a.table_floats >> 2);
    const unsigned long long *sthr = reinterpret_cast<const unsigned long long *>(sm4);
    const float *stab = reinterpret_cast<const float *>(sm4) + 4 * a.K;

    const int t = *a.frame;
    const uint32_t U1 = a.uni[2 * t];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    float lwj = -INFINITY;
    if (j < a.N) {
        const int src = a.anc ? a.anc[j] : j;
        const unsigned rowb = (unsigned)a.ld * 4u;
        const unsigned voff = (unsigned)src * 4u;
        const __amdgpu_buffer_rsrc_t rx = soa_rsrc(a.xin, rowb * D), rP = soa_rsrc(a.Pin, rowb * NP);
        float x[D], P[NP];
#pragma unroll
        for (int i = 0; i < D; ++i) x[i] = soa_ld(rx, voff, rowb * i);
#pragma unroll
        for (int p = 0; p < NP; ++p) P[p] = soa_ld(rP, voff, rowb * p);

        const int k = draw_component(sthr, a.K, (uint32_t)j, U1, a.pi_total);
        if (a.comp_out) a.comp_out[j] = k;
        const float *ck = stab + k * CS;
        const float g = ck[0], g2 = ck[1];
#pragma unroll
        for (int i = 0; i < D; ++i) x[i] = fmaf(g, x[i], ck[2 + i]);
#pragma unroll
        for (int p = 0; p < NP; ++p) P[p] = fmaf(g2, P[p], ck[2 + D + p]);

        const float *Hs = stab + a.K * CS;
        const float *BHs = Hs + a.m * D;
        const float *Rs = BHs + a.m;
        float acc = 0.f;
        const float LOG2PI = 1.8378770664093453f;
        for (int oi = 0; oi < a.J; ++oi) {
            const int b = (MODE == kModeRef) ? oi : (int)((a.order_packed >> (4 * oi)) & 15u);
            const bool clutter = (MODE == kModeCand) && ((a.clutter_mask >> b) & 1u);
            int cnt = 1;
            const float *cz = nullptr;
            if (MODE == kModeCand) {
                cnt = a.count[(size_t)t * a.J + b];
                cz = a.cand + ((size_t)t * a.J + b) * a.cap * 2;
                if (cnt <= 0 && !clutter) continue;
                if (cnt > a.cap) cnt = a.cap;
            }
            float w0[D], w1[D], y0, y1, il00, l10, il11, logl;
            Joint<D>::prepare(x, P, Hs + (2 * b) * D, Hs + (2 * b + 1) * D, BHs[2 * b], BHs[2 * b + 1],
                              Rs[3 * b], Rs[3 * b + 1], Rs[3 * b + 2], w0, w1, y0, y1, il00, l10,
                              il11, logl);
            if (!clutter) {
                float z0, z1;
                if (MODE == kModeRef) {
                    const float *Zt = a.Z + (size_t)t * a.z_frame_stride;
                    z0 = Zt[(size_t)(2 * b) * a.zld + j];
                    z1 = Zt[(size_t)(2 * b + 1) * a.zld + j];
                } else {
                    z0 = cz[0];
                    z1 = cz[1];
                }
                const float v0 = (z0 - y0) * il00;
                const float v1 = ((z1 - y1) - l10 * (z0 - y0) * il00) * il11;
                acc += -0.5f * fmaf(v0, v0, v1 * v1) - logl - LOG2PI;
                Joint<D>::apply(x, P, w0, w1, v0, v1, 1.f, 0.f, 1.f);
            } else {
                // PDAF over the gated detections (Bar-Shalom & Fortmann), in
                // whitened coordinates v_c = L^-1 (z_c - y).
                const float lnorm = -logl - LOG2PI;
                float se = 0.f, sv0 = 0.f, sv1 = 0.f, q00 = 0.f, q01 = 0.f, q11 = 0.f;
                for (int c = 0; c < cnt; ++c) {
                    const float n0 = cz[2 * c] - y0, n1 = cz[2 * c + 1] - y1;
                    const float v0 = n0 * il00;
                    const float v1 = (n1 - l10 * v0) * il11;
                    const float d2 = fmaf(v0, v0, v1 * v1);
                    if (d2 <= a.gate) {
                        const float e = a.pdl * __expf(lnorm - 0.5f * d2);
                        se += e;
                        sv0 = fmaf(e, v0, sv0);
                        sv1 = fmaf(e, v1, sv1);
                        q00 = fmaf(e * v0, v0, q00);
                        q01 = fmaf(e * v0, v1, q01);
                        q11 = fmaf(e * v1, v1, q11);
                    }
                }
                const float ell = a.b0 + se;
                const float inv = 1.f / ell;
                const float v0 = sv0 * inv, v1 = sv1 * inv;
                const float bsum = se * inv;  // 1 - beta_0
                const float m00 = bsum - (q00 * inv - v0 * v0);
                const float m01 = -(q01 * inv - v0 * v1);
                const float m11 = bsum - (q11 * inv - v1 * v1);
                acc += logf(ell);
                Joint<D>::apply(x, P, w0, w1, v0, v1, m00, m01, m11);
            }
        }
        lwj = acc;
        const unsigned doff = (unsigned)j * 4u;
        const __amdgpu_buffer_rsrc_t wx = soa_rsrc(a.xout, rowb * D), wP = soa_rsrc(a.Pout, rowb * NP);
#pragma unroll
        for (int i = 0; i < D; ++i) soa_st(wx, doff, rowb * i, x[i]);
#pragma unroll
        for (int p = 0; p < NP; ++p) soa_st(wP, doff, rowb * p, P[p]);
        a.lw[j] = lwj;
    }
    // block maximum of the finite log-weights -> one atomic per block
    double v = isfinite(lwj) ? (double)lwj : -INFINITY;
    v = wave_max(v);
    __shared__ double wmax[4];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) wmax[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = wmax[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = fmax(m, wmax[w]);
        if (isfinite(m)) atomicMax(&a.fs->maxkey, order_key(m));
    }
}

// ------------------------------------------------------------- resampling
// Systematic resampling (pf2DRao.cpp:175-210) with fixed-point weights:
//   q_i = floor(2^32 * w_i / max w) (or det_exp(lw_i - max lw)),
//   C = inclusive prefix sum (exact, order independent),
//   a_i = min{idx : C[idx] * N * 2^32 > (i * 2^32 + U) * C[L-1]}.
// Equivalently input idx owns the outputs [g(C[idx-1]), g(C[idx])) with
//   g(c) = min{i in [0, N] : (i * 2^32 + U) * T >= c * N * 2^32}.
// k_weights: per-block inclusive scans of q (no cross-block traffic).
// k_ancestors: block prefix from the block sums, g() of every input, then the
// block writes its contiguous output range (binary search in LDS).
enum WeightKind { kLogF32 = 0, kLogF64 = 1, kLinF64 = 2 };

struct WeightArgs {
    const float *lw32;
    const double *w64;
    unsigned long long *cloc;     // L block-local inclusive prefix sums
    unsigned long long *bsum;     // nb block sums
    double *bexp;                 // nb block sums of the FP64 weights
    FrameScratch *fs;
    const int *frame;             // frame counter read this frame
    int *frame_next;              // frame counter of the next frame (block 0 writes t + 1)
    int L;
};

template <int KIND>
__global__ void __launch_bounds__(kWeightsBlock) k_weights(WeightArgs a)
{
    __shared__ unsigned long long s_wave[4];
    __shared__ double s_dwave[4];
    const unsigned long long key = a.fs->maxkey;
    const bool degenerate = (key == 0ull);
    const double ref = degenerate ? 0.0 : key_value(key);
    const int i = blockIdx.x * kWeightsBlock + threadIdx.x;
    unsigned long long q = 0;
    double v = 0.0;
    if (i < a.L) {
        if (degenerate) {
            q = 1;
        } else if (KIND == kLogF32 || KIND == kLogF64) {
            const double x = (KIND == kLogF32) ? (double)a.lw32[i] : a.w64[i];
            v = isfinite(x) ? det_exp(x - ref) : 0.0;
            q = qlog(x, ref);
        } else {
            const double x = a.w64[i];
            const bool ok = isfinite(x) && x > 0.0;
            v = ok ? x : 0.0;
            q = ok ? fixed32(x / ref) : 0;
        }
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    unsigned long long s = q;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned long long n = __shfl_up(s, o, 64);
        if (lane >= o) s += n;
    }
    const double vs = wave_sum(v);
    if (lane == 63) s_wave[wid] = s;
    if (lane == 0) s_dwave[wid] = vs;
    __syncthreads();
    unsigned long long off = 0;
    for (int w = 0; w < wid; ++w) off += s_wave[w];
    s += off;
    if (i < a.L) a.cloc[i] = s;
    if (threadIdx.x == kWeightsBlock - 1) {
        a.bsum[blockIdx.x] = s;
        a.bexp[blockIdx.x] = (s_dwave[0] + s_dwave[1]) + (s_dwave[2] + s_dwave[3]);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && a.frame_next) *a.frame_next = *a.frame + 1;
}

struct AncArgs {
    const unsigned long long *cloc;
    const unsigned long long *bsum;
    const double *bexp;
    FrameScratch *fs;
    const int *frame;
    const uint32_t *uni;          // U = uni[t * ustride + uoff]
    double *lml;                  // optional: lml[t] = log mean weight
    int *anc;
    int L, N, nb, ustride, uoff, kind;
};

// g(c): first output whose systematic target reaches cumulative weight c.
__device__ __forceinline__ int first_output(unsigned long long c, uint32_t U, unsigned long long T, int N)
{
    const u128 cs = (u128)c * ((u128)(unsigned long long)N << 32);
    double est = (double)c * (double)N / (double)T - (double)U * 2.3283064365386963e-10;
    long long i0 = (long long)ceil(est);
    if (i0 < 0) i0 = 0;
    if (i0 > N) i0 = N;
    while (i0 > 0 && sys_target((uint32_t)(i0 - 1), U, T) >= cs) --i0;
    while (i0 < N && sys_target((uint32_t)i0, U, T) < cs) ++i0;
    return (int)i0;
}

__global__ void __launch_bounds__(kWeightsBlock) k_ancestors(AncArgs a)
{
    __shared__ unsigned long long s_pre[4], s_tot[4];
    __shared__ double s_e[4];
    __shared__ int s_hi[kWeightsBlock];
    __shared__ int s_lo;
    const int t = *a.frame;
    const uint32_t U = a.uni[(size_t)t * a.ustride + a.uoff];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int b = blockIdx.x;
    // prefix of the block sums before this block, and the total
    unsigned long long pre = 0, tot = 0;
    double e = 0.0;
    for (int k = threadIdx.x; k < a.nb; k += kWeightsBlock) {
        const unsigned long long v = a.bsum[k];
        tot += v;
        if (k < b) pre += v;
        if (b == 0) e += a.bexp[k];
    }
    pre = wave_sum(pre);
    tot = wave_sum(tot);
    e = wave_sum(e);
    if (lane == 0) { s_pre[wid] = pre; s_tot[wid] = tot; s_e[wid] = e; }
    __syncthreads();
    pre = (s_pre[0] + s_pre[1]) + (s_pre[2] + s_pre[3]);
    tot = (s_tot[0] + s_tot[1]) + (s_tot[2] + s_tot[3]);
    const int i0 = b * kWeightsBlock;
    const int idx = i0 + threadIdx.x;
    const int n_in = min(kWeightsBlock, a.L - i0);
    if (threadIdx.x < n_in) s_hi[threadIdx.x] = first_output(pre + a.cloc[idx], U, tot, a.N);
    if (threadIdx.x == 0) s_lo = first_output(pre, U, tot, a.N);
    __syncthreads();
    const int lo = s_lo, hi = s_hi[n_in - 1];
    for (int i = lo + threadIdx.x; i < hi; i += kWeightsBlock) {
        int l = 0, h = n_in - 1;           // smallest k with s_hi[k] > i
        while (l < h) {
            const int mid = (l + h) >> 1;
            if (s_hi[mid] > i) h = mid;
            else l = mid + 1;
        }
        a.anc[i] = i0 + l;
    }
    if (b == 0 && threadIdx.x == 0) {
        const unsigned long long key = a.fs->maxkey;
        const double se = (s_e[0] + s_e[1]) + (s_e[2] + s_e[3]);
        if (key == 0ull) a.fs->status |= 1u;
        if (a.lml) {
            double l;
            if (key == 0ull) l = -INFINITY;
            else if (a.kind == kLinF64) l = log(se / (double)a.N);
            else l = key_value(key) + log(se) - log((double)a.N);
            a.lml[t] = l;
        }
        a.fs->maxkey = 0ull;
    }
}

// block maximum of FP64 host weights (for the stand-alone resample entry points)
template <int KIND>
__global__ void __launch_bounds__(256) k_blockmax64(const double *w, int L, FrameScratch *fs)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v = -INFINITY;
    if (i < L) {
        const double x = w[i];
        if (isfinite(x) && (KIND == kLogF64 || x > 0.0)) v = x;
    }
    v = wave_max(v);
    if ((threadIdx.x & 63) == 0 && isfinite(v)) atomicMax(&fs->maxkey, order_key(v));
}

// ---------------------------------------------------------------- helpers
__global__ void k_set_int(int *p, int v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

__global__ void k_iota(int *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

// my_gmm::resetTracker (my_gmm.cpp:30-42) from the device copy of mu / Sigma.
template <int D>
__global__ void __launch_bounds__(256) k_reset(const int *bins, const float *mu, const float *sig,
                                               float *x, float *P, int N, int ld)
{
    constexpr int NP = npacked(D);
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    const int b = bins[j];
#pragma unroll
    for (int i = 0; i < D; ++i) x[(size_t)i * ld + j] = mu[b * D + i];
#pragma unroll
    for (int p = 0; p < NP; ++p) P[(size_t)p * ld + j] = sig[b * NP + p];
}

// Gather tracks[j] = updated[anc_j] (the copy of pf2DRao.cpp:153-156).
template <int D>
__global__ void __launch_bounds__(256) k_gather(const float *xin, const float *Pin, const int *anc,
                                                float *xout, float *Pout, int N, int ld)
{
    constexpr int NP = npacked(D);
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    const int s = anc ? anc[j] : j;
#pragma unroll
    for (int i = 0; i < D; ++i) xout[(size_t)i * ld + j] = xin[(size_t)i * ld + s];
#pragma unroll
    for (int p = 0; p < NP; ++p) Pout[(size_t)p * ld + j] = Pin[(size_t)p * ld + s];
}

// getEstimator (pf2DRao.cpp:23-31): xbar = sum_j x[anc_j] / N, FP64
// accumulation, fixed reduction order (block partials, then the last block).
template <int D>
__global__ void __launch_bounds__(256) k_estimate(const float *x, const int *anc, int N, int ld,
                                                  double *partial, FrameScratch *fs, double *out)
{
    __shared__ double s_w[4][D];
    __shared__ int s_last;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    double v[D];
    const int s = (j < N) ? (anc ? anc[j] : j) : 0;
#pragma unroll
    for (int i = 0; i < D; ++i) v[i] = (j < N) ? (double)x[(size_t)i * ld + s] : 0.0;
#pragma unroll
    for (int i = 0; i < D; ++i) {
        const double r = wave_sum(v[i]);
        if (lane == 0) s_w[wid][i] = r;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 0; i < D; ++i) {
            double r = 0.0;
            for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r += s_w[w][i];
            partial[(size_t)blockIdx.x * D + i] = r;
        }
        __threadfence();
        s_last = (atomicAdd(&fs->done_est, 1u) == gridDim.x - 1);
        if (s_last) __threadfence();
    }
    __syncthreads();
    if (!s_last) return;
    if (threadIdx.x < D) {
        double r = 0.0;
        for (int b = 0; b < (int)gridDim.x; ++b) r += ((volatile double *)partial)[(size_t)b * D + threadIdx.x];
        out[threadIdx.x] = r / (double)N;
    }
    if (threadIdx.x == 0) fs->done_est = 0u;
}

// getSamples (pf2DRao.cpp:85-103): out = y[0:2] + 0.8 scale n, y = Hp xbar + M.
__global__ void k_samples(const double *hp2, const double *m2, const double *xbar, int d,
                          const double *normals, int n, double scale, double *out)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    double y0 = m2[0], y1 = m2[1];
    for (int k = 0; k < d; ++k) {
        y0 = fma(hp2[k], xbar[k], y0);
        y1 = fma(hp2[d + k], xbar[k], y1);
    }
    const double sd = 0.8 * scale;
    out[c] = fma(sd, normals[c], y0);
    out[n + c] = fma(sd, normals[n + c], y1);
}

// getSampleProb (pf2DRao.cpp:105-122): N(in_c; y[0:2], 0.8 scale I) for two inputs.
__global__ void k_sample_prob(const double *hp2, const double *m2, const double *xbar, int d,
                              const double *in1, const double *in2, int n, double scale,
                              double *w1, double *w2)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    double y0 = m2[0], y1 = m2[1];
    for (int k = 0; k < d; ++k) {
        y0 = fma(hp2[k], xbar[k], y0);
        y1 = fma(hp2[d + k], xbar[k], y1);
    }
    const double var = 0.8 * scale;
    const double lnorm = -log(6.283185307179586 * var);
    const double a0 = in1[c] - y0, a1 = in1[n + c] - y1;
    const double b0 = in2[c] - y0, b1 = in2[n + c] - y1;
    w1[c] = exp(lnorm - 0.5 * (a0 * a0 + a1 * a1) / var);
    w2[c] = exp(lnorm - 0.5 * (b0 * b0 + b1 * b1) / var);
}

// KF_model::predict / update for n states sharing one model (generic F, H):
// one block per state, one thread per matrix element, FP64 (the class API of
// KF_model.h:8-16 is double precision; this is not the filter's hot loop).
__global__ void k_kf_predict(int d, const double *F, const double *B, const double *Q, double *x,
                             double *P)
{
    extern __shared__ double s[];
    double *sP = s, *sT = s + d * d, *sx = s + 2 * d * d;
    const int st = blockIdx.x;
    double *xs = x + (size_t)st * d, *Ps = P + (size_t)st * d * d;
    const int tid = threadIdx.x;
    for (int e = tid; e < d * d; e += blockDim.x) sP[e] = Ps[e];
    for (int e = tid; e < d; e += blockDim.x) sx[e] = xs[e];
    __syncthreads();
    for (int e = tid; e < d * d; e += blockDim.x) {  // T = F P
        const int i = e / d, jj = e % d;
        double acc = 0.0;
        for (int k = 0; k < d; ++k) acc = fma(F[i * d + k], sP[k * d + jj], acc);
        sT[e] = acc;
    }
    __syncthreads();
    for (int e = tid; e < d * d; e += blockDim.x) {  // P = T F^T + Q
        const int i = e / d, jj = e % d;
        double acc = Q[e];
        for (int k = 0; k < d; ++k) acc = fma(sT[i * d + k], F[jj * d + k], acc);
        Ps[e] = acc;
    }
    for (int e = tid; e < d; e += blockDim.x) {      // x = F x + B
        double acc = B[e];
        for (int k = 0; k < d; ++k) acc = fma(F[e * d + k], sx[k], acc);
        xs[e] = acc;
    }
}

__global__ void k_kf_update(int d, int m, const double *H, const double *BH, const double *R,
                            const double *z, double *x, double *P, int *status)
{
    extern __shared__ double s[];
    double *sP = s;                 // d*d
    double *sHP = sP + d * d;       // m*d
    double *sS = sHP + m * d;       // m*m  (becomes the Cholesky factor)
    double *sW = sS + m * m;        // m*d  W = L^-1 H P
    double *sx = sW + m * d;        // d
    double *sy = sx + d;            // m    y, then v = L^-1 y
    const int st = blockIdx.x;
    double *xs = x + (size_t)st * d, *Ps = P + (size_t)st * d * d;
    const double *zs = z + (size_t)st * m;
    const int tid = threadIdx.x;
    for (int e = tid; e < d * d; e += blockDim.x) sP[e] = Ps[e];
    for (int e = tid; e < d; e += blockDim.x) sx[e] = xs[e];
    __syncthreads();
    for (int e = tid; e < m * d; e += blockDim.x) {  // HP
        const int r = e / d, c = e % d;
        double acc = 0.0;
        for (int k = 0; k < d; ++k) acc = fma(H[r * d + k], sP[k * d + c], acc);
        sHP[e] = acc;
    }
    for (int e = tid; e < m; e += blockDim.x) {      // y = z - (H x + BH)
        double acc = BH[e];
        for (int k = 0; k < d; ++k) acc = fma(H[e * d + k], sx[k], acc);
        sy[e] = zs[e] - acc;
    }
    __syncthreads();
    for (int e = tid; e < m * m; e += blockDim.x) {  // S = HP H^T + R
        const int r = e / m, c = e % m;
        double acc = R[e];
        for (int k = 0; k < d; ++k) acc = fma(sHP[r * d + k], H[c * d + k], acc);
        sS[e] = acc;
    }
    __syncthreads();
    if (tid == 0) {                                  // Cholesky of S (m <= 16), v = L^-1 y
        bool ok = true;
        for (int jj = 0; jj < m && ok; ++jj) {
            double sum = sS[jj * m + jj];
            for (int k = 0; k < jj; ++k) sum -= sS[jj * m + k] * sS[jj * m + k];
            if (!(sum > 0.0)) { ok = false; break; }
            const double l = sqrt(sum);
            sS[jj * m + jj] = l;
            for (int i = jj + 1; i < m; ++i) {
                double t = sS[i * m + jj];
                for (int k = 0; k < jj; ++k) t -= sS[i * m + k] * sS[jj * m + k];
                sS[i * m + jj] = t / l;
            }
        }
        if (!ok) atomicOr(status, 1);
        for (int i = 0; i < m; ++i) {
            double t = sy[i];
            for (int k = 0; k < i; ++k) t -= sS[i * m + k] * sy[k];
            sy[i] = t / sS[i * m + i];
        }
    }
    __syncthreads();
    for (int c = tid; c < d; c += blockDim.x) {      // W = L^-1 HP, column c
        for (int i = 0; i < m; ++i) {
            double t = sHP[i * d + c];
            for (int k = 0; k < i; ++k) t -= sS[i * m + k] * sW[k * d + c];
            sW[i * d + c] = t / sS[i * m + i];
        }
    }
    __syncthreads();
    for (int e = tid; e < d; e += blockDim.x) {      // x += W^T v
        double acc = sx[e];
        for (int k = 0; k < m; ++k) acc = fma(sW[k * d + e], sy[k], acc);
        xs[e] = acc;
    }
    for (int e = tid; e < d * d; e += blockDim.x) {  // P -= W^T W
        const int i = e / d, c = e % d;
        double acc = 0.0;
        for (int k = 0; k < m; ++k) acc = fma(sW[k * d + i], sW[k * d + c], acc);
        Ps[e] = sP[e] - acc;
    }
}

}  // namespace mkfbt
