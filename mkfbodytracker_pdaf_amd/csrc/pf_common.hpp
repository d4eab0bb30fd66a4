// pf_common.hpp -- helpers shared by host and device code of the particle
// filter library: packed symmetric indexing, the deterministic exp used by the
// resampler, fixed-point weights and 128-bit threshold arithmetic.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

namespace mkfbt {

// fma(a, b, c) for the deterministic polynomials.  On the device it is the
// three-address v_fma_f64: the compiler's own FP64 FMA selection ties the
// addend to the result register, so a loop-invariant coefficient held in a
// register is copied before every Horner step; this form reads it in place.
// Either way one correctly rounded fused multiply-add, as on the host.
__host__ __device__ inline double dfma(double a, double b, double c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double r;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
#else
    return fma(a, b, c);
#endif
}

typedef unsigned __int128 u128;

// Packed upper-triangular index of (i, j), i <= j, row-major, n x n.
__host__ __device__ constexpr int pk(int n, int i, int j)
{
    return i <= j ? i * n - (i * (i - 1)) / 2 + (j - i) : j * n - (j * (j - 1)) / 2 + (i - j);
}

__host__ __device__ constexpr int npacked(int n) { return n * (n + 1) / 2; }

// exp(x) for x <= 0 from IEEE-exact operations only (rint, fma, ldexp):
// Cody-Waite reduction with the fdlibm split of ln 2 and the degree-12 Taylor
// polynomial on |r| <= ln2/2.  Host and device builds return identical bits,
// which makes the fixed-point resampling weights below reproducible anywhere.
__host__ __device__ inline double det_exp(double x)
{
    if (!(x >= -745.0)) return 0.0;
    if (x > 709.0) x = 709.0;
    const double n = rint(x * 1.4426950408889634074);
    double r = fma(-n, 6.93147180369123816490e-01, x);
    r = fma(-n, 1.90821492927058770002e-10, r);
    double p = 1.0 / 479001600.0;
    p = dfma(p, r, 1.0 / 39916800.0);
    p = dfma(p, r, 1.0 / 3628800.0);
    p = dfma(p, r, 1.0 / 362880.0);
    p = dfma(p, r, 1.0 / 40320.0);
    p = dfma(p, r, 1.0 / 5040.0);
    p = dfma(p, r, 1.0 / 720.0);
    p = dfma(p, r, 1.0 / 120.0);
    p = dfma(p, r, 1.0 / 24.0);
    p = dfma(p, r, 1.0 / 6.0);
    p = dfma(p, r, 0.5);
    p = dfma(p, r, 1.0);
    p = dfma(p, r, 1.0);
    return ldexp(p, (int)n);
}

// v in [0, 1] -> floor(v * 2^32), exactly.
__host__ __device__ inline uint64_t fixed32(double v)
{
    if (!(v > 0.0)) return 0;
    const double t = floor(ldexp(v, 32));
    if (t >= 4294967296.0) return (uint64_t)1 << 32;
    return (uint64_t)(uint32_t)t;
}

// Fixed-point weight of a log-weight relative to the maximum.
__host__ __device__ inline uint64_t qlog(double lw, double lmax)
{
    return (std::isfinite(lw) && std::isfinite(lmax)) ? fixed32(det_exp(lw - lmax)) : 0;
}

// Block-scaled log-weights (the resampler of log-weights).  A block of
// kWeightsBlock consecutive log-weights whose largest finite value is m is
// scaled by 2^-e with e = block_pow(m) = ceil(m log2 e), so that
// q_i = floor(2^32 exp(lw_i - e ln 2)) <= 2^32 and the block's largest q is
// at least 2^31; a block without a finite log-weight has e = kPowEmpty and
// q = 0.  Blocks are aligned by right shifts of E - e, E the largest e, so
// every prefix sum stays an exact integer and no pass over the weights needs
// the global maximum first.
constexpr int kPowEmpty = -2147483647 - 1;
constexpr int kPowLimit = 1 << 30;

__host__ __device__ inline int block_pow(double m)
{
    const double c = ceil(m * 1.4426950408889634);
    return c > (double)kPowLimit ? kPowLimit : (c < -(double)kPowLimit ? -kPowLimit : (int)c);
}

__host__ __device__ inline double pow_ref(int e) { return (double)e * 0.6931471805599453; }

__host__ __device__ inline uint64_t qlog_pow(double lw, int e)
{
    return (std::isfinite(lw) && e != kPowEmpty) ? fixed32(det_exp(lw - pow_ref(e))) : 0;
}

// Right shift aligning a block of exponent e to the largest exponent E (64:
// the block contributes nothing).
__host__ __device__ inline int pow_shift(int E, int e)
{
    if (e == kPowEmpty || E == kPowEmpty) return 64;
    const long long d = (long long)E - (long long)e;
    return d >= 64 ? 64 : (int)d;
}

__host__ __device__ inline uint64_t shr64(uint64_t x, int s) { return s >= 64 ? 0 : x >> s; }

// Systematic threshold of output i: (i * 2^32 + U) * total.
__host__ __device__ inline u128 sys_target(uint32_t i, uint32_t U, uint64_t total)
{
    return (u128)((((uint64_t)i) << 32) | U) * total;
}

// Model image of the step kernels: [component thresholds, 4K floats][H
// (m x D), BH (m), R 2x2 blocks (3J), padded to even][per joint the FP64
// innovation block (bh0, bh1, r00, r01, r11) as doubles (5J), padded to 4]
// [component rows, K x (2 + D + D(D+1)/2)].  Offset of the FP64 blocks (8-byte
// aligned) and floats of the part before the component rows:
__host__ __device__ constexpr int image_j64_floats(int K, int m, int D, int J)
{
    return 4 * K + ((m * D + m + 3 * J + 1) & ~1);
}
__host__ __device__ constexpr int image_head_floats(int K, int m, int D, int J)
{
    return 4 * K + ((((m * D + m + 3 * J + 1) & ~1) + 10 * J + 3) & ~3);
}

// A 2 x 2 innovation covariance S = [[s00, s01], [s01, s11]] whose FP32
// Cholesky is not trusted: a pivot that is not positive, or
// det S <= 1e-6 (tr S)^2, i.e. cond(S) >= ~1e6 (for a 2 x 2 SPD matrix
// det / tr^2 = c / (1 + c)^2 with c = 1 / cond).  Such joints are factored in FP64.
__host__ __device__ inline bool innovation_ill(float s00, float s01, float s11)
{
    const float tr = s00 + s11;
    const float det = fmaf(s00, s11, -s01 * s01);
    return !(s00 > 0.f) || !(det > 1e-6f * tr * tr);
}

// Status bits of the device work (FrameScratch::status and the batched
// engine's per-sequence status): the uniform resampling fallback (no positive
// weight), a joint factored in FP64, a joint whose innovation covariance is
// not positive definite even in FP64 (the particle's weight is zero).
constexpr unsigned kStatusDegenerate = 1u;
constexpr unsigned kStatusIllCond = 2u;
constexpr unsigned kStatusNotSpd = 4u;
// LDS budget of the staged image: above it only the head is staged and the
// component rows are read from global memory.
constexpr int kImageLdsBytes = 48 * 1024;

}  // namespace mkfbt
