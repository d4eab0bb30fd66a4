// pf_rng.hpp -- counter-based random numbers shared by host and device code.
//
// philox_block is the Philox4x32-10 block function (Salmon, Moraes, Dror and
// Shaw, SC'11); the word stream of pf_philox_u32 is word i of stream s =
// philox_block(ctr = {i/4 low, i/4 high, s, 0}, key = seed)[i % 4].
//
// normal_pair turns two 32-bit words into two standard normals by the
// Box-Muller transform, with its logarithm, sine and cosine evaluated from
// IEEE-exact operations only (fma, +, *, /, sqrt, frexp): host and device
// builds return identical bits, so the proposal positions a kernel draws can
// be recomputed anywhere from (seed, stream, word index).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pf_common.hpp"

namespace mkfbt {

__host__ __device__ inline void philox_block(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                             uint32_t k1, uint32_t out[4])
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// Block b of stream s under key seed.
__host__ __device__ inline void philox_stream_block(uint64_t seed, uint32_t stream, uint64_t b, uint32_t out[4])
{
    philox_block((uint32_t)b, (uint32_t)(b >> 32), stream, 0u, (uint32_t)seed, (uint32_t)(seed >> 32), out);
}

// log(x) for finite x > 0 in the normal range.  x = m 2^e with m in
// [sqrt(1/2), sqrt(2)); log m = 2 atanh(s), s = (m - 1)/(m + 1), |s| < 0.172,
// by its series through s^25; e ln 2 with the fdlibm two-part ln 2.
__host__ __device__ inline double det_log(double x)
{
#pragma clang fp contract(off)
    int e;
    double m = frexp(x, &e);
    if (m < 0.70710678118654752440) {
        m = m + m;
        e -= 1;
    }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double p = 1.0 / 25.0;
    p = dfma(p, z, 1.0 / 23.0);
    p = dfma(p, z, 1.0 / 21.0);
    p = dfma(p, z, 1.0 / 19.0);
    p = dfma(p, z, 1.0 / 17.0);
    p = dfma(p, z, 1.0 / 15.0);
    p = dfma(p, z, 1.0 / 13.0);
    p = dfma(p, z, 1.0 / 11.0);
    p = dfma(p, z, 1.0 / 9.0);
    p = dfma(p, z, 1.0 / 7.0);
    p = dfma(p, z, 1.0 / 5.0);
    p = dfma(p, z, 1.0 / 3.0);
    const double s2 = s + s;
    const double lm = fma(s2 * z, p, s2);
    const double de = (double)e;
    return fma(de, 6.93147180369123816490e-01, fma(de, 1.90821492927058770002e-10, lm));
}

// sin and cos of 2 pi w / 2^32.  The top two bits pick the quadrant; within
// it the angle (pi/2) g 2^-30 is taken on [0, pi/4] (complement above half a
// quadrant) where the Taylor series through x^19 (sin) and x^18 (cos) are
// below one unit in the last place.
__host__ __device__ inline void det_sincos_turn(uint32_t w, double &sn, double &cs)
{
#pragma clang fp contract(off)
    const uint32_t q = w >> 30;
    const uint32_t fr = w & 0x3FFFFFFFu;
    const bool flip = fr > 0x20000000u;
    const uint32_t g = flip ? 0x40000000u - fr : fr;
    const double x = (double)g * (1.5707963267948966 / 1073741824.0);
    const double z = x * x;
    double ps = -1.0 / 121645100408832000.0;
    ps = dfma(ps, z, 1.0 / 355687428096000.0);
    ps = dfma(ps, z, -1.0 / 1307674368000.0);
    ps = dfma(ps, z, 1.0 / 6227020800.0);
    ps = dfma(ps, z, -1.0 / 39916800.0);
    ps = dfma(ps, z, 1.0 / 362880.0);
    ps = dfma(ps, z, -1.0 / 5040.0);
    ps = dfma(ps, z, 1.0 / 120.0);
    ps = dfma(ps, z, -1.0 / 6.0);
    const double s0 = fma(x * z, ps, x);
    double pc = -1.0 / 6402373705728000.0;
    pc = dfma(pc, z, 1.0 / 20922789888000.0);
    pc = dfma(pc, z, -1.0 / 87178291200.0);
    pc = dfma(pc, z, 1.0 / 479001600.0);
    pc = dfma(pc, z, -1.0 / 3628800.0);
    pc = dfma(pc, z, 1.0 / 40320.0);
    pc = dfma(pc, z, -1.0 / 720.0);
    pc = dfma(pc, z, 1.0 / 24.0);
    pc = dfma(pc, z, -0.5);
    const double c0 = fma(z, pc, 1.0);
    const double a = flip ? c0 : s0;   // sin of the angle within the quadrant
    const double b = flip ? s0 : c0;   // cos
    switch (q) {
    case 0: sn = a; cs = b; break;
    case 1: sn = b; cs = -a; break;
    case 2: sn = -a; cs = -b; break;
    default: sn = -b; cs = a; break;
    }
}

// Two standard normals from two words: u1 = (w0 + 1) 2^-32 in (0, 1],
// r = sqrt(-2 log u1), (z0, z1) = r (cos, sin)(2 pi w1 2^-32).
__host__ __device__ inline void normal_pair(uint32_t w0, uint32_t w1, double &z0, double &z1)
{
#pragma clang fp contract(off)
    const double u1 = ((double)w0 + 1.0) * 2.3283064365386962890625e-10;
    const double r = sqrt(-2.0 * det_log(u1));
    double sn, cs;
    det_sincos_turn(w1, sn, cs);
    z0 = r * cs;
    z1 = r * sn;
}

// A word as a uniform in [0, 1): w 2^-32 (exact).
__host__ __device__ inline double unit_interval(uint32_t w) { return (double)w * 2.3283064365386962890625e-10; }

}  // namespace mkfbt
