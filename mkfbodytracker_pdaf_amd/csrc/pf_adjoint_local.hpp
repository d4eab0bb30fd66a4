// pf_adjoint_local.hpp -- the local function of one 2-row joint update and
// its reverse, for the adjoint of the RBPF log-marginal-likelihood.
//
// A joint with measured rows r (2 of them) updates a particle as
//   x += C g,  P -= C N C^T,  lw += ell,   C = P[:, r],
// where g (2), N (2 x 2 symmetric) and ell depend only on the local inputs
// y = x[r] + BH (2) and S = P[r, r] + R (symmetric 2 x 2: S00, S01, S11) and
// on the frame's detections:
//   certain joint (one detection z):  u = L^-1 (z - y), g = L^-T u,
//     N = S^-1, ell = -log(2 pi) - log det L - |u|^2 / 2;
//   PDAF joint (gated candidates z_c, Bar-Shalom & Fortmann), whitened
//     u_c = L^-1 (z_c - y), e_c = (P_D / lambda) N(z_c; y, S),
//     l = b0 + sum e_c, w = sum e_c u_c / l, M = (sum e_c / l) I - (sum e_c u_c u_c^T / l - w w^T),
//     g = L^-T w, N = L^-T M L^-1, ell = log l.
// (S = L L^T.)  This is what Joint::prepare_sel + joint_measure + apply do in
// the forward, in unwhitened form.
//
// Two implementations of the same derivatives d(g, N, ell)/d(y, S):
//   local_joint_dual -- forward mode, dual numbers along the 5 inputs;
//   local_joint_rev  -- hand-derived reverse mode (used by the adjoint kernel).
// Both are host- and device-callable, so they can be compared on the CPU.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace mkfbt {
namespace adj {

__host__ __device__ inline float fexp(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __expf(x);
#else
    return std::exp(x);
#endif
}

// ------------------------------------------------------------ dual numbers
struct D5 {
    float v, d[5];
};
__host__ __device__ inline D5 dc(float v)
{
    D5 r;
    r.v = v;
    for (int i = 0; i < 5; ++i) r.d[i] = 0.f;
    return r;
}
__host__ __device__ inline D5 dseed(float v, int i)
{
    D5 r = dc(v);
    r.d[i] = 1.f;
    return r;
}
__host__ __device__ inline D5 operator+(const D5 &a, const D5 &b)
{
    D5 r;
    r.v = a.v + b.v;
    for (int i = 0; i < 5; ++i) r.d[i] = a.d[i] + b.d[i];
    return r;
}
__host__ __device__ inline D5 operator-(const D5 &a, const D5 &b)
{
    D5 r;
    r.v = a.v - b.v;
    for (int i = 0; i < 5; ++i) r.d[i] = a.d[i] - b.d[i];
    return r;
}
__host__ __device__ inline D5 operator*(const D5 &a, const D5 &b)
{
    D5 r;
    r.v = a.v * b.v;
    for (int i = 0; i < 5; ++i) r.d[i] = fmaf(a.d[i], b.v, a.v * b.d[i]);
    return r;
}
__host__ __device__ inline D5 operator*(float s, const D5 &a)
{
    D5 r;
    r.v = s * a.v;
    for (int i = 0; i < 5; ++i) r.d[i] = s * a.d[i];
    return r;
}
__host__ __device__ inline D5 operator+(float s, const D5 &a)
{
    D5 r = a;
    r.v = s + a.v;
    return r;
}
__host__ __device__ inline D5 drcp(const D5 &a)
{
    D5 r;
    r.v = 1.f / a.v;
    const float g = -r.v * r.v;
    for (int i = 0; i < 5; ++i) r.d[i] = g * a.d[i];
    return r;
}
__host__ __device__ inline D5 dsqrt(const D5 &a)
{
    D5 r;
    r.v = sqrtf(a.v);
    const float g = 0.5f / r.v;
    for (int i = 0; i < 5; ++i) r.d[i] = g * a.d[i];
    return r;
}
__host__ __device__ inline D5 dlog(const D5 &a)
{
    D5 r;
    r.v = logf(a.v);
    const float g = 1.f / a.v;
    for (int i = 0; i < 5; ++i) r.d[i] = g * a.d[i];
    return r;
}
__host__ __device__ inline D5 dexp(const D5 &a)
{
    D5 r;
    r.v = fexp(a.v);
    for (int i = 0; i < 5; ++i) r.d[i] = r.v * a.d[i];
    return r;
}

struct LocalDual {
    D5 g0, g1, n00, n01, n11, ell;
};

// Constants of the PDAF likelihood: gate (chi2 threshold), pdl = P_D / lambda,
// b0 = 1 - P_D P_G.
struct PdafConst {
    float gate, pdl, b0;
};

__host__ __device__ inline LocalDual local_joint_dual(const PdafConst &k, float y0f, float y1f, float s00f,
                                                      float s01f, float s11f, bool clutter, int cnt, const float *cz)
{
    const float LOG2PI = 1.8378770664093453f;
    const D5 y0 = dseed(y0f, 0), y1 = dseed(y1f, 1);
    const D5 s00 = dseed(s00f, 2), s01 = dseed(s01f, 3), s11 = dseed(s11f, 4);
    const D5 l00 = dsqrt(s00);
    const D5 il00 = drcp(l00);
    const D5 l10 = s01 * il00;
    const D5 l11 = dsqrt(s11 - l10 * l10);
    const D5 il11 = drcp(l11);
    const D5 a10 = dc(0.f) - (l10 * il00) * il11;         // (L^-1)_{10}
    const D5 lnorm = dc(-LOG2PI) - dlog(l00 * l11);
    LocalDual o;
    if (!clutter) {
        const D5 n0 = dc(cz[0]) - y0, n1 = dc(cz[1]) - y1;
        const D5 u0 = n0 * il00;
        const D5 u1 = a10 * n0 + il11 * n1;
        o.g0 = il00 * u0 + a10 * u1;
        o.g1 = il11 * u1;
        o.n00 = il00 * il00 + a10 * a10;
        o.n01 = a10 * il11;
        o.n11 = il11 * il11;
        o.ell = lnorm - 0.5f * (u0 * u0 + u1 * u1);
        return o;
    }
    D5 se = dc(0.f), sv0 = dc(0.f), sv1 = dc(0.f), q00 = dc(0.f), q01 = dc(0.f), q11 = dc(0.f);
    for (int c = 0; c < cnt; ++c) {
        const D5 n0 = dc(cz[2 * c]) - y0, n1 = dc(cz[2 * c + 1]) - y1;
        const D5 u0 = n0 * il00;
        const D5 u1 = a10 * n0 + il11 * n1;
        const D5 d2 = u0 * u0 + u1 * u1;
        if (d2.v <= k.gate) {                               // gating is a discrete choice: held fixed
            const D5 e = k.pdl * dexp(lnorm - 0.5f * d2);
            se = se + e;
            sv0 = sv0 + e * u0;
            sv1 = sv1 + e * u1;
            q00 = q00 + (e * u0) * u0;
            q01 = q01 + (e * u0) * u1;
            q11 = q11 + (e * u1) * u1;
        }
    }
    const D5 ells = k.b0 + se;
    const D5 inv = drcp(ells);
    const D5 w0 = sv0 * inv, w1 = sv1 * inv;
    const D5 bs = se * inv;
    const D5 m00 = bs - (q00 * inv - w0 * w0);
    const D5 m01 = dc(0.f) - (q01 * inv - w0 * w1);
    const D5 m11 = bs - (q11 * inv - w1 * w1);
    o.g0 = il00 * w0 + a10 * w1;
    o.g1 = il11 * w1;
    const D5 A00 = m00 * il00 + m01 * a10, A01 = m01 * il11;
    const D5 A10 = m01 * il00 + m11 * a10, A11 = m11 * il11;
    o.n00 = il00 * A00 + a10 * A10;
    o.n01 = il00 * A01 + a10 * A11;
    o.n11 = il11 * A11;
    o.ell = dlog(ells);
    return o;
}

// ------------------------------------------------------------ reverse mode
// Values (g, N) and, given the adjoints gb = dF/dg, Nb = dF/dN (symmetric,
// full-matrix convention: dF = sum_ab Nb_ab dN_ab) and ellb = dF/d ell, the
// adjoints of the local inputs: yb (2) and Sb = (dF/dS00, dF/dS01, dF/dS11)
// with S01 the one off-diagonal parameter.
struct LocalRev {
    float g0, g1, n00, n01, n11;
    float yb0, yb1, sb00, sb01, sb11;
};

__host__ __device__ inline LocalRev local_joint_rev(const PdafConst &k, float y0, float y1, float s00, float s01,
                                                    float s11, bool clutter, int cnt, const float *cz, float gb0,
                                                    float gb1, float nb00, float nb01, float nb11, float ellb)
{
    const float LOG2PI = 1.8378770664093453f;
    // ---- forward: Cholesky factor and its inverse A = L^-1 = [[a00, 0], [a10, a11]]
    const float l00 = sqrtf(s00);
    const float a00 = 1.f / l00;
    const float l10 = s01 * a00;
    const float l11 = sqrtf(s11 - l10 * l10);
    const float a11 = 1.f / l11;
    const float a10 = -(l10 * a00) * a11;
    const float lnorm = -LOG2PI - logf(l00 * l11);
    LocalRev o;
    // adjoints of A (lower entries), of lnorm, of y
    float A00b = 0.f, A10b = 0.f, A11b = 0.f, lnb = 0.f, yb0 = 0.f, yb1 = 0.f;
    if (!clutter) {
        const float n0 = cz[0] - y0, n1 = cz[1] - y1;
        const float u0 = a00 * n0, u1 = fmaf(a10, n0, a11 * n1);
        o.g0 = fmaf(a00, u0, a10 * u1);
        o.g1 = a11 * u1;
        o.n00 = fmaf(a00, a00, a10 * a10);
        o.n01 = a10 * a11;
        o.n11 = a11 * a11;
        // g = A^T u: ub = A gb, Ab += u gb^T
        float ub0 = a00 * gb0;
        float ub1 = fmaf(a10, gb0, a11 * gb1);
        A00b = fmaf(u0, gb0, A00b);
        A10b = fmaf(u1, gb0, A10b);
        A11b = fmaf(u1, gb1, A11b);
        // N = A^T A: Ab += 2 A Nb (lower entries)
        A00b += 2.f * (a00 * nb00);
        A10b += 2.f * fmaf(a10, nb00, a11 * nb01);
        A11b += 2.f * fmaf(a10, nb01, a11 * nb11);
        // ell = lnorm - |u|^2 / 2
        lnb += ellb;
        ub0 = fmaf(-ellb, u0, ub0);
        ub1 = fmaf(-ellb, u1, ub1);
        // u = A n: Ab += ub n^T, nb = A^T ub, yb = -nb
        A00b = fmaf(ub0, n0, A00b);
        A10b = fmaf(ub1, n0, A10b);
        A11b = fmaf(ub1, n1, A11b);
        yb0 = -fmaf(a00, ub0, a10 * ub1);
        yb1 = -(a11 * ub1);
    } else {
        // forward sums over the gated candidates
        float se = 0.f, sv0 = 0.f, sv1 = 0.f, q00 = 0.f, q01 = 0.f, q11 = 0.f;
        for (int c = 0; c < cnt; ++c) {
            const float n0 = cz[2 * c] - y0, n1 = cz[2 * c + 1] - y1;
            const float u0 = a00 * n0, u1 = fmaf(a10, n0, a11 * n1);
            const float d2 = fmaf(u0, u0, u1 * u1);
            if (d2 <= k.gate) {
                const float e = k.pdl * fexp(fmaf(-0.5f, d2, lnorm));
                se += e;
                sv0 = fmaf(e, u0, sv0);
                sv1 = fmaf(e, u1, sv1);
                q00 = fmaf(e * u0, u0, q00);
                q01 = fmaf(e * u0, u1, q01);
                q11 = fmaf(e * u1, u1, q11);
            }
        }
        const float ls = k.b0 + se;
        const float inv = 1.f / ls;
        const float w0 = sv0 * inv, w1 = sv1 * inv;
        const float bs = se * inv;
        const float m00 = bs - (q00 * inv - w0 * w0);
        const float m01 = -(q01 * inv - w0 * w1);
        const float m11 = bs - (q11 * inv - w1 * w1);
        o.g0 = fmaf(a00, w0, a10 * w1);
        o.g1 = a11 * w1;
        const float B00 = fmaf(m00, a00, m01 * a10), B01 = m01 * a11;     // B = M A
        const float B10 = fmaf(m01, a00, m11 * a10), B11 = m11 * a11;
        o.n00 = fmaf(a00, B00, a10 * B10);
        o.n01 = fmaf(a00, B01, a10 * B11);
        o.n11 = a11 * B11;
        // ---- reverse
        // g = A^T w: wb = A gb, Ab += w gb^T
        float wb0 = a00 * gb0;
        float wb1 = fmaf(a10, gb0, a11 * gb1);
        A00b = fmaf(w0, gb0, A00b);
        A10b = fmaf(w1, gb0, A10b);
        A11b = fmaf(w1, gb1, A11b);
        // N = A^T M A: Mb = A Nb A^T, Ab += 2 M A Nb = 2 B Nb (lower entries)
        const float T00 = a00 * nb00, T01 = a00 * nb01;                   // T = A Nb
        const float T10 = fmaf(a10, nb00, a11 * nb01), T11 = fmaf(a10, nb01, a11 * nb11);
        const float mb00 = T00 * a00;                                     // Mb = T A^T
        const float mb01 = fmaf(T00, a10, T01 * a11);
        const float mb11 = fmaf(T10, a10, T11 * a11);
        A00b += 2.f * fmaf(B00, nb00, B01 * nb01);
        A10b += 2.f * fmaf(B10, nb00, B11 * nb01);
        A11b += 2.f * fmaf(B10, nb01, B11 * nb11);
        // M = bs I - Q inv + w w^T
        const float bsb = mb00 + mb11;
        const float qb00 = -inv * mb00, qb01 = -inv * mb01, qb11 = -inv * mb11;   // Qb (symmetric, full convention)
        float invb = -(mb00 * q00 + 2.f * mb01 * q01 + mb11 * q11);
        wb0 += 2.f * fmaf(mb00, w0, mb01 * w1);
        wb1 += 2.f * fmaf(mb01, w0, mb11 * w1);
        // w = sv inv, bs = se inv, ell = log ls, inv = 1 / ls, ls = b0 + se
        const float svb0 = inv * wb0, svb1 = inv * wb1;
        invb = fmaf(sv0, wb0, fmaf(sv1, wb1, invb));
        float seb = inv * bsb;
        invb = fmaf(se, bsb, invb);
        const float lsb = fmaf(ellb, inv, -inv * inv * invb);
        seb += lsb;
        // candidates, reverse
        for (int c = 0; c < cnt; ++c) {
            const float n0 = cz[2 * c] - y0, n1 = cz[2 * c + 1] - y1;
            const float u0 = a00 * n0, u1 = fmaf(a10, n0, a11 * n1);
            const float d2 = fmaf(u0, u0, u1 * u1);
            if (d2 <= k.gate) {
                const float e = k.pdl * fexp(fmaf(-0.5f, d2, lnorm));
                // eb = seb + svb . u + u^T Qb u
                const float qu0 = fmaf(qb00, u0, qb01 * u1), qu1 = fmaf(qb01, u0, qb11 * u1);
                const float eb = seb + fmaf(svb0, u0, svb1 * u1) + fmaf(u0, qu0, u1 * qu1);
                // ub = e svb + 2 e Qb u - eb e u   (the last term through d2)
                const float ee = eb * e;
                const float ub0 = fmaf(e, svb0, fmaf(2.f * e, qu0, -ee * u0));
                const float ub1 = fmaf(e, svb1, fmaf(2.f * e, qu1, -ee * u1));
                lnb += ee;
                // u = A n
                A00b = fmaf(ub0, n0, A00b);
                A10b = fmaf(ub1, n0, A10b);
                A11b = fmaf(ub1, n1, A11b);
                yb0 -= fmaf(a00, ub0, a10 * ub1);
                yb1 -= a11 * ub1;
            }
        }
    }
    // ---- A = L^-1, lnorm = -log 2pi - log l00 - log l11, L = chol(S)
    float a00b = A00b, a11b = A11b;
    const float l10b0 = -A10b * a00 * a11;
    a00b = fmaf(-A10b * l10, a11, a00b);
    a11b = fmaf(-A10b * l10, a00, a11b);
    float l00b = -a00b * a00 * a00 - lnb * a00;
    float l11b = -a11b * a11 * a11 - lnb * a11;
    float l10b = l10b0;
    // l11 = sqrt(s11 - l10^2)
    const float sb11 = l11b * 0.5f * a11;
    l10b = fmaf(-l11b * l10, a11, l10b);
    // l10 = s01 / l00
    const float sb01 = l10b * a00;
    l00b = fmaf(-l10b * l10, a00, l00b);
    // l00 = sqrt(s00)
    const float sb00 = l00b * 0.5f * a00;
    o.yb0 = yb0;
    o.yb1 = yb1;
    o.sb00 = sb00;
    o.sb01 = sb01;
    o.sb11 = sb11;
    return o;
}

}  // namespace adj
}  // namespace mkfbt
