// pf_quad24.hpp -- the batched step of the 24-dimensional state with four
// lanes per particle and the covariance held as 4 x 4 matrix-core tiles
// (reference: src/KF_model.cpp:11-25, src/pf2DRao.cpp:125-158; the same
// per-particle algorithm as particle_update in pf_kernels.hpp).
//
// Lane layout: a wave holds 16 particles, particle q in lanes 4q .. 4q + 3;
// lane c of a quad holds
//   x[4K + c]                      K = 0..5 (6 values)
//   P[4I + r][4K + c], r = 0..3    in tile (I, K), I <= K (21 tiles, 84 values)
// i.e. column c of every upper-triangular 4 x 4 tile of the 24 x 24
// covariance, one register per tile row.  This is the accumulator layout of
// v_mfma_f32_4x4x1f32 (16 blocks of 4 x 4, block = quad, register = row,
// lane = column), whose operands are column slices of the same layout: the
// rank-1 update D += a b^T of every particle's tile (I, K) takes a from the
// lane's value for column 4I + c and b from its value for column 4K + c.
//
// Per particle and frame: component draw, mean-reverting predict, then per
// joint (selection basis: rows 2 oi, 2 oi + 1) the 2 x 2 innovation
// Cholesky, W = L^-1 P[rows, :], the likelihood (PDAF over the gated
// candidates, split over the quad's four lanes and summed with quad
// butterflies) and the update x += W^T v, P -= W^T M W written as
// P -= l1 z1 z1^T + l2 z2 z2^T (eigen-decomposition of the 2 x 2 M): two
// rank-1 updates of the 21 tiles, as 42 matrix-core instructions per wave
// (MFMA = true) or as 168 FMAs per lane (MFMA = false, the same layout on the
// vector ALU).  Each rank-1 term is applied as (s z)(s z)^T with s =
// sqrt|l| and the sign on one operand, so the diagonal tiles stay exactly
// symmetric.
#pragma once

#include "pf_batch_impl.hpp"

namespace mkfbt {
namespace batch {

typedef float q4f __attribute__((ext_vector_type(4)));

constexpr int kQ24Tiles = 21;     // upper-triangular 4 x 4 tiles of 24 x 24

__host__ __device__ constexpr int qtile(int I, int K) { return I * 6 - (I * (I - 1)) / 2 + (K - I); }

// Value of lane S of each quad, in all four lanes of the quad.
template <int S>
__device__ __forceinline__ float qbcast(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), S * 0x55, 0xF, 0xF, false));
}

// Sum over the quad, bit-identical in its four lanes ((a + b) + (c + d) in
// every lane: addition commutes exactly).
__device__ __forceinline__ float qsum(float v)
{
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));  // [1,0,3,2]
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, false));  // [2,3,0,1]
    return v;
}

__device__ __forceinline__ float qsel(int c, float v0, float v1, float v2, float v3)
{
    return c == 0 ? v0 : (c == 1 ? v1 : (c == 2 ? v2 : v3));
}

// Packed row (pk(24, ., .)) of entry (4I + r, 4I + c) of diagonal tile I.
__device__ __forceinline__ int qdiag_row(int I, int r, int c)
{
    const int lo = r < c ? r : c, hi = r < c ? c : r;
    return pk(24, 4 * I + lo, 4 * I + lo) + (hi - lo);
}

// Tile (I, K) += a b^T for every particle of the wave (a, b: the lanes'
// column slices).
template <bool MFMA>
__device__ __forceinline__ q4f qrank1(float a, float b, q4f acc)
{
    if constexpr (MFMA) {
        return __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, acc, 0, 0, 0);
    } else {
        acc[0] = fmaf(qbcast<0>(a), b, acc[0]);
        acc[1] = fmaf(qbcast<1>(a), b, acc[1]);
        acc[2] = fmaf(qbcast<2>(a), b, acc[2]);
        acc[3] = fmaf(qbcast<3>(a), b, acc[3]);
        return acc;
    }
}

// One 256-particle tile per block; each wave takes four rounds of 16
// particles (round r, wave w: tile particles 64 r + 16 w + q).
template <int SELJ, bool MFMA>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 8))) k_batch_step_q24(BatchArgs a)
{
    constexpr int D = 24, NP = npacked(24), CS = 2 + D + NP;
    extern __shared__ float4 sm4[];
    const int t = *a.frame;
    const int tile = xcd_tile();
    if (tile >= a.n_tiles[t]) return;                   // block-uniform: inactive sequences' tiles
    const bool odd = (t & 1) != 0;
    const float *xin = odd ? a.buf_x[1] : a.buf_x[0], *Pin = odd ? a.buf_P[1] : a.buf_P[0];
    float *xout = odd ? a.buf_x[0] : a.buf_x[1], *Pout = odd ? a.buf_P[0] : a.buf_P[1];
    const unsigned rowb = (unsigned)a.ld * 4u;
    unsigned long long *wsp = reinterpret_cast<unsigned long long *>(reinterpret_cast<float *>(sm4) + a.stage_floats);
    int *s_src = reinterpret_cast<int *>(wsp + resolve_lds_words(a.nb_max));
    const int s = uni_int(a.tile_seq[tile]);
    const int jt0 = uni_int(a.tile_j0[tile]);
    const SeqDesc sd = uniform_seq(a.seqs, s);
    int src[1];
    if (a.fused) {
        tile_ancestors<1>(a, sd, s, t - 1, jt0, wsp, src, a.rec_anc ? a.rec_anc + (size_t)(t - 1) * a.ld : nullptr,
                          nullptr, a.lml + sd.tofs + (t - 1));
    } else {
        const int j = jt0 + (int)threadIdx.x;
        src[0] = j < sd.N ? a.anc[sd.pofs + j] : 0;
    }
    s_src[threadIdx.x] = src[0];
    const float *fcand = a.cand + ((size_t)sd.tofs + t) * a.J * a.cap * 2;
    CandPrefetch cpf;
    cand_prefetch(cpf, a.count + ((size_t)sd.tofs + t) * a.J, fcand, a.J, a.cap);
    stage_table(sm4, reinterpret_cast<const float4 *>(a.tables + (size_t)sd.table_ofs), a.table_floats >> 2);
    cand_store(cpf, reinterpret_cast<float *>(sm4) + a.table_floats, fcand, a.J, a.cap);
    __syncthreads();

    const float *sm = reinterpret_cast<const float *>(sm4);
    const unsigned long long *sthr = reinterpret_cast<const unsigned long long *>(sm);
    const float *BHs = sm + 4 * a.K + a.m * D;
    const float *Rs = BHs + a.m;
    const double *J64 = reinterpret_cast<const double *>(sm + image_j64_floats(a.K, a.m, D, a.J));
    const float *comp_global = seq_comp_global(a, sd);
    const float *stab = comp_global ? comp_global : sm + image_head_floats(a.K, a.m, D, a.J);
    const uint32_t U1 = a.uni[(size_t)(sd.tofs + t) * 2];
    int *comp_out = a.rec_comp ? a.rec_comp + (size_t)t * a.ld + sd.pofs : nullptr;
    const int lane = (int)threadIdx.x & 63, w = (int)threadIdx.x >> 6, q = lane >> 2, c = lane & 3;
    const __amdgpu_buffer_rsrc_t rx = soa_rsrc(xin, rowb * D), rP = soa_rsrc(Pin, rowb * NP);
    const __amdgpu_buffer_rsrc_t wx = soa_rsrc(xout, rowb * D), wP = soa_rsrc(Pout, rowb * NP);
    const unsigned crow = (unsigned)c * rowb;
    // packed rows of the diagonal tiles' entries held by this lane
    int drow[6][4];
#pragma unroll
    for (int I = 0; I < 6; ++I)
#pragma unroll
        for (int r = 0; r < 4; ++r) drow[I][r] = qdiag_row(I, r, c);

#pragma unroll 1
    for (int round = 0; round < 4; ++round) {
        const int jl = 64 * round + 16 * w + q;
        const int j = jt0 + jl;
        const bool live = j < sd.N;                        // quad-uniform; dead particles read ancestor 0
        const unsigned vb = (unsigned)(sd.pofs + s_src[jl]) * 4u;
        float x[6];
        q4f Pt[kQ24Tiles];
#pragma unroll
        for (int K = 0; K < 6; ++K) x[K] = soa_ld(rx, vb + crow, rowb * (unsigned)(4 * K));
#pragma unroll
        for (int I = 0; I < 6; ++I)
#pragma unroll
            for (int K = I; K < 6; ++K)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    Pt[qtile(I, K)][r] = (K > I) ? soa_ld(rP, vb + crow, rowb * (unsigned)pk(24, 4 * I + r, 4 * K))
                                                 : soa_ld(rP, vb + (unsigned)drow[I][r] * rowb, 0);

        // component draw (pf2DRao.cpp:128) and predict (KF_model.cpp:11-15)
        const int k = draw_component(sthr, a.K, (uint32_t)j, U1, sd.pi_total);
        const float *ck = stab + k * CS;
        const float g = ck[0], g2 = ck[1];
#pragma unroll
        for (int K = 0; K < 6; ++K) x[K] = fmaf(g, x[K], ck[2 + 4 * K + c]);
        const float *cq = ck + 2 + D;
#pragma unroll
        for (int I = 0; I < 6; ++I)
#pragma unroll
            for (int K = I; K < 6; ++K)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const float qv = (K > I) ? cq[pk(24, 4 * I + r, 4 * K) + c] : cq[drow[I][r]];
                    Pt[qtile(I, K)][r] = fmaf(g2, Pt[qtile(I, K)][r], qv);
                }

        float acc = 0.f;
        unsigned pflags = 0;
#pragma unroll
        for (int oi = 0; oi < SELJ; ++oi) {
            const int b = (int)((a.order_packed >> (4 * oi)) & 15u);
            int cnt;
            const float *cz;
            bool clutter;
            if (!joint_inputs<kModeCand>(a.base, t, b, cnt, cz, clutter, sm)) continue;
            const int r0 = 2 * oi, I0 = r0 >> 2;
            // rows r0, r1 of P over the lane's columns 4K + c
            float R0[6], R1[6];
#pragma unroll
            for (int K = 0; K < 6; ++K) {
                if (K >= I0) {
                    R0[K] = Pt[qtile(I0, K)][r0 & 3];
                    R1[K] = Pt[qtile(I0, K)][(r0 & 3) + 1];
                } else {
                    // P[r0][4K + c] = P[4K + c][r0]: row c of tile (K, I0) in lane r0 & 3
                    const q4f &T = Pt[qtile(K, I0)];
                    if ((r0 & 3) == 0) {
                        R0[K] = qsel(c, qbcast<0>(T[0]), qbcast<0>(T[1]), qbcast<0>(T[2]), qbcast<0>(T[3]));
                        R1[K] = qsel(c, qbcast<1>(T[0]), qbcast<1>(T[1]), qbcast<1>(T[2]), qbcast<1>(T[3]));
                    } else {
                        R0[K] = qsel(c, qbcast<2>(T[0]), qbcast<2>(T[1]), qbcast<2>(T[2]), qbcast<2>(T[3]));
                        R1[K] = qsel(c, qbcast<3>(T[0]), qbcast<3>(T[1]), qbcast<3>(T[2]), qbcast<3>(T[3]));
                    }
                }
            }
            float y0, y1, s00, s01, s11;
            if ((r0 & 3) == 0) {
                y0 = qbcast<0>(x[I0]);
                y1 = qbcast<1>(x[I0]);
                s00 = qbcast<0>(R0[I0]);
                s01 = qbcast<1>(R0[I0]);
                s11 = qbcast<1>(R1[I0]);
            } else {
                y0 = qbcast<2>(x[I0]);
                y1 = qbcast<3>(x[I0]);
                s00 = qbcast<2>(R0[I0]);
                s01 = qbcast<3>(R0[I0]);
                s11 = qbcast<3>(R1[I0]);
            }
            // the FP32 sums; the FP64 path rebuilds them from the same state values
            const float py0 = y0, py1 = y1, ps00 = s00, ps01 = s01, ps11 = s11;
            y0 += BHs[2 * b];
            y1 += BHs[2 * b + 1];
            s00 += Rs[3 * b];
            s01 += Rs[3 * b + 1];
            s11 += Rs[3 * b + 2];
            float w0[6], w1[6];
            float v0, v1, m00, m01, m11;
            // likelihood and the update parameters (joint_measure of pf_kernels.hpp), the
            // candidates split over the quad
            auto measure = [&](const auto &wh, float logl) {
                const float LOG2PI = 1.8378770664093453f;
                if (!clutter) {
                    wh(cz[0], cz[1], v0, v1);
                    m00 = 1.f;
                    m01 = 0.f;
                    m11 = 1.f;
                    acc += -0.5f * fmaf(v0, v0, v1 * v1) - logl - LOG2PI;
                    return;
                }
                const float lnorm = -logl - LOG2PI;
                float se = 0.f, sv0 = 0.f, sv1 = 0.f, q00 = 0.f, q01 = 0.f, q11 = 0.f;
                for (int cc = c; cc < cnt; cc += 4) {
                    float u0, u1;
                    wh(cz[2 * cc], cz[2 * cc + 1], u0, u1);
                    const float d2 = fmaf(u0, u0, u1 * u1);
                    if (d2 <= a.gate) {
                        const float e = a.pdl * __expf(lnorm - 0.5f * d2);
                        se += e;
                        sv0 = fmaf(e, u0, sv0);
                        sv1 = fmaf(e, u1, sv1);
                        q00 = fmaf(e * u0, u0, q00);
                        q01 = fmaf(e * u0, u1, q01);
                        q11 = fmaf(e * u1, u1, q11);
                    }
                }
                se = qsum(se);
                sv0 = qsum(sv0);
                sv1 = qsum(sv1);
                q00 = qsum(q00);
                q01 = qsum(q01);
                q11 = qsum(q11);
                const float ell = a.b0 + se;
                const float inv = 1.f / ell;
                v0 = sv0 * inv;
                v1 = sv1 * inv;
                const float bs = se * inv;
                m00 = bs - (q00 * inv - v0 * v0);
                m01 = -(q01 * inv - v0 * v1);
                m11 = bs - (q11 * inv - v1 * v1);
                acc += logf(ell);
            };
            if (!innovation_ill(s00, s01, s11)) {         // quad-uniform
                const float l00 = sqrtf(s00);
                const float il00 = 1.f / l00;
                const float l10 = s01 * il00;
                const float l11 = sqrtf(s11 - l10 * l10);
                const float il11 = 1.f / l11;
#pragma unroll
                for (int K = 0; K < 6; ++K) {
                    w0[K] = R0[K] * il00;
                    w1[K] = (R1[K] - l10 * w0[K]) * il11;
                }
                measure(Whiten32{y0, y1, il00, l10, il11}, logf(l00 * l11));
            } else {
                // FP64 factor, W and whitening (joint_factor64 of pf_kernels.hpp)
                pflags |= kStatusIllCond;
                const double *jb = J64 + 5 * b;
                const double s00d = (double)ps00 + jb[2], s01d = (double)ps01 + jb[3], s11d = (double)ps11 + jb[4];
                const double d11 = s00d > 0.0 ? s11d - s01d * s01d / s00d : 0.0;
                if (!(s00d > 0.0) || !(d11 > 0.0)) {
                    pflags |= kStatusNotSpd;
#pragma unroll
                    for (int K = 0; K < 6; ++K) w0[K] = w1[K] = 0.f;
                    v0 = v1 = m00 = m01 = m11 = 0.f;
                    acc = -INFINITY;
                } else {
                    const double l00 = sqrt(s00d), l11 = sqrt(d11);
                    Whiten64 wh;
                    wh.y0 = (double)py0 + jb[0];
                    wh.y1 = (double)py1 + jb[1];
                    wh.il00 = 1.0 / l00;
                    wh.l10 = s01d * wh.il00;
                    wh.il11 = 1.0 / l11;
#pragma unroll
                    for (int K = 0; K < 6; ++K) {
                        const double a0 = (double)R0[K] * wh.il00;
                        w0[K] = (float)a0;
                        w1[K] = (float)(((double)R1[K] - wh.l10 * a0) * wh.il11);
                    }
                    measure(wh, (float)log(l00 * l11));
                }
            }
            // x += W^T v; P -= W^T M W = l1 z1 z1^T + l2 z2 z2^T
#pragma unroll
            for (int K = 0; K < 6; ++K) x[K] = fmaf(w0[K], v0, fmaf(w1[K], v1, x[K]));
            const float h = 0.5f * (m00 - m11), mid = 0.5f * (m00 + m11);
            const float rr = sqrtf(fmaf(h, h, m01 * m01));
            const float l1 = mid + rr, l2 = mid - rr;
            float e0 = 1.f, e1 = 0.f;
            if (rr > 0.f) {
                const float a0 = h >= 0.f ? h + rr : m01, a1 = h >= 0.f ? m01 : rr - h;
                const float inv = 1.f / sqrtf(fmaf(a0, a0, a1 * a1));
                e0 = a0 * inv;
                e1 = a1 * inv;
            }
            const float sq1 = sqrtf(fabsf(l1)), sq2 = sqrtf(fabsf(l2));
            float b1[6], b2[6], a1v[6], a2v[6];
#pragma unroll
            for (int K = 0; K < 6; ++K) {
                b1[K] = sq1 * fmaf(w0[K], e0, w1[K] * e1);
                b2[K] = sq2 * fmaf(w1[K], e0, -w0[K] * e1);
                a1v[K] = l1 >= 0.f ? -b1[K] : b1[K];
                a2v[K] = l2 >= 0.f ? -b2[K] : b2[K];
            }
#pragma unroll
            for (int I = 0; I < 6; ++I)
#pragma unroll
                for (int K = I; K < 6; ++K) Pt[qtile(I, K)] = qrank1<MFMA>(a1v[I], b1[K], Pt[qtile(I, K)]);
#pragma unroll
            for (int I = 0; I < 6; ++I)
#pragma unroll
                for (int K = I; K < 6; ++K) Pt[qtile(I, K)] = qrank1<MFMA>(a2v[I], b2[K], Pt[qtile(I, K)]);
        }

        if (live) {
            const unsigned vo = (unsigned)(sd.pofs + j) * 4u;
#pragma unroll
            for (int K = 0; K < 6; ++K) soa_st_partial(wx, vo + crow, rowb * (unsigned)(4 * K), x[K]);
#pragma unroll
            for (int I = 0; I < 6; ++I)
#pragma unroll
                for (int K = I; K < 6; ++K)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        if (K > I)
                            soa_st_partial(wP, vo + crow, rowb * (unsigned)pk(24, 4 * I + r, 4 * K), Pt[qtile(I, K)][r]);
                        else if (r <= c)
                            soa_st_partial(wP, vo + (unsigned)drow[I][r] * rowb, 0, Pt[qtile(I, K)][r]);
                    }
            if (c == 0) {
                a.lw[sd.pofs + j] = acc;
                if (comp_out) comp_out[j] = k;
            }
        }
        report_flags(live ? pflags : 0u, &a.status[s]);
    }
}

}  // namespace batch
}  // namespace mkfbt
