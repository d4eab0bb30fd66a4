// pf_adjoint.hip -- hand-written adjoint of the batched RBPF's
// log-marginal-likelihood (C ABI pf_batch_lml_grad in include/pf_batch.h).
//
// Objective of sequence s: F = sum_t lml_t, lml_t = log((1/N) sum_j exp(lw_tj)),
// with resampling a stop-gradient (the ancestors and component draws of the
// forward run are held fixed) and log pi_k - stopgrad(log pi_k) carried in
// every log-weight (value 0; its gradient is the score of the draws).
// Parameters: GMM means, logit weights, reversion rates r = 1 - gamma and the
// Cholesky factors of the component covariances (Q_k = (1 - g^2) L L^T).
//
// Forward: the batched step with components and ancestors recorded and the
// input state of every C-th frame checkpointed.  Backward, per segment of C
// frames, last segment first: replay the segment's states from its
// checkpoint, then one reverse kernel per frame, last frame first.  Per
// particle the reverse kernel recomputes its frame from the ancestor's state
// (recording the 2-row joint inputs), then propagates the adjoint of the
// updated state back through the joint updates (reverse mode over the
// d-dimensional state and, per joint, over its 5 local inputs y, S:
// pf_adjoint_local.hpp) and the mean-reverting predict.  Adjoints of a frame's inputs flow to
// the parents by the recorded ancestors: ancestors are sorted, so each
// parent's children are a contiguous range and the scatter is a deterministic
// gather.  Per-component gradient terms are reduced per wave (by component),
// per block in LDS and per 256-particle tile in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "pf_adjoint_local.hpp"
#include "pf_batch_impl.hpp"

using namespace mkfbt;
using namespace mkfbt::batch;

namespace {

constexpr int D = kBD;
constexpr int NP = npacked(kBD);
constexpr int kGV = D + NP + 2;       // per-component terms: mu' (D), Sigma' packed (NP), gamma, sum of wbar
constexpr int kMaxAdjK = 64;
#ifndef PF_ADJ_TILE
#define PF_ADJ_TILE 128
#endif
constexpr int kAdjTile = PF_ADJ_TILE;    // particles (threads) per tile of the adjoint kernels

// Inputs of one joint recorded on the way forward.
struct JointRec {
    float c0[D], c1[D];
    float y0, y1, s00, s01, s11;
};

struct AdjArgs {
    BatchArgs b;
    const float *in_x, *in_P;          // updated states of frame t-1 (checkpoint or segment slot)
    float *out_x, *out_P;              // replay: updated states of frame t
    const int *anc_prev;               // ancestors of frame t-1's resampling (null: frame 0, identity)
    const float *adj_in_x, *adj_in_P;  // adjoint of frame t+1's inputs, per particle
    float *adj_out_x, *adj_out_P;      // adjoint of frame t's inputs, per particle
    int *first;                        // children of particle j: [first[f + j], first[f + j + 1]), f = pofs + i
    const int *anc_t;                  // ancestors of frame t's resampling
    const float *mus, *sigs;           // state-basis mu' and packed Sigma' per sequence (K x D, K x NP)
    const int *t_seq, *t_j0;           // 256-particle tiles, active sequences first
    const int *bins;                   // initial components (reset)
    double *tacc;                      // per tile, K x kGV
    int t, t0;
    int kc;                            // components per group of the deterministic term reduction
};

// LDS floats of the term reduction's wave slots for groups of kc components
// (nwaves x kc x kGV values, then nwaves 64-bit masks).
__host__ __device__ constexpr int term_slot_floats(int kc, int nwaves)
{
    return ((nwaves * kc * kGV + 1) & ~1) + 2 * nwaves;
}

// One reduce-scatter step inside the rows of 16 lanes: the lanes with the
// mask bit clear keep the first of each pair (lo[i], hi[i]) and the others
// the second, each adding the value its DPP partner (a lane of the other
// half) holds for it.
template <int CTRL>
__device__ __forceinline__ float rs_step(bool upper, float lo, float hi)
{
    const float keep = upper ? hi : lo, send = upper ? lo : hi;
    return keep + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(send), CTRL, 0xF, 0xF, true));
}

// The wave sums of the kGV = 92 terms of the lanes with sel, into dst[0..91],
// by a fixed reduce-scatter butterfly: the two lane halves by
// v_permlane32_swap (46 pairs), the rows 0/1 and 2/3 by v_permlane16_swap
// (23 pairs; row r then holds terms 23 r .. 23 r + 22), then inside each row
// of 16 by DPP row_mirror, row_half_mirror, quad_perm [2,3,0,1] and
// quad_perm [1,0,3,2].  Lane l of row r ends with the sums of terms
// 23 r + o and 23 r + o + 1, o = 12 b3 + 6 b2 + 3 b1 + 2 b0 from the bits of
// l mod 16 (the second only for b0 = 0; o = 23 is padding).  The tree is
// fixed by the lane positions, so every run adds in the same order.
__device__ __forceinline__ void wave_sums_scatter(bool sel, const float (&term)[kGV], float *dst)
{
    static_assert(kGV == 4 * 23, "the butterfly splits the terms as 4 rows of 23");
    const int lane = threadIdx.x & 63, row = lane >> 4;
    const bool b3 = (lane & 8) != 0, b2 = (lane & 4) != 0, b1 = (lane & 2) != 0, b0 = (lane & 1) != 0;
    float u[46];
#pragma unroll
    for (int i = 0; i < 46; ++i) {
        const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(sel ? term[i] : 0.f),
                                                        __float_as_uint(sel ? term[i + 46] : 0.f), false, false);
        u[i] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
    float w[24];
#pragma unroll
    for (int i = 0; i < 23; ++i) {
        const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(u[i]), __float_as_uint(u[i + 23]), false,
                                                        false);
        w[i] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
    w[23] = 0.f;
    float x[12], y[6], z[4], q[2];
#pragma unroll
    for (int i = 0; i < 12; ++i) x[i] = rs_step<0x140>(b3, w[i], w[i + 12]);
#pragma unroll
    for (int i = 0; i < 6; ++i) y[i] = rs_step<0x141>(b2, x[i], x[i + 6]);
#pragma unroll
    for (int i = 0; i < 3; ++i) z[i] = rs_step<0x4E>(b1, y[i], y[i + 3]);
    z[3] = 0.f;
#pragma unroll
    for (int i = 0; i < 2; ++i) q[i] = rs_step<0xB1>(b0, z[i], z[i + 2]);
    const int o = 12 * b3 + 6 * b2 + 3 * b1 + 2 * b0;
    if (o < 23) dst[23 * row + o] = q[0];
    if (!b0) dst[23 * row + o + 1] = q[1];
}

// The per-particle component terms of a tile added to its FP64 accumulator
// tacc (K x kGV) in an order fixed by the tile alone, so gradients are
// bit-reproducible: components in groups of kc; per distinct component of
// the group in a wave, the wave's sums (fixed butterfly) go to that wave's
// LDS slot; after a barrier the slots are added in wave order, in FP64, to
// tacc.  Every thread of the block calls it.
__device__ __forceinline__ void reduce_terms(bool live, int k, const float (&term)[kGV], float *slots, double *tacc,
                                             int K, int kc)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
    unsigned long long *wmask = reinterpret_cast<unsigned long long *>(slots + ((nw * kc * kGV + 1) & ~1));
    for (int g0 = 0; g0 < K; g0 += kc) {
        const int g1 = min(K, g0 + kc);
        unsigned long long rem = __ballot(live && k >= g0 && k < g1), mine = 0;
        while (rem) {
            const int leader = __ffsll((long long)rem) - 1;
            const int kk = __shfl(k, leader, 64);
            const bool sel = live && k == kk;
            rem &= ~__ballot(sel);
            wave_sums_scatter(sel, term, slots + ((size_t)wid * kc + (kk - g0)) * kGV);
            mine |= 1ull << (kk - g0);
        }
        if (lane == 0) wmask[wid] = mine;
        __syncthreads();
        for (int i = threadIdx.x; i < (g1 - g0) * kGV; i += blockDim.x) {
            const int c = i / kGV;
            double sum = 0.0;
            bool any = false;
            for (int w = 0; w < nw; ++w)
                if ((wmask[w] >> c) & 1ull) {
                    sum += (double)slots[((size_t)w * kc + c) * kGV + (i - c * kGV)];
                    any = true;
                }
            if (any) tacc[(size_t)g0 * kGV + i] += sum;
        }
        __syncthreads();
    }
}

__device__ __forceinline__ void load_state(const float *xs, const float *Ps, int ld, int col, float (&x)[D],
                                           float (&P)[NP])
{
    const unsigned rowb = (unsigned)ld * 4u;
    const __amdgpu_buffer_rsrc_t rx = soa_rsrc(xs, rowb * D), rP = soa_rsrc(Ps, rowb * NP);
    const unsigned off = (unsigned)col * 4u;
#pragma unroll
    for (int i = 0; i < D; ++i) x[i] = soa_ld(rx, off, rowb * i);
#pragma unroll
    for (int p = 0; p < NP; ++p) P[p] = soa_ld(rP, off, rowb * p);
}

__device__ __forceinline__ void store_state(float *xs, float *Ps, int ld, int col, const float (&x)[D],
                                            const float (&P)[NP])
{
    const unsigned rowb = (unsigned)ld * 4u;
    const __amdgpu_buffer_rsrc_t rx = soa_rsrc(xs, rowb * D), rP = soa_rsrc(Ps, rowb * NP);
    const unsigned off = (unsigned)col * 4u;
#pragma unroll
    for (int i = 0; i < D; ++i) soa_st(rx, off, rowb * i, x[i]);
#pragma unroll
    for (int p = 0; p < NP; ++p) soa_st(rP, off, rowb * p, P[p]);
}

// Replay of frame t (recorded ancestors, same component draw and update as
// the forward): writes the updated states of frame t into the segment slot.
// Three waves per SIMD: it then holds ~150 registers without spills (177
// when the compiler may use two waves' budget), and it shares the GPU with
// the reverse kernels on the other stream.
template <int SELJ>
__global__ void __launch_bounds__(kAdjTile) __attribute__((amdgpu_waves_per_eu(3, 8))) k_adj_replay(AdjArgs a)
{
    extern __shared__ float4 sm4[];
    const BatchArgs &ba = a.b;
    const int s = a.t_seq[blockIdx.x];
    const int j0 = a.t_j0[blockIdx.x];
    const SeqDesc sd = ba.seqs[s];
    const StepArgs v = seq_view(ba, sd);
    const float *fcand = v.cand + (size_t)a.t * ba.J * ba.cap * 2;
    CandPrefetch cpf;
    cand_prefetch(cpf, v.count + (size_t)a.t * ba.J, fcand, ba.J, ba.cap);
    stage_table(sm4, reinterpret_cast<const float4 *>(ba.tables + (size_t)sd.table_ofs), ba.table_floats >> 2);
    cand_store(cpf, reinterpret_cast<float *>(sm4) + ba.table_floats, fcand, ba.J, ba.cap);
    __syncthreads();
    const int j = j0 + (int)threadIdx.x;
    if (j >= sd.N || a.t >= sd.T) return;
    const int src = a.anc_prev ? a.anc_prev[sd.pofs + j] : j;
    float x[D], P[NP];
    load_state(a.in_x, a.in_P, ba.ld, sd.pofs + src, x, P);
    const uint32_t U1 = ba.uni[(size_t)(sd.tofs + a.t) * 2];
    unsigned flags = 0;   // the recorded forward run reported them
    (void)particle_update<D, kModeCand, SELJ>(ba.base, a.t, j, U1, reinterpret_cast<const float *>(sm4), x, P,
                                              sd.pi_total, seq_comp_global(ba, sd), nullptr, flags);
    store_state(a.out_x, a.out_P, ba.ld, sd.pofs + j, x, P);
}

// Children ranges of frame t's resampling: first[f + i] = first child of
// parent i = lower_bound(anc, i) (ancestors are non-decreasing), i in [0, N].
__global__ void __launch_bounds__(kAdjTile) k_adj_first(AdjArgs a)
{
    const BatchArgs &ba = a.b;
    const int s = a.t_seq[blockIdx.x];
    const int j0 = a.t_j0[blockIdx.x];
    const SeqDesc sd = ba.seqs[s];
    const int i = j0 + (int)threadIdx.x;
    if (a.t + 1 >= sd.T) return;
    const int *anc = a.anc_t + sd.pofs;
    const int f = sd.pofs + s;
    auto lower = [&](int key) {
        int lo = 0, hi = sd.N;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (anc[mid] < key) lo = mid + 1;
            else hi = mid;
        }
        return lo;
    };
    if (i < sd.N) a.first[f + i] = lower(i);
    if (i == sd.N - 1) a.first[f + sd.N] = sd.N;
}

// Reverse kernel of frame t.
template <int SELJ>
__global__ void __launch_bounds__(kAdjTile) k_adj_step(AdjArgs a)
{
    extern __shared__ float4 sm4[];
    const BatchArgs &ba = a.b;
    const int K = ba.K;
    const int tile = blockIdx.x;
    const int s = a.t_seq[tile];
    const int j0 = a.t_j0[tile];
    const SeqDesc sd = ba.seqs[s];
    float *sm = reinterpret_cast<float *>(sm4);
    float *smu = sm + ba.stage_floats;
    float *ssg = smu + K * D;
    float *slots = ssg + K * NP;
    const StepArgs v = seq_view(ba, sd);
    const float *fcand = v.cand + (size_t)a.t * ba.J * ba.cap * 2;
    CandPrefetch cpf;
    cand_prefetch(cpf, v.count + (size_t)a.t * ba.J, fcand, ba.J, ba.cap);
    stage_table(sm4, reinterpret_cast<const float4 *>(ba.tables + (size_t)sd.table_ofs), ba.table_floats >> 2);
    cand_store(cpf, sm + ba.table_floats, fcand, ba.J, ba.cap);
    for (int i = threadIdx.x; i < K * D; i += blockDim.x) smu[i] = a.mus[(size_t)s * K * D + i];
    for (int i = threadIdx.x; i < K * NP; i += blockDim.x) ssg[i] = a.sigs[(size_t)s * K * NP + i];
    __syncthreads();

    const int t = a.t;
    const int j = j0 + (int)threadIdx.x;
    const bool live = j < sd.N && t < sd.T;
    float term[kGV];
#pragma unroll
    for (int i = 0; i < kGV; ++i) term[i] = 0.f;
    int k = 0;
    if (live) {
        const int src = a.anc_prev ? a.anc_prev[sd.pofs + j] : j;
        float x[D], P[NP];
        load_state(a.in_x, a.in_P, ba.ld, sd.pofs + src, x, P);
        // ---- forward recompute, recording each joint's inputs
        constexpr int CS = 2 + D + NP;
        const unsigned long long *sthr = reinterpret_cast<const unsigned long long *>(sm);
        const float *Hs = sm + 4 * K;
        const float *BHs = Hs + ba.m * D;
        const float *Rs = BHs + ba.m;
        const float *stab = v.comp_global ? v.comp_global : sm + image_head_floats(K, ba.m, D, ba.J);
        const uint32_t U1 = ba.uni[(size_t)(sd.tofs + t) * 2];
        k = draw_component(sthr, K, (uint32_t)j, U1, sd.pi_total);
        const float *ck = stab + k * CS;
        const float g = ck[0], g2 = ck[1];
#pragma unroll
        for (int i = 0; i < D; ++i) x[i] = fmaf(g, x[i], ck[2 + i]);
#pragma unroll
        for (int p = 0; p < NP; ++p) P[p] = fmaf(g2, P[p], ck[2 + D + p]);
        JointRec rec[SELJ];
        bool on[SELJ];
        float lw = 0.f;
#pragma unroll
        for (int oi = 0; oi < SELJ; ++oi) {
            const int b = (int)((ba.order_packed >> (4 * oi)) & 15u);
            int cnt;
            const float *cz;
            bool clutter;
            on[oi] = joint_inputs<kModeCand>(v, t, b, cnt, cz, clutter, sm);
            if (!on[oi]) continue;
            const int r0 = 2 * oi, r1 = r0 + 1;
            if (oi > 0) {            // the first joint's inputs are recomputed from the ancestor's state
#pragma unroll
                for (int i = 0; i < D; ++i) {
                    rec[oi].c0[i] = P[pk(D, r0, i)];
                    rec[oi].c1[i] = P[pk(D, r1, i)];
                }
                rec[oi].y0 = x[r0] + BHs[2 * b];
                rec[oi].y1 = x[r1] + BHs[2 * b + 1];
                rec[oi].s00 = P[pk(D, r0, r0)] + Rs[3 * b];
                rec[oi].s01 = P[pk(D, r0, r1)] + Rs[3 * b + 1];
                rec[oi].s11 = P[pk(D, r1, r1)] + Rs[3 * b + 2];
            }
            float w0[D], w1[D], y0, y1, il00, l10, il11, logl;
            Joint<D>::prepare_sel(x, P, r0, BHs[2 * b], BHs[2 * b + 1], Rs[3 * b], Rs[3 * b + 1], Rs[3 * b + 2], w0,
                                  w1, y0, y1, il00, l10, il11, logl);
            float v0, v1, m00, m01, m11;
            lw += joint_measure<kModeCand>(v, t, j, b, clutter, cnt, cz, y0, y1, il00, l10, il11, logl, v0, v1,
                                           m00, m01, m11);
            Joint<D>::apply(x, P, w0, w1, v0, v1, m00, m01, m11);
        }
        // ---- d F / d lw_tj
        const double lml = ba.lml[sd.tofs + t];
        const float wbar = isfinite(lml) && isfinite(lw) ? (float)exp((double)lw - lml - log((double)sd.N)) : 0.f;
        // ---- adjoint of the updated state: sum over the children at frame t + 1
        float xb[D], Pb[NP];
#pragma unroll
        for (int i = 0; i < D; ++i) xb[i] = 0.f;
#pragma unroll
        for (int p = 0; p < NP; ++p) Pb[p] = 0.f;
        if (t + 1 < sd.T) {
            const int f = sd.pofs + s;
            const int c0 = a.first[f + j], c1 = a.first[f + j + 1];
            for (int c = c0; c < c1; ++c) {
                float cx[D], cP[NP];
                load_state(a.adj_in_x, a.adj_in_P, ba.ld, sd.pofs + c, cx, cP);
#pragma unroll
                for (int i = 0; i < D; ++i) xb[i] += cx[i];
#pragma unroll
                for (int p = 0; p < NP; ++p) Pb[p] += cP[p];
            }
        }
        // ---- joints in reverse order
#pragma unroll
        for (int oi = SELJ - 1; oi >= 0; --oi) {
            if (!on[oi]) continue;
            const int b = (int)((ba.order_packed >> (4 * oi)) & 15u);
            int cnt;
            const float *cz;
            bool clutter;
            (void)joint_inputs<kModeCand>(v, t, b, cnt, cz, clutter, sm);
            const int r0 = 2 * oi, r1 = r0 + 1;
            JointRec R0;
            if (oi == 0) {           // rows r0, r1 of the predicted state, as the forward computed them
                const unsigned rowb = (unsigned)ba.ld * 4u;
                const __amdgpu_buffer_rsrc_t rx = soa_rsrc(a.in_x, rowb * D), rP = soa_rsrc(a.in_P, rowb * NP);
                const unsigned off = (unsigned)(sd.pofs + src) * 4u;
#pragma unroll
                for (int i = 0; i < D; ++i) {
                    R0.c0[i] = fmaf(g2, soa_ld(rP, off, rowb * pk(D, r0, i)), ck[2 + D + pk(D, r0, i)]);
                    R0.c1[i] = fmaf(g2, soa_ld(rP, off, rowb * pk(D, r1, i)), ck[2 + D + pk(D, r1, i)]);
                }
                R0.y0 = fmaf(g, soa_ld(rx, off, rowb * r0), ck[2 + r0]) + BHs[2 * b];
                R0.y1 = fmaf(g, soa_ld(rx, off, rowb * r1), ck[2 + r1]) + BHs[2 * b + 1];
                R0.s00 = R0.c0[r0] + Rs[3 * b];
                R0.s01 = R0.c0[r1] + Rs[3 * b + 1];
                R0.s11 = R0.c1[r1] + Rs[3 * b + 2];
            }
            const JointRec &R = (oi == 0) ? R0 : rec[oi];
            // GC = G+ C with G+ the symmetric gradient of the updated P
            // (packed off-diagonal adjoints count both entries).
            float gc0[D], gc1[D];
#pragma unroll
            for (int i = 0; i < D; ++i) {
                float a0 = 0.f, a1 = 0.f;
#pragma unroll
                for (int q = 0; q < D; ++q) {
                    const float gq = (i == q) ? Pb[pk(D, i, q)] : 0.5f * Pb[pk(D, i, q)];
                    a0 = fmaf(gq, R.c0[q], a0);
                    a1 = fmaf(gq, R.c1[q], a1);
                }
                gc0[i] = a0;
                gc1[i] = a1;
            }
            float gb0 = 0.f, gb1 = 0.f, nb00 = 0.f, nb01 = 0.f, nb11 = 0.f;
#pragma unroll
            for (int i = 0; i < D; ++i) {
                gb0 = fmaf(R.c0[i], xb[i], gb0);
                gb1 = fmaf(R.c1[i], xb[i], gb1);
                nb00 = fmaf(-R.c0[i], gc0[i], nb00);
                nb01 = fmaf(-R.c0[i], gc1[i], nb01);
                nb11 = fmaf(-R.c1[i], gc1[i], nb11);
            }
            const adj::LocalRev L = adj::local_joint_rev(adj::PdafConst{v.gate, v.pdl, v.b0}, R.y0, R.y1, R.s00,
                                                         R.s01, R.s11, clutter, cnt, cz, gb0, gb1, nb00, nb01, nb11,
                                                         wbar);
            const float din[5] = {L.yb0, L.yb1, L.sb00, L.sb01, L.sb11};
            // C-bar = xbar g^T - 2 GC N, into the packed entries of P read by C
#pragma unroll
            for (int i = 0; i < D; ++i) {
                const float cb0 = xb[i] * L.g0 - 2.f * (gc0[i] * L.n00 + gc1[i] * L.n01);
                const float cb1 = xb[i] * L.g1 - 2.f * (gc0[i] * L.n01 + gc1[i] * L.n11);
                Pb[pk(D, i, r0)] += cb0;
                Pb[pk(D, i, r1)] += cb1;
            }
            xb[r0] += din[0];
            xb[r1] += din[1];
            Pb[pk(D, r0, r0)] += din[2];
            Pb[pk(D, r0, r1)] += din[3];
            Pb[pk(D, r1, r1)] += din[4];
        }
        // ---- predict x = g x + (1 - g) mu', P = g^2 P + (1 - g^2) Sigma'
        const float *mu = smu + k * D, *sg = ssg + k * NP;
        const unsigned rowb = (unsigned)ba.ld * 4u;
        const __amdgpu_buffer_rsrc_t rx = soa_rsrc(a.in_x, rowb * D), rP = soa_rsrc(a.in_P, rowb * NP);
        const unsigned off = (unsigned)(sd.pofs + src) * 4u;
        float gbar = 0.f;
#pragma unroll
        for (int i = 0; i < D; ++i) {
            gbar = fmaf(xb[i], soa_ld(rx, off, rowb * i) - mu[i], gbar);
            term[i] = (1.f - g) * xb[i];
            xb[i] *= g;
        }
        float gpp = 0.f;
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            gpp = fmaf(Pb[p], soa_ld(rP, off, rowb * p) - sg[p], gpp);
            term[D + p] = (1.f - g2) * Pb[p];
            Pb[p] *= g2;
        }
        term[D + NP] = fmaf(2.f * g, gpp, gbar);
        term[D + NP + 1] = wbar;
        store_state(a.adj_out_x, a.adj_out_P, ba.ld, sd.pofs + j, xb, Pb);
    }
    reduce_terms(live, k, term, slots, a.tacc + (size_t)tile * K * kGV, K, a.kc);
}

// Initial states x_0j = mu'_{bins_j}, P_0j = Sigma'_{bins_j}: the adjoint of
// frame 0's inputs goes to the components of the reset.
__global__ void __launch_bounds__(kAdjTile) k_adj_reset(AdjArgs a)
{
    extern __shared__ float4 sm4[];
    const BatchArgs &ba = a.b;
    const int K = ba.K;
    const int tile = blockIdx.x;
    const int s = a.t_seq[tile];
    const int j0 = a.t_j0[tile];
    const SeqDesc sd = ba.seqs[s];
    float *slots = reinterpret_cast<float *>(sm4);
    const int j = j0 + (int)threadIdx.x;
    const bool live = j < sd.N;
    float term[kGV];
#pragma unroll
    for (int i = 0; i < kGV; ++i) term[i] = 0.f;
    int k = 0;
    if (live) {
        k = a.bins[sd.pofs + j];
        float xb[D], Pb[NP];
        load_state(a.adj_in_x, a.adj_in_P, ba.ld, sd.pofs + j, xb, Pb);
#pragma unroll
        for (int i = 0; i < D; ++i) term[i] = xb[i];
#pragma unroll
        for (int p = 0; p < NP; ++p) term[D + p] = Pb[p];
    }
    reduce_terms(live, k, term, slots, a.tacc + (size_t)tile * K * kGV, K, a.kc);
}

// Per-sequence sums of the tile accumulators (fixed order, FP64).
__global__ void k_adj_sum(const double *tacc, const int *t_first, int S, int K, double *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int per = K * kGV;
    if (i >= S * per) return;
    const int s = i / per, e = i % per;
    double r = 0.0;
    for (int tl = t_first[s]; tl < t_first[s + 1]; ++tl) r += tacc[(size_t)tl * per + e];
    out[i] = r;
}

__global__ void k_fill_f64(double *p, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// ------------------------------------------------------------------- host
template <typename T>
int ensure(T **p, size_t n, size_t &have)
{
    if (*p && have >= n) return PF_OK;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    have = 0;
    HIP_TRYB(hipMalloc((void **)p, sizeof(T) * std::max<size_t>(n, 1)));
    have = n;
    return PF_OK;
}

int setup(pf_batch *b, int C)
{
    std::vector<size_t> &h = b->adj_have;
    const size_t ld = (size_t)b->ld, xw = (size_t)D * ld, Pw = (size_t)NP * ld;
    b->ckpt_every = C;
    b->n_ckpt = (b->Tmax + C - 1) / C;
    b->seg_slots = std::max(1, C - 1);
    int st;
    if ((st = ensure(&b->rec_comp, (size_t)b->Tmax * ld, h[0]))) return st;
    if ((st = ensure(&b->rec_anc, (size_t)b->Tmax * ld, h[1]))) return st;
    if ((st = ensure(&b->ck_x, (size_t)b->n_ckpt * xw, h[2]))) return st;
    if ((st = ensure(&b->ck_P, (size_t)b->n_ckpt * Pw, h[3]))) return st;
    if ((st = ensure(&b->seg_x[0], (size_t)b->seg_slots * xw, h[4]))) return st;
    if ((st = ensure(&b->seg_P[0], (size_t)b->seg_slots * Pw, h[5]))) return st;
    if ((st = ensure(&b->seg_x[1], (size_t)b->seg_slots * xw, h[16]))) return st;
    if ((st = ensure(&b->seg_P[1], (size_t)b->seg_slots * Pw, h[17]))) return st;
    if ((st = ensure(&b->adj_x[0], xw, h[6]))) return st;
    if ((st = ensure(&b->adj_x[1], xw, h[7]))) return st;
    if ((st = ensure(&b->adj_P[0], Pw, h[8]))) return st;
    if ((st = ensure(&b->adj_P[1], Pw, h[9]))) return st;
    if ((st = ensure(&b->first, 2 * (size_t)C * (ld + b->S + 1), h[10]))) return st;
    if (!b->stream2) HIP_TRYB(hipStreamCreateWithFlags(&b->stream2, hipStreamNonBlocking));
    if (!b->ev_fwd) HIP_TRYB(hipEventCreateWithFlags(&b->ev_fwd, hipEventDisableTiming));
    while ((int)b->ev_replay.size() < b->n_ckpt) {
        hipEvent_t e1, e2;
        HIP_TRYB(hipEventCreateWithFlags(&e1, hipEventDisableTiming));
        HIP_TRYB(hipEventCreateWithFlags(&e2, hipEventDisableTiming));
        b->ev_replay.push_back(e1);
        b->ev_back.push_back(e2);
    }
    // 256-particle tiles, sequences in length order (active ones first)
    std::vector<int> ts, tj, tf(b->S + 1);
    for (int i = 0; i < b->S; ++i) {
        tf[i] = (int)ts.size();
        for (int j0 = 0; j0 < b->h_seqs[i].N; j0 += kAdjTile) {
            ts.push_back(i);
            tj.push_back(j0);
        }
    }
    tf[b->S] = (int)ts.size();
    b->t256_total = (int)ts.size();
    b->h_t256_count.assign((size_t)b->Tmax, 0);
    for (int t = 0; t < b->Tmax; ++t) {
        int act = 0;
        while (act < b->S && b->h_seqs[act].T > t) ++act;
        b->h_t256_count[t] = tf[act];
    }
    if ((st = ensure(&b->t256_seq, ts.size(), h[11]))) return st;
    if ((st = ensure(&b->t256_j0, tj.size(), h[12]))) return st;
    if ((st = ensure(&b->t256_first, tf.size(), h[13]))) return st;
    HIP_TRYB(hipMemcpy(b->t256_seq, ts.data(), sizeof(int) * ts.size(), hipMemcpyHostToDevice));
    HIP_TRYB(hipMemcpy(b->t256_j0, tj.data(), sizeof(int) * tj.size(), hipMemcpyHostToDevice));
    HIP_TRYB(hipMemcpy(b->t256_first, tf.data(), sizeof(int) * tf.size(), hipMemcpyHostToDevice));
    if ((st = ensure(&b->tacc, (size_t)b->t256_total * b->K * kGV, h[14]))) return st;
    if ((st = ensure(&b->gsum, (size_t)b->S * b->K * kGV, h[15]))) return st;
    return PF_OK;
}

template <int SELJ>
void launch_adj_t(const AdjArgs &a, int grid, size_t lds_replay, size_t lds_step, bool replay, hipStream_t st)
{
    if (replay)
        hipLaunchKernelGGL((k_adj_replay<SELJ>), dim3(grid), dim3(kAdjTile), lds_replay, st, a);
    else
        hipLaunchKernelGGL((k_adj_step<SELJ>), dim3(grid), dim3(kAdjTile), lds_step, st, a);
}

int launch_adj(pf_batch *b, const AdjArgs &a, int grid, bool replay, hipStream_t st)
{
    if (grid <= 0) return PF_OK;
    const size_t lds_replay = 4 * (size_t)b->stage_floats();
    const size_t lds_step = 4 * ((size_t)b->stage_floats() + (size_t)b->K * (D + NP) +
                                 (size_t)term_slot_floats(a.kc, kAdjTile / 64));
    switch (b->J) {
    case 1: launch_adj_t<1>(a, grid, lds_replay, lds_step, replay, st); break;
    case 2: launch_adj_t<2>(a, grid, lds_replay, lds_step, replay, st); break;
    case 3: launch_adj_t<3>(a, grid, lds_replay, lds_step, replay, st); break;
    case 4: launch_adj_t<4>(a, grid, lds_replay, lds_step, replay, st); break;
    case 5: launch_adj_t<5>(a, grid, lds_replay, lds_step, replay, st); break;
    case 6: launch_adj_t<6>(a, grid, lds_replay, lds_step, replay, st); break;
    default: return set_error(PF_ERR_UNSUPPORTED, "adjoint: %d joints", b->J);
    }
    HIP_TRYB(hipGetLastError());
    return PF_OK;
}

// Gauss elimination Cholesky (FP64, lower) for the host-side chain rule.
bool chol_lower(int n, const double *A, double *L)
{
    for (int i = 0; i < n * n; ++i) L[i] = 0.0;
    for (int j = 0; j < n; ++j) {
        double s = A[j * n + j];
        for (int k = 0; k < j; ++k) s -= L[j * n + k] * L[j * n + k];
        if (!(s > 0.0)) return false;
        L[j * n + j] = std::sqrt(s);
        for (int i = j + 1; i < n; ++i) {
            double r = A[i * n + j];
            for (int k = 0; k < j; ++k) r -= L[i * n + k] * L[j * n + k];
            L[i * n + j] = r / L[j * n + j];
        }
    }
    return true;
}

}  // namespace

extern "C" {

int pf_batch_lml_grad(pf_batch *b, const int32_t *count, const float *cand, const uint32_t *uniforms, double *lml,
                      int32_t checkpoint_every, double *grad_means, double *grad_logits, double *grad_rates,
                      double *grad_chol)
{
    int st = bcheck(b);
    if (st) return st;
    if (!b->tracks) return set_error(PF_ERR_STATE, "reset every sequence before a run");
    if (!count || !cand || !uniforms || !lml) return set_error(PF_ERR_ARG, "null run argument");
    if (checkpoint_every < 1) return set_error(PF_ERR_ARG, "checkpoint_every must be >= 1");
    if (b->K > kMaxAdjK) return set_error(PF_ERR_UNSUPPORTED, "adjoint supports K <= %d", kMaxAdjK);
    if (b->D != D) return set_error(PF_ERR_UNSUPPORTED, "adjoint is compiled for d = %d", D);
    for (int s = 0; s < b->S; ++s)
        if (!b->sel_ok[s]) return set_error(PF_ERR_UNSUPPORTED, "adjoint needs a full-rank measurement matrix");
    if ((size_t)step_lds_bytes(b->stage_floats(), b->nb_max) > (size_t)kMaxLdsBytes)
        return set_error(PF_ERR_UNSUPPORTED, "model image and resampler workspace exceed the LDS budget");
    const int C = std::min<int>(checkpoint_every, b->Tmax);
    if ((st = setup(b, C))) return st;
    b->tracks = false;
    std::fill(b->fresh.begin(), b->fresh.end(), 0);
    const size_t ld = (size_t)b->ld, xw = (size_t)D * ld, Pw = (size_t)NP * ld;

    hipEvent_t e0, e1, e2;
    HIP_TRYB(hipEventCreate(&e0));
    HIP_TRYB(hipEventCreate(&e1));
    HIP_TRYB(hipEventCreate(&e2));
    // ---- forward, recorded
    HIP_TRYB(hipEventRecord(e0, b->stream));
    if ((st = run_all(b, count, cand, uniforms, lml, b->Tmax, false, true))) return st;
    {
        BatchArgs fa = batch_args(b, count, cand, uniforms, lml, 0, 1);
        fa.rec_anc = b->rec_anc;
        if ((st = launch_batch_finish(b, fa))) return st;
    }
    HIP_TRYB(hipEventRecord(e1, b->stream));
    {
        // The reverse sweep differentiates the FP32 factorization of every
        // joint; a forward run that factored an innovation in FP64 (cond > ~1e6)
        // or met a non-SPD one has no gradient here.
        std::vector<unsigned> w((size_t)b->S);
        HIP_TRYB(hipMemcpy(w.data(), b->status, sizeof(unsigned) * b->S, hipMemcpyDeviceToHost));
        unsigned word = 0;
        for (unsigned v : w) word |= v;
        if (word & (kStatusIllCond | kStatusNotSpd)) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            (void)hipEventDestroy(e2);
            return set_error(PF_ERR_UNSUPPORTED, "the forward run factored an innovation covariance in FP64 (cond > "
                                                 "1e6) or met a non-SPD one; the adjoint differentiates the FP32 "
                                                 "factorization only");
        }
    }
    // ---- backward
    const size_t nt = (size_t)b->t256_total * b->K * kGV;
    hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, b->stream, b->tacc, nt);
    AdjArgs a{};
    a.b = batch_args(b, count, cand, uniforms, lml, 0, 0);
    a.mus = b->mus;
    a.sigs = b->sigs;
    a.t_seq = b->t256_seq;
    a.t_j0 = b->t256_j0;
    a.bins = b->rec_bins;
    a.tacc = b->tacc;
    {
        // components per reduction group: the step kernel's LDS within 64 KB
        const long avail = 16384L - (long)b->stage_floats() - (long)b->K * (D + NP) - 2 * (kAdjTile / 64) - 1;
        a.kc = (int)std::max(1L, std::min<long>({(long)b->K, 64L, avail / ((kAdjTile / 64) * kGV)}));
    }
    // Segments, last first.  The replay of segment s-1 (and its child ranges)
    // runs on the second stream while the reverse kernels of segment s run on
    // the first; segment buffers alternate by segment parity.
    hipStream_t M = b->stream, R = b->stream2;
    const size_t fst = ld + (size_t)b->S + 1;
    HIP_TRYB(hipEventRecord(b->ev_fwd, M));
    HIP_TRYB(hipStreamWaitEvent(R, b->ev_fwd, 0));
    auto seg_input_x = [&](int seg, int t, int t0) -> const float * {
        return t == t0 ? b->ck_x + (size_t)seg * xw : b->seg_x[seg & 1] + (size_t)(t - 1 - t0) * xw;
    };
    auto seg_input_P = [&](int seg, int t, int t0) -> const float * {
        return t == t0 ? b->ck_P + (size_t)seg * Pw : b->seg_P[seg & 1] + (size_t)(t - 1 - t0) * Pw;
    };
    // states = false: the recorded forward already wrote the segment's states
    // (the last segment), only its child ranges are computed.
    auto replay_seg = [&](int seg, bool states) -> int {
        const int t0 = seg * C, t1 = std::min(t0 + C, b->Tmax);
        AdjArgs r = a;
        for (int t = t0; states && t + 1 < t1; ++t) {   // states of frames t0 .. t1-2
            r.t = t;
            r.t0 = t0;
            r.in_x = seg_input_x(seg, t, t0);
            r.in_P = seg_input_P(seg, t, t0);
            r.out_x = b->seg_x[seg & 1] + (size_t)(t - t0) * xw;
            r.out_P = b->seg_P[seg & 1] + (size_t)(t - t0) * Pw;
            r.anc_prev = t > 0 ? b->rec_anc + (size_t)(t - 1) * ld : nullptr;
            const int rs = launch_adj(b, r, b->h_t256_count[t], true, R);
            if (rs) return rs;
        }
        for (int t = t0; t < t1; ++t) {              // child ranges of every frame's resampling
            const int grid = b->h_t256_count[t];
            if (grid <= 0) continue;
            r.t = t;
            r.anc_t = b->rec_anc + (size_t)t * ld;
            r.first = b->first + ((size_t)(seg & 1) * C + (t - t0)) * fst;
            hipLaunchKernelGGL(k_adj_first, dim3(grid), dim3(kAdjTile), 0, R, r);
            HIP_TRYB(hipGetLastError());
        }
        HIP_TRYB(hipEventRecord(b->ev_replay[seg], R));
        return PF_OK;
    };
    auto backward_seg = [&](int seg) -> int {
        const int t0 = seg * C, t1 = std::min(t0 + C, b->Tmax);
        HIP_TRYB(hipStreamWaitEvent(M, b->ev_replay[seg], 0));
        for (int t = t1 - 1; t >= t0; --t) {
            a.t = t;
            a.t0 = t0;
            a.in_x = seg_input_x(seg, t, t0);
            a.in_P = seg_input_P(seg, t, t0);
            a.anc_prev = t > 0 ? b->rec_anc + (size_t)(t - 1) * ld : nullptr;
            a.first = b->first + ((size_t)(seg & 1) * C + (t - t0)) * fst;
            a.adj_in_x = b->adj_x[(t + 1) & 1];
            a.adj_in_P = b->adj_P[(t + 1) & 1];
            a.adj_out_x = b->adj_x[t & 1];
            a.adj_out_P = b->adj_P[t & 1];
            const int rs = launch_adj(b, a, b->h_t256_count[t], false, M);
            if (rs) return rs;
        }
        HIP_TRYB(hipEventRecord(b->ev_back[seg], M));
        return PF_OK;
    };
    const int last = b->n_ckpt - 1;
    if ((st = replay_seg(last, false))) return st;
    for (int seg = last; seg >= 0; --seg) {
        if (seg > 0) {
            if (seg + 1 <= last) HIP_TRYB(hipStreamWaitEvent(R, b->ev_back[seg + 1], 0));
            if ((st = replay_seg(seg - 1, true))) return st;
        }
        if ((st = backward_seg(seg))) return st;
    }
    a.adj_in_x = b->adj_x[0];
    a.adj_in_P = b->adj_P[0];
    const size_t lds_reset = 4 * (size_t)term_slot_floats(a.kc, kAdjTile / 64);
    hipLaunchKernelGGL(k_adj_reset, dim3(b->t256_total), dim3(kAdjTile), lds_reset, b->stream, a);
    HIP_TRYB(hipGetLastError());
    const int ng = b->S * b->K * kGV;
    hipLaunchKernelGGL(k_adj_sum, dim3((ng + 255) / 256), dim3(256), 0, b->stream, b->tacc, b->t256_first, b->S,
                       b->K, b->gsum);
    HIP_TRYB(hipGetLastError());
    HIP_TRYB(hipEventRecord(e2, b->stream));
    std::vector<double> g((size_t)ng);
    HIP_TRYB(hipMemcpyAsync(g.data(), b->gsum, sizeof(double) * g.size(), hipMemcpyDeviceToHost, b->stream));
    HIP_TRYB(hipStreamSynchronize(b->stream));
    float f_ms = 0.f, b_ms = 0.f;
    HIP_TRYB(hipEventElapsedTime(&f_ms, e0, e1));
    HIP_TRYB(hipEventElapsedTime(&b_ms, e1, e2));
    b->last_fwd_ms = f_ms;
    b->last_bwd_ms = b_ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
    b->recorded = true;
    b->rec_frames = b->Tmax;

    // ---- chain rule to the parameters (FP64, model basis), per user sequence
    const int K = b->K;
    for (int s = 0; s < b->S; ++s) {
        const int i = b->rank[s];
        const HostModel &hm = b->hm[s];
        const double *T = b->h_T.data() + (size_t)i * D * D;
        const double *gi = g.data() + (size_t)i * K * kGV;
        double wsum = 0.0, asum = 0.0;
        for (int k = 0; k < K; ++k) {
            wsum += hm.pi[k];
            asum += gi[k * kGV + D + NP + 1];
        }
        for (int k = 0; k < K; ++k) {
            const double *gk = gi + (size_t)k * kGV;
            if (grad_means)
                for (int c = 0; c < D; ++c) {             // mu = T^-1 mu'  ->  mubar = T^T mubar'
                    double r = 0.0;
                    for (int q = 0; q < D; ++q) r += T[q * D + c] * gk[q];
                    grad_means[((size_t)s * K + k) * D + c] = r;
                }
            if (grad_logits) grad_logits[(size_t)s * K + k] = gk[D + NP + 1] - hm.pi[k] / wsum * asum;
            if (grad_rates) grad_rates[(size_t)s * K + k] = -gk[D + NP];
            if (grad_chol) {
                double Gp[D * D], G[D * D], TG[D * D], L[D * D];
                for (int r = 0; r < D; ++r)
                    for (int c = 0; c < D; ++c) {
                        const double v = gk[D + pk(D, std::min(r, c), std::max(r, c))];
                        Gp[r * D + c] = (r == c) ? v : 0.5 * v;
                    }
                for (int r = 0; r < D; ++r)                 // G = T^T G' T
                    for (int c = 0; c < D; ++c) {
                        double acc = 0.0;
                        for (int q = 0; q < D; ++q) acc += T[q * D + r] * Gp[q * D + c];
                        TG[r * D + c] = acc;
                    }
                for (int r = 0; r < D; ++r)
                    for (int c = 0; c < D; ++c) {
                        double acc = 0.0;
                        for (int q = 0; q < D; ++q) acc += TG[r * D + q] * T[q * D + c];
                        G[r * D + c] = acc;
                    }
                if (!chol_lower(D, hm.sig.data() + (size_t)k * D * D, L))
                    return set_error(PF_ERR_NOT_SPD, "sequence %d component %d: covariance not SPD", s, k);
                double *out = grad_chol + ((size_t)s * K + k) * D * D;
                for (int r = 0; r < D; ++r)                 // Lbar = 2 G L, lower triangle
                    for (int c = 0; c < D; ++c) {
                        double acc = 0.0;
                        if (c <= r)
                            for (int q = 0; q < D; ++q) acc += G[r * D + q] * L[q * D + c];
                        out[r * D + c] = 2.0 * acc;
                    }
            }
        }
    }
    return PF_OK;
}

int pf_batch_get_choices(pf_batch *b, int32_t s, int32_t *comp, int32_t *anc)
{
    int st = bcheck(b);
    if (st) return st;
    if (!b->recorded) return set_error(PF_ERR_STATE, "no recorded run (call pf_batch_lml_grad first)");
    if (s < 0 || s >= b->S) return set_error(PF_ERR_ARG, "sequence out of range");
    const SeqDesc &sd = b->h_seqs[b->rank[s]];
    for (int t = 0; t < sd.T; ++t) {
        const size_t off = (size_t)t * b->ld + sd.pofs;
        if (comp) HIP_TRYB(hipMemcpy(comp + (size_t)t * sd.N, b->rec_comp + off, sizeof(int) * sd.N,
                                     hipMemcpyDeviceToHost));
        if (anc) HIP_TRYB(hipMemcpy(anc + (size_t)t * sd.N, b->rec_anc + off, sizeof(int) * sd.N,
                                    hipMemcpyDeviceToHost));
    }
    return PF_OK;
}

int pf_batch_last_adjoint_ms(pf_batch *b, double *forward_ms, double *backward_ms)
{
    if (!b) return set_error(PF_ERR_ARG, "null batch");
    if (forward_ms) *forward_ms = b->last_fwd_ms;
    if (backward_ms) *backward_ms = b->last_bwd_ms;
    return PF_OK;
}

}  // extern "C"
