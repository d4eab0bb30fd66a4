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
// d-dimensional state, forward-mode duals over each joint's 5 local inputs
// y, S) and the mean-reverting predict.  Adjoints of a frame's inputs flow to
// the parents by the recorded ancestors: ancestors are sorted, so each
// parent's children are a contiguous range and the scatter is a deterministic
// gather.  Per-component gradient terms are reduced per wave (by component),
// per block in LDS and per 256-particle tile in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "pf_batch_impl.hpp"

using namespace mkfbt;
using namespace mkfbt::batch;

namespace {

constexpr int D = kBD;
constexpr int NP = npacked(kBD);
constexpr int kGV = D + NP + 2;       // per-component terms: mu' (D), Sigma' packed (NP), gamma, sum of wbar
constexpr int kMaxAdjK = 64;
constexpr int kAdjTile = 256;

// ------------------------------------------------------------ dual numbers
// Value and derivatives along the 5 local inputs of a joint (y0, y1, S00,
// S01, S11), FP32.
struct D5 {
    float v, d[5];
};
__device__ __forceinline__ D5 dc(float v)
{
    D5 r;
    r.v = v;
#pragma unroll
    for (int i = 0; i < 5; ++i) r.d[i] = 0.f;
    return r;
}
__device__ __forceinline__ D5 dseed(float v, int i)
{
    D5 r = dc(v);
    r.d[i] = 1.f;
    return r;
}
__device__ __forceinline__ D5 operator+(const D5 &a, const D5 &b)
{
    D5 r;
    r.v = a.v + b.v;
#pragma unroll
    for (int i = 0; i < 5; ++i) r.d[i] = a.d[i] + b.d[i];
    return r;
}
__device__ __forceinline__ D5 operator-(const D5 &a, const D5 &b)
{
    D5 r;
    r.v = a.v - b.v;
#pragma unroll
    for (int i = 0; i < 5; ++i) r.d[i] = a.d[i] - b.d[i];
    return r;
}
__device__ __forceinline__ D5 operator*(const D5 &a, const D5 &b)
{
    D5 r;
    r.v = a.v * b.v;
#pragma unroll
    for (int i = 0; i < 5; ++i) r.d[i] = fmaf(a.d[i], b.v, a.v * b.d[i]);
    return r;
}
__device__ __forceinline__ D5 operator*(float s, const D5 &a)
{
    D5 r;
    r.v = s * a.v;
#pragma unroll
    for (int i = 0; i < 5; ++i) r.d[i] = s * a.d[i];
    return r;
}
__device__ __forceinline__ D5 operator+(float s, const D5 &a)
{
    D5 r = a;
    r.v = s + a.v;
    return r;
}
__device__ __forceinline__ D5 operator-(float s, const D5 &a)
{
    D5 r;
    r.v = s - a.v;
#pragma unroll
    for (int i = 0; i < 5; ++i) r.d[i] = -a.d[i];
    return r;
}
__device__ __forceinline__ D5 drcp(const D5 &a)
{
    D5 r;
    r.v = 1.f / a.v;
    const float g = -r.v * r.v;
#pragma unroll
    for (int i = 0; i < 5; ++i) r.d[i] = g * a.d[i];
    return r;
}
__device__ __forceinline__ D5 dsqrt(const D5 &a)
{
    D5 r;
    r.v = sqrtf(a.v);
    const float g = 0.5f / r.v;
#pragma unroll
    for (int i = 0; i < 5; ++i) r.d[i] = g * a.d[i];
    return r;
}
__device__ __forceinline__ D5 dlog(const D5 &a)
{
    D5 r;
    r.v = logf(a.v);
    const float g = 1.f / a.v;
#pragma unroll
    for (int i = 0; i < 5; ++i) r.d[i] = g * a.d[i];
    return r;
}
__device__ __forceinline__ D5 dexp(const D5 &a)
{
    D5 r;
    r.v = __expf(a.v);
#pragma unroll
    for (int i = 0; i < 5; ++i) r.d[i] = r.v * a.d[i];
    return r;
}

// The local function of one joint: from (y, S) to the update x += C g,
// P -= C N C^T (C = P[:, rows]) and the log-likelihood term ell, as the
// forward computes them (Joint::prepare_sel + joint_measure + apply, with
// g = L^-T v and N = L^-T M L^-1).
struct Local {
    D5 g0, g1, n00, n01, n11, ell;
};

__device__ __forceinline__ Local local_joint(const StepArgs &v, float y0f, float y1f, float s00f, float s01f,
                                             float s11f, bool clutter, int cnt, const float *cz)
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
    Local o;
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
        if (d2.v <= v.gate) {                               // gating is a discrete choice: held fixed
            const D5 e = v.pdl * dexp(lnorm - 0.5f * d2);
            se = se + e;
            sv0 = sv0 + e * u0;
            sv1 = sv1 + e * u1;
            q00 = q00 + (e * u0) * u0;
            q01 = q01 + (e * u0) * u1;
            q11 = q11 + (e * u1) * u1;
        }
    }
    const D5 ells = v.b0 + se;
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
};

// Wave-then-block reduction of the per-particle component terms into the
// tile's FP64 accumulator.  acc is the block's LDS accumulator (K x kGV,
// zeroed by the caller), mask the set of components present.
__device__ __forceinline__ void reduce_terms(bool live, int k, const float (&term)[kGV], float *acc,
                                             unsigned long long *mask)
{
    const int lane = threadIdx.x & 63;
    unsigned long long rem = __ballot(live);
    while (rem) {
        const int leader = __ffsll((long long)rem) - 1;
        const int kk = __shfl(k, leader, 64);
        const bool sel = live && k == kk;
        rem &= ~__ballot(sel);
        float keep0 = 0.f, keep1 = 0.f;
#pragma unroll
        for (int v = 0; v < kGV; ++v) {
            const float s = wave_sum(sel ? term[v] : 0.f);
            if ((v & 63) == lane) {
                if (v < 64) keep0 = s;
                else keep1 = s;
            }
        }
        atomicAdd(acc + kk * kGV + lane, keep0);
        if (lane < kGV - 64) atomicAdd(acc + kk * kGV + 64 + lane, keep1);
        if (lane == 0) atomicOr(mask, 1ull << kk);
    }
}

__device__ __forceinline__ void flush_terms(const float *acc, const unsigned long long *mask, double *tacc, int K)
{
    const unsigned long long m = *mask;
    for (int i = threadIdx.x; i < K * kGV; i += blockDim.x)
        if ((m >> (i / kGV)) & 1ull) tacc[i] += (double)acc[i];
}

__device__ __forceinline__ void clear_terms(float *acc, unsigned long long *mask, int K)
{
    for (int i = threadIdx.x; i < K * kGV; i += blockDim.x) acc[i] = 0.f;
    if (threadIdx.x == 0) *mask = 0ull;
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
template <int SELJ>
__global__ void __launch_bounds__(256) k_adj_replay(AdjArgs a)
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
    (void)particle_update<D, kModeCand, SELJ>(v, a.t, j, U1, reinterpret_cast<const float *>(sm4), x, P);
    store_state(a.out_x, a.out_P, ba.ld, sd.pofs + j, x, P);
}

// Children ranges of frame t's resampling: first[f + i] = first child of
// parent i = lower_bound(anc, i) (ancestors are non-decreasing), i in [0, N].
__global__ void __launch_bounds__(256) k_adj_first(AdjArgs a)
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
__global__ void __launch_bounds__(256) k_adj_step(AdjArgs a)
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
    float *acc = ssg + K * NP;
    unsigned long long *mask = reinterpret_cast<unsigned long long *>(acc + ((K * kGV + 1) & ~1));
    const StepArgs v = seq_view(ba, sd);
    const float *fcand = v.cand + (size_t)a.t * ba.J * ba.cap * 2;
    CandPrefetch cpf;
    cand_prefetch(cpf, v.count + (size_t)a.t * ba.J, fcand, ba.J, ba.cap);
    stage_table(sm4, reinterpret_cast<const float4 *>(ba.tables + (size_t)sd.table_ofs), ba.table_floats >> 2);
    cand_store(cpf, sm + ba.table_floats, fcand, ba.J, ba.cap);
    for (int i = threadIdx.x; i < K * D; i += blockDim.x) smu[i] = a.mus[(size_t)s * K * D + i];
    for (int i = threadIdx.x; i < K * NP; i += blockDim.x) ssg[i] = a.sigs[(size_t)s * K * NP + i];
    clear_terms(acc, mask, K);
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
        const float *stab = sm + 4 * K;
        const uint32_t U1 = ba.uni[(size_t)(sd.tofs + t) * 2];
        k = draw_component(sthr, K, (uint32_t)j, U1, sd.pi_total);
        const float *ck = stab + k * CS;
        const float g = ck[0], g2 = ck[1];
#pragma unroll
        for (int i = 0; i < D; ++i) x[i] = fmaf(g, x[i], ck[2 + i]);
#pragma unroll
        for (int p = 0; p < NP; ++p) P[p] = fmaf(g2, P[p], ck[2 + D + p]);
        const float *Hs = stab + K * CS;
        const float *BHs = Hs + ba.m * D;
        const float *Rs = BHs + ba.m;
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
            const JointRec &R = rec[oi];
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
            const Local L = local_joint(v, R.y0, R.y1, R.s00, R.s01, R.s11, clutter, cnt, cz);
            float din[5];
#pragma unroll
            for (int q = 0; q < 5; ++q)
                din[q] = gb0 * L.g0.d[q] + gb1 * L.g1.d[q] + nb00 * L.n00.d[q] + 2.f * nb01 * L.n01.d[q] +
                         nb11 * L.n11.d[q] + wbar * L.ell.d[q];
            // C-bar = xbar g^T - 2 GC N, into the packed entries of P read by C
#pragma unroll
            for (int i = 0; i < D; ++i) {
                const float cb0 = xb[i] * L.g0.v - 2.f * (gc0[i] * L.n00.v + gc1[i] * L.n01.v);
                const float cb1 = xb[i] * L.g1.v - 2.f * (gc0[i] * L.n01.v + gc1[i] * L.n11.v);
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
        float xp[D], Pp[NP];
        load_state(a.in_x, a.in_P, ba.ld, sd.pofs + src, xp, Pp);
        const float *mu = smu + k * D, *sg = ssg + k * NP;
        float gbar = 0.f;
#pragma unroll
        for (int i = 0; i < D; ++i) {
            gbar = fmaf(xb[i], xp[i] - mu[i], gbar);
            term[i] = (1.f - g) * xb[i];
            xb[i] *= g;
        }
        float gpp = 0.f;
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            gpp = fmaf(Pb[p], Pp[p] - sg[p], gpp);
            term[D + p] = (1.f - g2) * Pb[p];
            Pb[p] *= g2;
        }
        term[D + NP] = fmaf(2.f * g, gpp, gbar);
        term[D + NP + 1] = wbar;
        store_state(a.adj_out_x, a.adj_out_P, ba.ld, sd.pofs + j, xb, Pb);
    }
    reduce_terms(live, k, term, acc, mask);
    __syncthreads();
    flush_terms(acc, mask, a.tacc + (size_t)tile * K * kGV, K);
}

// Initial states x_0j = mu'_{bins_j}, P_0j = Sigma'_{bins_j}: the adjoint of
// frame 0's inputs goes to the components of the reset.
__global__ void __launch_bounds__(256) k_adj_reset(AdjArgs a)
{
    extern __shared__ float4 sm4[];
    const BatchArgs &ba = a.b;
    const int K = ba.K;
    const int tile = blockIdx.x;
    const int s = a.t_seq[tile];
    const int j0 = a.t_j0[tile];
    const SeqDesc sd = ba.seqs[s];
    float *acc = reinterpret_cast<float *>(sm4);
    unsigned long long *mask = reinterpret_cast<unsigned long long *>(acc + ((K * kGV + 1) & ~1));
    clear_terms(acc, mask, K);
    __syncthreads();
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
    reduce_terms(live, k, term, acc, mask);
    __syncthreads();
    flush_terms(acc, mask, a.tacc + (size_t)tile * K * kGV, K);
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
    if ((st = ensure(&b->seg_x, (size_t)b->seg_slots * xw, h[4]))) return st;
    if ((st = ensure(&b->seg_P, (size_t)b->seg_slots * Pw, h[5]))) return st;
    if ((st = ensure(&b->adj_x[0], xw, h[6]))) return st;
    if ((st = ensure(&b->adj_x[1], xw, h[7]))) return st;
    if ((st = ensure(&b->adj_P[0], Pw, h[8]))) return st;
    if ((st = ensure(&b->adj_P[1], Pw, h[9]))) return st;
    if ((st = ensure(&b->first, ld + b->S + 1, h[10]))) return st;
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
void launch_adj_t(pf_batch *b, const AdjArgs &a, int grid, size_t lds_replay, size_t lds_step, bool replay)
{
    if (replay)
        hipLaunchKernelGGL((k_adj_replay<SELJ>), dim3(grid), dim3(kAdjTile), lds_replay, b->stream, a);
    else
        hipLaunchKernelGGL((k_adj_step<SELJ>), dim3(grid), dim3(kAdjTile), lds_step, b->stream, a);
}

int launch_adj(pf_batch *b, const AdjArgs &a, int grid, bool replay)
{
    if (grid <= 0) return PF_OK;
    const size_t lds_replay = 4 * (size_t)b->stage_floats();
    const size_t lds_step = 4 * ((size_t)b->stage_floats() + (size_t)b->K * (D + NP) + (((size_t)b->K * kGV + 1) & ~1)) + 8;
    switch (b->J) {
    case 1: launch_adj_t<1>(b, a, grid, lds_replay, lds_step, replay); break;
    case 2: launch_adj_t<2>(b, a, grid, lds_replay, lds_step, replay); break;
    case 3: launch_adj_t<3>(b, a, grid, lds_replay, lds_step, replay); break;
    case 4: launch_adj_t<4>(b, a, grid, lds_replay, lds_step, replay); break;
    case 5: launch_adj_t<5>(b, a, grid, lds_replay, lds_step, replay); break;
    case 6: launch_adj_t<6>(b, a, grid, lds_replay, lds_step, replay); break;
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
    a.first = b->first;
    for (int seg = b->n_ckpt - 1; seg >= 0; --seg) {
        const int t0 = seg * C, t1 = std::min(t0 + C, b->Tmax);
        const float *ckx = b->ck_x + (size_t)seg * xw, *ckP = b->ck_P + (size_t)seg * Pw;
        auto input_x = [&](int t) { return t == t0 ? ckx : b->seg_x + (size_t)(t - 1 - t0) * xw; };
        auto input_P = [&](int t) { return t == t0 ? ckP : b->seg_P + (size_t)(t - 1 - t0) * Pw; };
        for (int t = t0; t + 1 < t1; ++t) {          // states of frames t0 .. t1-2
            a.t = t;
            a.t0 = t0;
            a.in_x = input_x(t);
            a.in_P = input_P(t);
            a.out_x = b->seg_x + (size_t)(t - t0) * xw;
            a.out_P = b->seg_P + (size_t)(t - t0) * Pw;
            a.anc_prev = t > 0 ? b->rec_anc + (size_t)(t - 1) * ld : nullptr;
            if ((st = launch_adj(b, a, b->h_t256_count[t], true))) return st;
        }
        for (int t = t1 - 1; t >= t0; --t) {
            a.t = t;
            a.t0 = t0;
            a.in_x = input_x(t);
            a.in_P = input_P(t);
            a.anc_prev = t > 0 ? b->rec_anc + (size_t)(t - 1) * ld : nullptr;
            a.anc_t = b->rec_anc + (size_t)t * ld;
            a.adj_in_x = b->adj_x[(t + 1) & 1];
            a.adj_in_P = b->adj_P[(t + 1) & 1];
            a.adj_out_x = b->adj_x[t & 1];
            a.adj_out_P = b->adj_P[t & 1];
            const int grid = b->h_t256_count[t];
            if (grid > 0) {
                hipLaunchKernelGGL(k_adj_first, dim3(grid), dim3(kAdjTile), 0, b->stream, a);
                HIP_TRYB(hipGetLastError());
            }
            if ((st = launch_adj(b, a, grid, false))) return st;
        }
    }
    a.adj_in_x = b->adj_x[0];
    a.adj_in_P = b->adj_P[0];
    const size_t lds_reset = 4 * (((size_t)b->K * kGV + 1) & ~1) + 8;
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
