// pf_kernels.hpp -- gfx950 kernels of the Rao-Blackwellised particle filter
// (reference: src/pf2DRao.cpp, src/KF_model.cpp, src/my_gmm.cpp).
//
// Device layout (one filter):
//   state   x (D) and the packed upper triangle of the symmetric
//           covariance P (D(D+1)/2), FP32, either as per-particle records or
//           as rows x[D][ld], P[D(D+1)/2][ld] (Lay below; ld = particle
//           capacity).  Two copies: the step reads copy A at the ancestor
//           index and writes copy B at the particle index (the resampling
//           gather of pf2DRao.cpp:153-156 is folded into the next frame's
//           loads).
//   table   LDS image of the model: 2K words of component thresholds
//           C_pi[k] * N * 2^32 (128-bit), then per component k
//           [g, g^2, (1-g) mu_k (D), (1-g^2) Sigma_k packed], then the shared
//           H (m x D), BH (m) and the 2x2 joint blocks of R (r00, r01, r11).
//
// Frame pipeline (device-resident runs: one launch per frame):
//   k_step    resolves each particle's ancestor from the previous frame's
//             pending resampling (or reads a materialised ancestor array),
//             gathers, predicts, updates, writes the log-weight and, in
//             blocks that are the resampler's blocks, the frame's
//             block-scaled fixed-point weights, scans and exponent; block 0
//             records the resampling uniform and advances the frame counter.
//   k_weights the same weights as a separate launch (host-driven steps,
//             other callers of the filter).
// k_ancestors materialises the pending ancestors (end of a run, host steps,
// stand-alone resampling).
#pragma once

#include "pf_common.hpp"

namespace mkfbt {

// Ordered 64-bit key of a double (unsigned comparison = numeric order), for
// maxima taken with 64-bit atomicMax, and its inverse.
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

// Cross-lane primitives without LDS (every lane of the wave must be active).
// xor_lane<O>(v): the value of lane l ^ O, O = 1, 2, 4, 8, 16, 32: quad
// permutations, row rotation by 8 and row_half_mirror + quad reversal (DPP),
// and the row-pair / lane-half swaps of v_permlane16_swap and
// v_permlane32_swap.
template <int O>
__device__ __forceinline__ uint32_t xor_lane_u32(uint32_t v)
{
    static_assert(O == 1 || O == 2 || O == 4 || O == 8 || O == 16 || O == 32, "xor distance");
    const int x = (int)v;
    if constexpr (O == 1) return (uint32_t)__builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, false);
    if constexpr (O == 2) return (uint32_t)__builtin_amdgcn_update_dpp(0, x, 0x4E, 0xF, 0xF, false);
    if constexpr (O == 4) {
        const int y = __builtin_amdgcn_update_dpp(0, x, 0x141, 0xF, 0xF, false);     // l -> 7 - l in each 8
        return (uint32_t)__builtin_amdgcn_update_dpp(0, y, 0x1B, 0xF, 0xF, false);  // then reversed in each 4
    }
    if constexpr (O == 8) return (uint32_t)__builtin_amdgcn_update_dpp(0, x, 0x128, 0xF, 0xF, false);   // row_ror:8
    const int lane = (int)(threadIdx.x & 63);
    if constexpr (O == 16) {
        const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        return (lane & 16) ? r[0] : r[1];
    }
    const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return (lane & 32) ? r[0] : r[1];
}

template <int O, typename T>
__device__ __forceinline__ T xor_lane(T v)
{
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit values");
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, xor_lane_u32<O>(__builtin_bit_cast(uint32_t, v)));
    } else {
        const uint64_t u = __builtin_bit_cast(uint64_t, v);
        const uint64_t lo = xor_lane_u32<O>((uint32_t)u), hi = xor_lane_u32<O>((uint32_t)(u >> 32));
        return __builtin_bit_cast(T, (hi << 32) | lo);
    }
}

// x shifted by DPP control CTRL on the rows in ROWS (other rows and
// out-of-row sources give 0).
template <int CTRL, int ROWS>
__device__ __forceinline__ unsigned long long dpp_u64(unsigned long long x)
{
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)x, CTRL, ROWS, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(x >> 32), CTRL, ROWS, 0xF, true);
    return ((unsigned long long)(uint32_t)hi << 32) | (uint32_t)lo;
}

// Inclusive prefix sum over the wave (lane l: lanes 0 .. l): row_shr 1, 2,
// 4, 8 inside each row of 16, then row_bcast:15 and row_bcast:31 carry the
// rows' totals forward (integers: the same sums as any other order).
__device__ __forceinline__ unsigned long long wave_incl_scan(unsigned long long v)
{
    v += dpp_u64<0x111, 0xF>(v);
    v += dpp_u64<0x112, 0xF>(v);
    v += dpp_u64<0x114, 0xF>(v);
    v += dpp_u64<0x118, 0xF>(v);
    v += dpp_u64<0x142, 0xA>(v);
    v += dpp_u64<0x143, 0xC>(v);
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v)
{
    v = fmax(v, xor_lane<32>(v));
    v = fmax(v, xor_lane<16>(v));
    v = fmax(v, xor_lane<8>(v));
    v = fmax(v, xor_lane<4>(v));
    v = fmax(v, xor_lane<2>(v));
    v = fmax(v, xor_lane<1>(v));
    return v;
}

constexpr int kMaxJoints = 8;
constexpr int kWeightsBlock = 256;     // input block of the resampler scans
constexpr int kStageBlocks = 12;       // input blocks the ancestor search stages in LDS
constexpr int kMaxScanBlocks = 4096;   // => at most 2^20 resampled weights
constexpr int kMaxLdsBytes = 64 * 1024;  // per-workgroup LDS budget

enum StepMode { kModeRef = 0, kModeCand = 1 };
enum AncMode { kAncArray = 0, kAncFused = 1 };

struct FrameScratch {
    unsigned long long maxkey[2];  // per frame parity: ordered key of the largest finite weight (0 = none)
    unsigned int spare;
    unsigned int status;           // bit 0: a degenerate (all-zero) resampling happened
    uint32_t U_pending;            // uniform of the pending resampling
    uint32_t pad;
    int pow_max[3];                // frame t mod 3: largest block exponent of frame t's log-weights
                                   // (raised by the step kernels that write their own weights)
    int pad2;
};

// The pending systematic resampling of one weight vector of length L into N
// outputs: block-local inclusive prefix sums and block sums of q.
struct Pending {
    const unsigned long long *__restrict__ cloc;
    const unsigned long long *__restrict__ bsum;
    const double *__restrict__ bexp;
    const int *__restrict__ bpow = nullptr;      // block exponents of block-scaled log-weights (null: none)
    const int *pow_max = nullptr;                // their maximum, when its producer recorded it (else reduced)
    int L, nb;
    int bshift = 8;                              // log2 of the inputs per block (kWeightsBlock)
    int stage = kStageBlocks * kWeightsBlock;    // LDS elements of the staged search
};

// Inputs of a device-resident run, written to device memory once per run
// (k_run_begin), so that the frames captured into a graph do not depend on
// the run's buffers and one captured graph serves every run of the handle.
struct RunDesc {
    const float *Z;          // reference mode: Z[t][m][N]
    const int *count;        // candidate mode: count[t][J]
    const float *cand;       //                 cand[t][J][cap][2]
    const uint32_t *uni;     // uniforms[t][2]: U_comp, U_res
    double *lml;             // lml[t], the frame's log mean weight
    int flip;                // state buffer read by frame 0 (0: xin/Pin, 1: xout/Pout)
    int pad;
};

struct StepArgs {
    const float *__restrict__ xin;
    const float *__restrict__ Pin;
    float *__restrict__ xout;
    float *__restrict__ Pout;
    int *__restrict__ anc;                // read (kAncArray) or written (kAncFused)
    float *__restrict__ lw;
    int *__restrict__ comp_out;           // optional
    FrameScratch *fs;
    const float *__restrict__ table;
    const int *__restrict__ frame;
    const uint32_t *__restrict__ uni;     // 2 per frame: U_comp, U_res
    double *lml;                          // kAncFused: lml[t-1] of the consumed resampling
    Pending pend;                         // kAncFused: previous frame's weights
    // reference mode: Z[t][r][zld]
    const float *__restrict__ Z;
    long long z_frame_stride;
    // candidate mode: count[t][J], cand[t][J][cap][2]
    const int *__restrict__ count;
    const float *__restrict__ cand;
    unsigned long long pi_total;
    int table_floats;   // floats of the LDS model image (multiple of 4)
    int stage_floats;   // LDS floats before the resolver workspace: image + staged detections (candidate mode)
    const float *comp_global;  // component rows in global memory when the image head alone is staged, else null
    int K, N, ld, m, J, cap, zld;
    int parity;         // frame parity: selects maxkey[parity]
    int anc_mode;
    unsigned int order_packed;   // joint processing order, 4 bits per slot
    unsigned int clutter_mask;
    float pdl;    // P_D / lambda
    float b0;     // 1 - P_D * P_G
    float gate;   // chi-square gate
    // Optional: the frame's resampler weights written by the step itself
    // (blocks of kWeightsBlock threads = resampler blocks); block 0 also
    // records the pending uniform uni[2 t + 1] and writes t + 1 to frame_next.
    unsigned long long *w_cloc, *w_bsum;
    double *w_bexp;
    int *w_bpow;
    int *frame_next;
    // Device-resident runs (run != null): the frame t = *frame (block 0 writes
    // t + 1 to *frame_next) chooses everything else in the kernel: the state
    // buffer read is (t + run->flip) & 1 (xin/Pin = buffer 0, xout/Pout =
    // buffer 1) and the other is written; frame t writes its resampler
    // weights to set t & 1 (w_* = set 0, w1_* = set 1) and resolves frame
    // t - 1's from the other set (pend views set 1, pend1 set 0); frame 0
    // reads the ancestor array.  Inputs come from *run.
    const RunDesc *run;
    Pending pend1;
    unsigned long long *w1_cloc, *w1_bsum;
    double *w1_bexp;
    int *w1_bpow;
};

// The StepArgs of frame t of a device-resident run (see StepArgs::run), with
// t returned; outside runs a itself and t = *a.frame.
__device__ __forceinline__ StepArgs frame_args(const StepArgs &ain, int &t)
{
    StepArgs a = ain;
    t = *a.frame;
    if (!a.run) return a;
    const RunDesc rd = *a.run;
    if ((t + rd.flip) & 1) {
        const float *x0 = a.xin, *P0 = a.Pin;
        a.xin = a.xout;
        a.Pin = a.Pout;
        a.xout = const_cast<float *>(x0);
        a.Pout = const_cast<float *>(P0);
    }
    if (t & 1) {
        a.w_cloc = a.w1_cloc;
        a.w_bsum = a.w1_bsum;
        a.w_bexp = a.w1_bexp;
        a.w_bpow = a.w1_bpow;
        a.pend = a.pend1;
    }
    a.anc_mode = (t == 0) ? kAncArray : kAncFused;
    a.parity = t & 1;
    a.Z = rd.Z;
    a.count = rd.count;
    a.cand = rd.cand;
    a.uni = rd.uni;
    a.lml = rd.lml;
    return a;
}

// Start of a device-resident run: the run's inputs, frame 0 in the counter
// that even frames read, and the exponent slot of frame 0 cleared (the
// frames raise it).
static __global__ void k_run_begin(RunDesc *d, RunDesc v, int *ctr, FrameScratch *fs)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        *d = v;
        ctr[0] = 0;
        ctr[1] = 0;
        fs->pow_max[0] = kPowEmpty;
    }
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
    v += xor_lane<32>(v);      // the butterfly order xor 32, 16, ..., 1
    v += xor_lane<16>(v);
    v += xor_lane<8>(v);
    v += xor_lane<4>(v);
    v += xor_lane<2>(v);
    v += xor_lane<1>(v);
    return v;
}

// ------------------------------------------------------------- resampling
// Systematic resampling (pf2DRao.cpp:175-210) with fixed-point weights q_i:
//   linear weights: q_i = floor(2^32 * w_i / max w) (or a known bound),
//   log-weights: block-scaled, q_i = floor(2^32 exp(lw_i - e_b ln 2)) with the
//     block exponent e_b of pf_common.hpp, aligned by >> (E - e_b),
//   C = inclusive prefix sum of the aligned values (exact integers),
//   a_i = min{idx : C[idx] * N * 2^32 > (i * 2^32 + U) * C[L-1]}.
// The prefix sum is held as block-local scans (cloc) plus block sums (bsum)
// and exponents (bpow): C at input k of block b is pre_b + (cloc_k >> s_b),
// pre_b the sum of the aligned block sums before b.  resolve_ancestor()
// finds a_i for one block of outputs; C[L-1] = 0 (no positive weight) gives
// the uniform resampling.

// LDS workspace of resolve_ancestor: nb + 1 prefix words, the staged input
// elements (Pending::stage), 16 broadcast / wave slots, nb block shifts.
__host__ __device__ constexpr int resolve_lds_words(int nb, int stage = kStageBlocks * kWeightsBlock)
{
    return nb + 1 + stage + 16 + (nb + 1) / 2;
}

// A load through a pointer known to address global memory (a global rather
// than a flat instruction: it counts only against the vector-memory counter,
// so later waits need not cover it unless its value is used).
template <typename T>
__device__ __forceinline__ T gload(const T *p)
{
    return *(const __attribute__((address_space(1))) T *)p;
}

struct NoHook {
    __device__ void operator()() const {}
};

// Phase 1 of the ancestor resolution: the largest block exponent E, the
// exclusive prefix of the aligned block sums in LDS (pre[0..nb], pre[nb] =
// total) with each block's shift, and, if want_es, the FP64 sum of the block
// weight sums scaled to 2^E in es.  Every thread of the block must call it
// (want_es block-uniform).  hook() runs once the block sums' loads are
// issued: a caller's own loads placed there share their latency.
template <typename Hook = NoHook>
__device__ __forceinline__ unsigned long long resolve_prefix(const Pending &pd, unsigned long long *w, bool want_es,
                                                             double &es_out, int &pow_out, Hook hook = Hook())
{
    const int tid = threadIdx.x, nt = blockDim.x, nb = pd.nb;
    const int lane = tid & 63, wid = tid >> 6, nwaves = (nt + 63) >> 6;
    unsigned long long *pre = w;
    unsigned long long *stg = w + nb + 1;
    unsigned long long *misc = stg + pd.stage;
    int *shf = reinterpret_cast<int *>(misc + 16);
    const int per = (nb + nt - 1) / nt;
    const int c0 = min(nb, tid * per), c1 = min(nb, c0 + per);
    // A thread's blocks c0 .. c1 - 1; up to 8 of them stay in registers from
    // one round of loads (more are loaded again in each pass).
    constexpr int kReg = 8;
    unsigned long long rs[kReg];
    int re[kReg];
#pragma unroll
    for (int u = 0; u < kReg; ++u) {
        rs[u] = (c0 + u < c1) ? pd.bsum[c0 + u] : 0ull;
        re[u] = (pd.bpow && c0 + u < c1) ? pd.bpow[c0 + u] : kPowEmpty;
    }
    // 0. E, the largest block exponent (0 without exponents)
    int E = 0;
    if (pd.bpow && pd.pow_max) E = gload(pd.pow_max);
    hook();
    if (pd.bpow && !pd.pow_max) {
        int em = kPowEmpty;
#pragma unroll
        for (int u = 0; u < kReg; ++u) em = max(em, re[u]);
        for (int k = c0 + kReg; k < c1; ++k) em = max(em, pd.bpow[k]);
        em = max(em, xor_lane<32>(em));
        em = max(em, xor_lane<16>(em));
        em = max(em, xor_lane<8>(em));
        em = max(em, xor_lane<4>(em));
        em = max(em, xor_lane<2>(em));
        em = max(em, xor_lane<1>(em));
        __syncthreads();                 // LDS of a previous use is free
        if (lane == 0) misc[8 + wid] = (unsigned long long)(long long)em;
        __syncthreads();
        E = kPowEmpty;
        for (int k = 0; k < nwaves; ++k) E = max(E, (int)(long long)misc[8 + k]);
    }
    auto shift_of = [&](int e) { return pd.bpow ? pow_shift(E, e) : 0; };
    // 1. the aligned block sums of this thread's blocks
    unsigned long long loc = 0;
#pragma unroll
    for (int u = 0; u < kReg; ++u)
        if (c0 + u < c1) loc += shr64(rs[u], shift_of(re[u]));
    for (int k = c0 + kReg; k < c1; ++k) loc += shr64(pd.bsum[k], shift_of(pd.bpow ? pd.bpow[k] : 0));
    unsigned long long sc = loc;
    sc = wave_incl_scan(sc);
    // FP64 weight sum in an order fixed by nb alone (lane l of wave 0 takes
    // blocks l, l + 64, ...), so it does not depend on the block size; block
    // b's sum is relative to 2^e_b and scaled to 2^E (exactly).
    double es = 0.0;
    if (want_es && wid == 0) {
        for (int k = lane; k < nb; k += 8 * 64) {       // same order, 8 loads in flight
            double t8[8];
            int e8[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                t8[u] = (k + 64 * u < nb) ? pd.bexp[k + 64 * u] : 0.0;
                e8[u] = (pd.bpow && k + 64 * u < nb) ? pd.bpow[k + 64 * u] : 0;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (k + 64 * u < nb) {
                    if (!pd.bpow) es += t8[u];
                    else if (e8[u] != kPowEmpty) es += ldexp(t8[u], e8[u] - E);
                }
        }
        es = wave_sum(es);
    }
    __syncthreads();                     // LDS of a previous use is free
    if (lane == 63) misc[wid] = sc;
    if (want_es && tid == 0) reinterpret_cast<double *>(stg)[0] = es;
    __syncthreads();
    unsigned long long run = sc - loc;
    for (int k = 0; k < wid; ++k) run += misc[k];
    unsigned long long total = 0;
    for (int k = 0; k < nwaves; ++k) total += misc[k];
    if (want_es) es_out = reinterpret_cast<double *>(stg)[0];
#pragma unroll
    for (int u = 0; u < kReg; ++u)
        if (c0 + u < c1) {
            const int sh = shift_of(re[u]);
            pre[c0 + u] = run;
            shf[c0 + u] = sh;
            run += shr64(rs[u], sh);
        }
    for (int k = c0 + kReg; k < c1; ++k) {
        const int sh = shift_of(pd.bpow ? pd.bpow[k] : 0);
        pre[k] = run;
        shf[k] = sh;
        run += shr64(pd.bsum[k], sh);
    }
    if (tid == 0) pre[nb] = total;
    __syncthreads();
    pow_out = E;
    return total;
}

// Phase 2: a_j for outputs j in [j0, j0 + blockDim) (j = j0 + threadIdx.x),
// given the prefix of phase 1.  Every thread of the block must call it.
__device__ int resolve_search(const Pending &pd, int N, uint32_t U, unsigned long long total, int j0,
                              unsigned long long *w)
{
    const int tid = threadIdx.x, nt = blockDim.x, nb = pd.nb;
    unsigned long long *pre = w;
    unsigned long long *stg = w + nb + 1;
    unsigned long long *misc = stg + pd.stage;
    const int *shf = reinterpret_cast<const int *>(misc + 16);
    const int bs = 1 << pd.bshift;
    // 2. the input block of each output: largest b with pre[b] * S <= t_j
    const int j = j0 + tid;
    const int jc = min(j, N - 1);
    if (total == 0ull) {
        // every weight zero: the uniform resampling of q_i = 1 (C_i = i + 1,
        // total L), whose a_j = floor((j 2^32 + U) L / (N 2^32)) in closed form
        const u128 x = (u128)((((unsigned long long)(uint32_t)jc) << 32) | U) * (u128)(unsigned)pd.L;
        return (int)((unsigned long long)(x >> 32) / (unsigned long long)N);
    }
    const u128 S = (u128)(unsigned long long)N << 32;
    const u128 tj = sys_target((uint32_t)jc, U, total);
    int lo = 0, hi = nb - 1;
    if ((nt & 63) == 0) {
        // t_j grows with j, so every output's block lies between the blocks
        // of the block's first and last outputs.  Each wave finds those two
        // by 64-way searches: lane l probes position lo + l * step of the
        // range (non-decreasing in l, lane 0 at a position known to pass),
        // and the ballot's count of passing lanes narrows the range to one
        // step (two levels up to 4096 input blocks, one LDS read per lane
        // each).  The output's own search then covers that span only.
        const int lane = tid & 63;
        const int jl = min(j0 + nt, N) - 1;
        const u128 tf = sys_target((uint32_t)j0, U, total), tl = sys_target((uint32_t)jl, U, total);
        int flo = 0, fhi = nb - 1, llo = 0, lhi = nb - 1;
        while (flo < fhi || llo < lhi) {
            const int fs = (fhi - flo + 64) >> 6, ls = (lhi - llo + 64) >> 6;
            const int fp = min(flo + lane * fs, fhi), lp = min(llo + lane * ls, lhi);
            const int fc = __popcll(__ballot((u128)pre[fp] * S <= tf));
            const int lc = __popcll(__ballot((u128)pre[lp] * S <= tl));
            const int fn = min(flo + (fc - 1) * fs, fhi), ln = min(llo + (lc - 1) * ls, lhi);
            fhi = fc < 64 ? min(flo + fc * fs, fhi) - 1 : fhi;
            lhi = lc < 64 ? min(llo + lc * ls, lhi) - 1 : lhi;
            flo = fn;
            llo = ln;
        }
        lo = flo;
        hi = llo;
    }
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((u128)pre[mid] * S <= tj) lo = mid;
        else hi = mid - 1;
    }
    const int bj = lo;
    const int jlast = min(j0 + nt, N) - 1;
    __syncthreads();                     // stg / misc of a previous search are free
    if (j == j0) misc[4] = (unsigned long long)bj;
    if (j == jlast) misc[5] = (unsigned long long)bj;
    __syncthreads();
    const int blo = (int)misc[4], bhi = (int)misc[5];
    // 3. a span of input blocks that fits the stage (pd.stage elements) is
    //    staged in LDS (C = pre + aligned cloc, one memory latency) and
    //    searched there; wider spans (strongly concentrated weights) search
    //    the input block in global memory, one dependent probe per level.
    int a;
    if (((bhi - blo + 1) << pd.bshift) <= pd.stage) {
        const int base = blo << pd.bshift;
        const int n = min(pd.L, (bhi + 1) << pd.bshift) - base;
        for (int k0 = tid; k0 < n; k0 += 4 * nt) {      // 4 loads in flight per lane
            unsigned long long c4[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) c4[u] = (k0 + u * nt < n) ? pd.cloc[base + k0 + u * nt] : 0ull;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int k = k0 + u * nt;
                if (k < n) {
                    const int b = blo + (k >> pd.bshift);
                    stg[k] = pre[b] + (pd.bpow ? shr64(c4[u], shf[b]) : c4[u]);
                }
            }
        }
        __syncthreads();
        int l = (bj - blo) * bs, h = min(n, (bj - blo + 1) * bs) - 1;
        while (l < h) {
            const int mid = (l + h) >> 1;
            if ((u128)stg[mid] * S > tj) h = mid;
            else l = mid + 1;
        }
        a = base + l;
    } else {
        const unsigned long long pb = pre[bj];
        const int sh = pd.bpow ? shf[bj] : 0;
        int l = bj * bs, h = min(pd.L, (bj + 1) * bs) - 1;
        while (l < h) {
            const int mid = (l + h) >> 1;
            if ((u128)(pb + shr64(pd.cloc[mid], sh)) * S > tj) h = mid;
            else l = mid + 1;
        }
        a = l;
    }
    return a;
}

// The frame statistics of a resolved resampling: the largest block exponent
// and the total of the aligned weights (0: every weight zero).
struct ResolveStats {
    unsigned long long total;
    int pow;
};

// Both phases for one block of outputs (hook: as for resolve_prefix).
template <typename Hook = NoHook>
__device__ __forceinline__ int resolve_ancestor(const Pending &pd, int N, uint32_t U, int j0, unsigned long long *w,
                                                bool want_es, double &es_out, ResolveStats &st, Hook hook = Hook())
{
    st.total = resolve_prefix(pd, w, want_es, es_out, st.pow, hook);
    return resolve_search(pd, N, U, st.total, j0, w);
}

// Frame statistics of a consumed resampling (block 0, thread 0): the log
// mean weight, the degenerate flag (no positive weight), and the reset of the
// maximum slot (linear weights).  bexp_sum is the block weight sums' total,
// relative to 2^E for log-weights.
__device__ __forceinline__ void finish_resampling(FrameScratch *fs, int parity, double bexp_sum, const ResolveStats &st,
                                                  int N, int kind_linear, double *lml_slot)
{
    if (st.total == 0ull) atomicOr(&fs->status, kStatusDegenerate);
    if (lml_slot) {
        double l;
        if (st.total == 0ull) l = -INFINITY;
        else if (kind_linear) l = log(bexp_sum / (double)N);
        else l = pow_ref(st.pow) + log(bexp_sum) - log((double)N);
        *lml_slot = l;
    }
    fs->maxkey[parity] = 0ull;
}

// Block-scaled fixed-point weights of one resampler block (kWeightsBlock
// threads, log-weight x of thread i; invalid threads and non-finite x count
// as empty): the block exponent e of its largest finite x (pf_common.hpp),
// q = floor(2^32 exp(x - e ln 2)), the block-local inclusive scan of q to
// *cloc (valid threads), and from the last thread the block sum, the FP64
// sum of exp(x - e ln 2) and e; returns e.  Every thread of the block calls it.
__device__ __forceinline__ int block_weights_log(double x, bool valid, unsigned long long *cloc,
                                                 unsigned long long *bsum, double *bexp, int *bpow)
{
    __shared__ double s_m[kWeightsBlock / 64];
    __shared__ unsigned long long s_wave[kWeightsBlock / 64];
    __shared__ double s_dwave[kWeightsBlock / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const bool fin = valid && isfinite(x);
    double m = wave_max(fin ? x : -INFINITY);
    if (lane == 0) s_m[wid] = m;
    __syncthreads();
    double bm = s_m[0];
#pragma unroll
    for (int w = 1; w < kWeightsBlock / 64; ++w) bm = fmax(bm, s_m[w]);
    const int e = isfinite(bm) ? block_pow(bm) : kPowEmpty;
    const double v = (fin && e != kPowEmpty) ? det_exp(x - pow_ref(e)) : 0.0;
    const unsigned long long q = fixed32(v);
    unsigned long long sc = q;
    sc = wave_incl_scan(sc);
    const double vs = wave_sum(v);
    if (lane == 63) s_wave[wid] = sc;
    if (lane == 0) s_dwave[wid] = vs;
    __syncthreads();
    unsigned long long off = 0;
    for (int w = 0; w < wid; ++w) off += s_wave[w];
    sc += off;
    if (valid) *cloc = sc;
    if (threadIdx.x == kWeightsBlock - 1) {
        *bsum = sc;
        *bexp = (s_dwave[0] + s_dwave[1]) + (s_dwave[2] + s_dwave[3]);
        *bpow = e;
    }
    return e;
}

// --------------------------------------------------------------- step kernel
// One lane = one particle.  Fuses, for particle j of frame t:
//   ancestor (pf2DRao.cpp:152-156), component draw resample(gmm.weight, N)
//   (:128, systematic, thresholds precomputed), KF_model::predict with
//   F = g I (KF_model.cpp:11-15, my_gmm.cpp:53-58), the innovation likelihood
//   (:138) and KF_model::update (KF_model.cpp:17-25), processed as sequential
//   2-row joint blocks (exact for block-diagonal R), then the log-weight write
//   and the block maximum for the resampler.
template <int D>
struct Joint {
    static constexpr int NP = npacked(D);

    // W = L^-1 H_b P (2 x D), y = H_b x + BH_b, S = L L^T.  Returns false
    // when S is ill-conditioned for FP32 (innovation_ill; the outputs are then
    // not used: the joint is redone in FP64).
    __device__ static __forceinline__ bool prepare(const float (&x)[D], const float (&P)[NP],
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
        return !innovation_ill(s00, s01, s11);
    }

    // Same quantities when the state basis makes H_b the selection of rows
    // r0, r0 + 1: H_b P is the pair of rows of P, y the pair of entries of x.
    __device__ static __forceinline__ bool prepare_sel(const float (&x)[D], const float (&P)[NP], int r0,
                                                       float bh0, float bh1, float r00, float r01, float r11,
                                                       float (&w0)[D], float (&w1)[D], float &y0, float &y1,
                                                       float &il00, float &l10, float &il11, float &logl)
    {
        const int r1 = r0 + 1;
        y0 = x[r0] + bh0;
        y1 = x[r1] + bh1;
        const float s00 = P[pk(D, r0, r0)] + r00, s01 = P[pk(D, r0, r1)] + r01, s11 = P[pk(D, r1, r1)] + r11;
        const float l00 = sqrtf(s00);
        il00 = 1.f / l00;
        l10 = s01 * il00;
        const float l11 = sqrtf(s11 - l10 * l10);
        il11 = 1.f / l11;
        logl = logf(l00 * l11);
#pragma unroll
        for (int c = 0; c < D; ++c) {
            w0[c] = P[pk(D, r0, c)] * il00;
            w1[c] = (P[pk(D, r1, c)] - l10 * w0[c]) * il11;
        }
        return !innovation_ill(s00, s01, s11);
    }

    // x += W^T v ; P -= W^T M W with M = [[m00, m01], [m01, m11]].
    __device__ static __forceinline__ void apply(float (&x)[D], float (&P)[NP], float (&w0)[D], float (&w1)[D],
                                                 float v0, float v1, float m00, float m01, float m11)
    {
        if constexpr (D > 16) {
            // Large states: M = l1 e1 e1^T + l2 e2 e2^T, so the update is
            // P -= l1 z1 z1^T + l2 z2 z2^T with z = W^T e overwriting W
            // (half the temporaries of the W^T (M W) form).
#pragma unroll
            for (int c = 0; c < D; ++c) x[c] = fmaf(w0[c], v0, fmaf(w1[c], v1, x[c]));
            const float h = 0.5f * (m00 - m11), mid = 0.5f * (m00 + m11);
            const float r = sqrtf(fmaf(h, h, m01 * m01));
            const float l1 = mid + r, l2 = mid - r;
            float e0 = 1.f, e1 = 0.f;
            if (r > 0.f) {
                const float a0 = h >= 0.f ? h + r : m01, a1 = h >= 0.f ? m01 : r - h;
                const float inv = 1.f / sqrtf(fmaf(a0, a0, a1 * a1));
                e0 = a0 * inv;
                e1 = a1 * inv;
            }
#pragma unroll
            for (int c = 0; c < D; ++c) {
                const float z1 = fmaf(w0[c], e0, w1[c] * e1);
                const float z2 = fmaf(w1[c], e0, -w0[c] * e1);
                w0[c] = z1;
                w1[c] = z2;
            }
#pragma unroll
            for (int i = 0; i < D; ++i) {
                const float a1 = l1 * w0[i], a2 = l2 * w1[i];
#pragma unroll
                for (int k = i; k < D; ++k) {
                    float p = fmaf(-a1, w0[k], P[pk(D, i, k)]);
                    P[pk(D, i, k)] = fmaf(-a2, w1[k], p);
                }
            }
            return;
        }
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
            for (int k = i; k < D; ++k) {
                float p = fmaf(-w0[i], u0[k], P[pk(D, i, k)]);
                P[pk(D, i, k)] = fmaf(-w1[i], u1[k], p);
            }
    }
};

// Copies n4 float4 of the model table into LDS: all loads of a chunk are
// issued before the first LDS write so the copy costs one memory latency.
// The loads are unconditional buffer loads (out-of-range lanes read 0), so
// the chunk stays in registers.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void stage_table(float4 *dst, const float4 *__restrict__ src, int n4, int start = 0)
{
    constexpr int kChunk = 4;
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float4 *>(src), (short)0, n4 * 16, 0x00020000);
    u32x4 *d4 = reinterpret_cast<u32x4 *>(dst);
    for (int base = start; base < n4; base += kChunk * (int)blockDim.x) {
        u32x4 r[kChunk];
#pragma unroll
        for (int it = 0; it < kChunk; ++it) {
            const int i = base + it * (int)blockDim.x + (int)threadIdx.x;
            r[it] = __builtin_amdgcn_raw_buffer_load_b128(rs, i * 16, 0, 0);
        }
#pragma unroll
        for (int it = 0; it < kChunk; ++it) {
            const int i = base + it * (int)blockDim.x + (int)threadIdx.x;
            if (i < n4) d4[i] = r[it];
        }
    }
}

// The first 4 x blockDim float4 of the image loaded into registers ahead of
// their use (out-of-range chunks read as zero), so their latency overlaps
// other loads; table_store writes them, and stages any rest, like
// stage_table.
struct TablePrefetch {
    u32x4 r[4];
};
__device__ __forceinline__ void table_prefetch(TablePrefetch &tp, const float4 *__restrict__ src, int n4)
{
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float4 *>(src), (short)0, n4 * 16, 0x00020000);
#pragma unroll
    for (int it = 0; it < 4; ++it)
        tp.r[it] = __builtin_amdgcn_raw_buffer_load_b128(rs, (it * (int)blockDim.x + (int)threadIdx.x) * 16, 0, 0);
}
__device__ __forceinline__ void table_store(const TablePrefetch &tp, float4 *dst, const float4 *__restrict__ src,
                                            int n4)
{
    u32x4 *d4 = reinterpret_cast<u32x4 *>(dst);
#pragma unroll
    for (int it = 0; it < 4; ++it) {
        const int i = it * (int)blockDim.x + (int)threadIdx.x;
        if (i < n4) d4[i] = tp.r[it];
    }
    stage_table(dst, src, n4, 4 * (int)blockDim.x);
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

// State stores of the one-particle-per-lane kernels carry device scope (cache
// policy sc1): the new state of a frame is read by the next launch's blocks
// on every XCD, and each store instruction writes whole 128-byte lines (64
// consecutive particles of one row).  Measured at N = 100k, d = 12: 33.1 ->
// 31.9 us per frame against default-scope stores, same results bit for bit.
// Stores that write partial lines (the four-lane d = 24 kernels: 16
// particles per instruction) keep the default policy (soa_st_partial): with
// device scope they were 5% slower.
constexpr int kStoreDeviceScope = 16;   // CPol::SC1

__device__ __forceinline__ void soa_st(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_int(v), r, (int)voff, (int)soff, kStoreDeviceScope);
}

__device__ __forceinline__ void soa_st_partial(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_int(v), r, (int)voff, (int)soff, 0);
}

// Layout of a particle state buffer: element r of particle j of x (r < D)
// or of packed P (r < NP) at base[r * row + j * part], P's base its own.
// Rows (structure of arrays: row = ld, part = 1, P's base after the D
// rows of x) or one record per particle (x, then packed P, padded to 16
// bytes: row = 1, part = rec_floats(D), P's base = x's + D).  A resampled
// particle's record is ~3 cache lines (d = 12) wherever its ancestor lies,
// where rows cost one line per row for ancestors spread thinly; the record
// gather and store go through LDS and hold more registers.  The batched
// engine keeps rows; a single filter with d <= 12 may use records.
struct Lay {
    int row, part;
};
__host__ __device__ constexpr int rec_floats(int D) { return (D + npacked(D) + 3) & ~3; }
__host__ __device__ constexpr bool rec_layout(int D) { return D <= 12; }     // records compiled for this d
__host__ __device__ inline size_t lidx(const Lay &L, int r, int j) { return (size_t)r * L.row + (size_t)j * L.part; }

// LDS of the record gather and store of the step: per wave the records of
// 16 particles (d <= 12: at most 16 x 23 chunks of 16 bytes), 4 waves.
constexpr int kRecStageChunks = 16 * 23;
constexpr int kRecStageBytes = 4 * kRecStageChunks * 16;

// Wave-cooperative gather of the records of the wave's 64 particles, in four
// rounds of 16: lane k of a round loads 16-byte chunk k % NCH of particle
// k / NCH's ancestor src (consecutive lanes read consecutive bytes of one
// record, then of the next); the four rounds' loads are issued together,
// then each round's chunks go to the wave's LDS slice gw in turn and the
// round's 16 lanes take their particle's chunks into x and P.  Particles
// j0w + p >= N read nothing (the buffer resource's bound returns zeros for
// them).  Every lane of the wave calls it.
template <int D>
__device__ __forceinline__ void rec_gather(const float *rec, int N, int j0w, int src, float4 *gw, float (&x)[D],
                                           float (&P)[npacked(D)])
{
    constexpr int NP = npacked(D), RS = rec_floats(D), NCH = RS / 4;
    static_assert(16 * NCH <= kRecStageChunks, "a round of 16 records exceeds the LDS slice");
    const int lane = threadIdx.x & 63;
    const unsigned lim = (unsigned)N * RS * 4u;
    const __amdgpu_buffer_rsrc_t ra =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(rec), (short)0, (int)lim, 0x00020000);
    // every round's loads first: the wave's 64 records share one memory
    // latency; the rounds then pass through the LDS slice in turn
    constexpr int NI = (16 * NCH + 63) / 64;
    u32x4 c[4][NI];
#pragma unroll
    for (int g = 0; g < 4; ++g) {
#pragma unroll
        for (int i = 0; i < NI; ++i) {
            const int k = i * 64 + lane;
            const int p = k / NCH, ch = k - p * NCH;
            const int sp = __shfl(src, 16 * g + (p < 16 ? p : 15), 64);
            const bool ok = k < 16 * NCH && j0w + 16 * g + p < N;
            c[g][i] = __builtin_amdgcn_raw_buffer_load_b128(ra, ok ? (int)((unsigned)(sp * RS + ch * 4) * 4u) : (int)lim,
                                                            0, 0);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int g = 0; g < 4; ++g) {
#pragma unroll
        for (int i = 0; i < NI; ++i) {
            const int k = i * 64 + lane;
            if (k < 16 * NCH) reinterpret_cast<u32x4 *>(gw)[k] = c[g][i];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if ((lane >> 4) == g) {
#pragma unroll
            for (int cc = 0; cc < NCH; ++cc) {
                const float4 v = gw[(lane & 15) * NCH + cc];
                const float ve[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int f = cc * 4 + e;
                    if (f < D) x[f] = ve[e];
                    else if (f < D + NP) P[f - D] = ve[e];
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// The inverse: the wave's 64 records (particles j0w .. j0w + 63, contiguous
// in the buffer) written from x and P through gw in four rounds of 16; each
// round stores 16 whole records as consecutive 16-byte chunks, so the stores
// write whole lines and carry device scope like the row stores.  Every lane
// of the wave calls it.
template <int D>
__device__ __forceinline__ void rec_store(float *rec, int N, int j0w, float4 *gw, const float (&x)[D],
                                          const float (&P)[npacked(D)])
{
    constexpr int NP = npacked(D), RS = rec_floats(D), NCH = RS / 4;
    static_assert(16 * NCH <= kRecStageChunks, "a round of 16 records exceeds the LDS slice");
    const int lane = threadIdx.x & 63;
    const unsigned lim = (unsigned)N * RS * 4u;
    const __amdgpu_buffer_rsrc_t wa = __builtin_amdgcn_make_buffer_rsrc(rec, (short)0, (int)lim, 0x00020000);
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        if ((lane >> 4) == g) {
#pragma unroll
            for (int cc = 0; cc < NCH; ++cc) {
                float ve[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int f = cc * 4 + e;
                    ve[e] = f < D ? x[f] : (f < D + NP ? P[f - D] : 0.f);
                }
                gw[(lane & 15) * NCH + cc] = make_float4(ve[0], ve[1], ve[2], ve[3]);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const unsigned base = (unsigned)(j0w + 16 * g) * RS * 4u;
        const int nvalid = min(16, N - (j0w + 16 * g)) * NCH;
#pragma unroll
        for (int i = 0; i < (16 * NCH + 63) / 64; ++i) {
            const int k = i * 64 + lane;
            const bool ok = k < 16 * NCH && k < nvalid;
            const u32x4 v = reinterpret_cast<const u32x4 *>(gw)[k < 16 * NCH ? k : 0];
            __builtin_amdgcn_raw_buffer_store_b128(v, wa, ok ? (int)(base + (unsigned)k * 16u) : (int)lim, 0,
                                                   kStoreDeviceScope);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

__device__ __forceinline__ int draw_component(const unsigned long long *thr, int K, uint32_t j,
                                              uint32_t U, unsigned long long total)
{
    // The first k < K - 1 whose threshold exceeds the target, else K - 1.
    // The thresholds are cumulative (non-decreasing), so k is the count of
    // thresholds at or below the target: they are read in groups of eight
    // (one LDS latency per group), and the wave stops after the first group
    // whose last threshold exceeds every lane's target.
    const u128 t = sys_target(j, U, total);
    int k = 0;
    for (int k0 = 0; k0 < K - 1; k0 += 8) {
        u128 th[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int i = min(k0 + u, K - 2);
            th[u] = ((u128)thr[2 * i + 1] << 64) | (u128)thr[2 * i];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (k0 + u < K - 1 && th[u] <= t) ++k;
        if (!__any(th[7] <= t)) break;
    }
    return k;
}

// Detections of joint b in frame t (candidate mode) or the per-particle
// measurement (reference mode).  Returns false if a certain joint has no
// detection this frame (the joint is skipped).
// Candidate mode stages the frame's detections in LDS after the model image:
// J x cap x 2 floats, then J int32 counts (padded to an even count).
__host__ __device__ constexpr int cand_lds_floats(int J, int cap) { return J * cap * 2 + ((J + 1) & ~1); }

// Loads of the staging are issued before the model image's and stored after
// it, so both share one memory latency.
struct CandPrefetch {
    float v[4];
    int c;
};
__device__ __forceinline__ void cand_prefetch(CandPrefetch &pf, const int *cnt, const float *cand, int J, int cap)
{
    const int n = J * cap * 2, bd = blockDim.x, tid = threadIdx.x;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i = tid + r * bd;
        pf.v[r] = i < n ? cand[i] : 0.f;
    }
    pf.c = tid < J ? cnt[tid] : 0;
}
__device__ __forceinline__ void cand_store(const CandPrefetch &pf, float *dst, const float *cand, int J, int cap)
{
    const int n = J * cap * 2, bd = blockDim.x, tid = threadIdx.x;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i = tid + r * bd;
        if (i < n) dst[i] = pf.v[r];
    }
    for (int i = tid + 4 * bd; i < n; i += bd) dst[i] = cand[i];
    if (tid < J) reinterpret_cast<int *>(dst + n)[tid] = pf.c;
}

// The staged detection count of joint b (candidate mode).
__device__ __forceinline__ int staged_count(const StepArgs &a, int b, const float *sm)
{
    return reinterpret_cast<const int *>(sm + a.table_floats + a.J * a.cap * 2)[b];
}

// joint_inputs with the joint's staged count n already read.
template <int MODE>
__device__ __forceinline__ bool joint_inputs_n(const StepArgs &a, int b, int n, int &cnt, const float *&cz,
                                               bool &clutter, const float *sm)
{
    clutter = (MODE == kModeCand) && ((a.clutter_mask >> b) & 1u);
    cnt = 1;
    cz = nullptr;
    if (MODE == kModeCand) {
        cnt = n;
        cz = sm + a.table_floats + b * a.cap * 2;
        if (cnt <= 0 && !clutter) return false;
        if (cnt > a.cap) cnt = a.cap;
    }
    return true;
}

// Detections of joint b in frame t (candidate mode, from the LDS staging at
// sm + table_floats) or the per-particle measurement (reference mode).
// Returns false if a certain joint has no detection this frame (the joint is
// skipped).
template <int MODE>
__device__ __forceinline__ bool joint_inputs(const StepArgs &a, int t, int b, int &cnt, const float *&cz,
                                             bool &clutter, const float *sm)
{
    (void)t;
    return joint_inputs_n<MODE>(a, b, MODE == kModeCand ? staged_count(a, b, sm) : 1, cnt, cz, clutter, sm);
}

// Whitening of an innovation, u = L^-1 (z - y) with S = L L^T, L = [[l00,
// 0], [l10, l11]]: in FP32 from the FP32 factor, or in FP64 from the FP64
// factor of an ill-conditioned S (the cancellation is in this step; the
// whitened values are well scaled and rounded to FP32).
struct Whiten32 {
    float y0, y1, il00, l10, il11;
    __device__ __forceinline__ void operator()(float z0, float z1, float &u0, float &u1) const
    {
        u0 = (z0 - y0) * il00;
        u1 = ((z1 - y1) - l10 * u0) * il11;
    }
};
struct Whiten64 {
    double y0, y1, il00, l10, il11;
    __device__ __forceinline__ void operator()(float z0, float z1, float &u0, float &u1) const
    {
        const double a0 = ((double)z0 - y0) * il00;
        u0 = (float)a0;
        u1 = (float)((((double)z1 - y1) - l10 * a0) * il11);
    }
};

// Likelihood of one prepared joint block and the parameters of its update
// x += W^T v, P -= W^T M W (M = I for a certain measurement); returns the
// joint's log-likelihood term.  logl = log(l00 l11).
template <int MODE, typename WH>
__device__ __forceinline__ float joint_measure(const StepArgs &a, int t, int j, int b, bool clutter, int cnt,
                                               const float *cz, const WH &wh, float logl, float &v0, float &v1,
                                               float &m00, float &m01, float &m11)
{
    const float LOG2PI = 1.8378770664093453f;
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
        wh(z0, z1, v0, v1);
        m00 = 1.f;
        m01 = 0.f;
        m11 = 1.f;
        return -0.5f * fmaf(v0, v0, v1 * v1) - logl - LOG2PI;
    }
    // PDAF over the gated detections (Bar-Shalom & Fortmann), in whitened
    // coordinates u_c = L^-1 (z_c - y).
    const float lnorm = -logl - LOG2PI;
    float se = 0.f, sv0 = 0.f, sv1 = 0.f, q00 = 0.f, q01 = 0.f, q11 = 0.f;
    auto one = [&](float z0, float z1) {
        float u0, u1;
        wh(z0, z1, u0, u1);
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
    };
    // detections in groups of four read before they are used (one LDS
    // latency per group), in the same order as one at a time
    const float2 *cz2 = reinterpret_cast<const float2 *>(cz);
    int c = 0;
    for (; c + 4 <= cnt; c += 4) {
        float2 g[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) g[u] = cz2[c + u];
#pragma unroll
        for (int u = 0; u < 4; ++u) one(g[u].x, g[u].y);
    }
    for (; c < cnt; ++c) one(cz[2 * c], cz[2 * c + 1]);
    const float ell = a.b0 + se;
    const float inv = 1.f / ell;
    v0 = sv0 * inv;
    v1 = sv1 * inv;
    const float bsum = se * inv;  // 1 - beta_0
    m00 = bsum - (q00 * inv - v0 * v0);
    m01 = -(q01 * inv - v0 * v1);
    m11 = bsum - (q11 * inv - v1 * v1);
    return logf(ell);
}

template <int MODE>
__device__ __forceinline__ float joint_measure(const StepArgs &a, int t, int j, int b, bool clutter, int cnt,
                                               const float *cz, float y0, float y1, float il00, float l10,
                                               float il11, float logl, float &v0, float &v1, float &m00,
                                               float &m01, float &m11)
{
    return joint_measure<MODE>(a, t, j, b, clutter, cnt, cz, Whiten32{y0, y1, il00, l10, il11}, logl, v0, v1, m00,
                               m01, m11);
}

// The FP64 factor of one joint whose innovation covariance FP32 cannot
// factor reliably (innovation_ill): S = H_b P H_b^T + R_b and y = H_b x +
// BH_b from the FP32 state and the FP64 innovation block j64 = (bh0, bh1,
// r00, r01, r11), the FP64 Cholesky, W = L^-1 H_b P rounded to FP32 (the
// update P - W^T M W stays within the FP32 rounding of P: W^T W <= P) and
// the FP64 whitening.  SEL: H_b selects rows r0, r0 + 1 (h0, h1 unused),
// otherwise h0, h1 are its rows.  Sets kStatusIllCond; returns false (and
// sets kStatusNotSpd) if S is not positive definite in FP64 either.
template <int D, bool SEL>
__device__ __forceinline__ bool joint_factor64(const float (&x)[D], const float (&P)[npacked(D)], int r0,
                                               const float *h0, const float *h1, const double *j64, Whiten64 &wh,
                                               float &logl, float (&w0)[D], float (&w1)[D], unsigned &flags)
{
    flags |= kStatusIllCond;
    auto hp = [&](int row, int c) -> double {       // (H_b P)[row][c]
        if constexpr (SEL) {
            return (double)P[pk(D, r0 + row, c)];
        } else {
            const float *h = row ? h1 : h0;
            double acc = 0.0;
            for (int k = 0; k < D; ++k) acc = fma((double)h[k], (double)P[pk(D, k, c)], acc);
            return acc;
        }
    };
    double y0 = j64[0], y1 = j64[1], s00 = j64[2], s01 = j64[3], s11 = j64[4];
    if constexpr (SEL) {
        y0 += (double)x[r0];
        y1 += (double)x[r0 + 1];
        s00 += (double)P[pk(D, r0, r0)];
        s01 += (double)P[pk(D, r0, r0 + 1)];
        s11 += (double)P[pk(D, r0 + 1, r0 + 1)];
    } else {
        for (int c = 0; c < D; ++c) {
            const double p0 = hp(0, c), p1 = hp(1, c);
            y0 = fma((double)h0[c], (double)x[c], y0);
            y1 = fma((double)h1[c], (double)x[c], y1);
            s00 = fma((double)h0[c], p0, s00);
            s01 = fma((double)h1[c], p0, s01);
            s11 = fma((double)h1[c], p1, s11);
        }
    }
    const double d11 = s00 > 0.0 ? s11 - s01 * s01 / s00 : 0.0;
    if (!(s00 > 0.0) || !(d11 > 0.0)) {
        flags |= kStatusNotSpd;
        return false;
    }
    const double l00 = sqrt(s00), l11 = sqrt(d11);
    wh.y0 = y0;
    wh.y1 = y1;
    wh.il00 = 1.0 / l00;
    wh.l10 = s01 * wh.il00;
    wh.il11 = 1.0 / l11;
    logl = (float)log(l00 * l11);
    // W rows: w0 = hp0 / l00 and w1 = (hp1 - rho hp0) / l11 with rho = s01 / s00
    // as a double-float pair, so the cancellation in hp1 - rho hp0 happens
    // inside two FMAs (each rounds once) and the FP32 rows stay accurate.
    const double rho = s01 / s00;
    const float rhi = (float)rho, rlo = (float)(rho - (double)rhi);
    const float fil00 = (float)wh.il00, fil11 = (float)wh.il11;
#pragma unroll
    for (int c = 0; c < D; ++c) {
        float p0, p1;
        if constexpr (SEL) {
            p0 = P[pk(D, r0, c)];
            p1 = P[pk(D, r0 + 1, c)];
        } else {
            p0 = (float)hp(0, c);
            p1 = (float)hp(1, c);
        }
        w0[c] = p0 * fil00;
        w1[c] = fmaf(-rlo, p0, fmaf(-rhi, p0, p1)) * fil11;
    }
    return true;
}

// ORs the wave's particle flags (kStatusIllCond, kStatusNotSpd) into *status
// with one atomic per wave, only when a flag is raised.  Every lane of the
// wave calls it.
__device__ __forceinline__ void report_flags(unsigned flags, unsigned int *status)
{
    const unsigned long long ill = __ballot((flags & kStatusIllCond) != 0u);
    const unsigned long long bad = __ballot((flags & kStatusNotSpd) != 0u);
    if ((ill | bad) && (threadIdx.x & 63) == 0)
        atomicOr(status, (ill ? kStatusIllCond : 0u) | (bad ? kStatusNotSpd : 0u));
}

// LDS bytes of k_step: model image + ancestor workspace (fused mode).
__host__ __device__ constexpr int step_lds_bytes(int stage_floats, int nb_pending)
{
    return 4 * stage_floats + (nb_pending > 0 ? 8 * resolve_lds_words(nb_pending) : 0);
}

// The single filter's step with per-particle records: the record gather's
// LDS follows the staged image, 16-byte aligned, sharing the resolver's
// workspace.
__host__ __device__ constexpr int step_lds_bytes_rec(bool rec, int stage_floats, int nb_pending)
{
    return !rec ? step_lds_bytes(stage_floats, nb_pending)
                : 4 * ((stage_floats + 3) & ~3) +
                      (nb_pending > 0 && 8 * resolve_lds_words(nb_pending) > kRecStageBytes
                           ? 8 * resolve_lds_words(nb_pending)
                           : kRecStageBytes);
}

// Model image layout: [component thresholds, 4K floats][H (m x D), BH (m),
// R 2x2 blocks (3J), padded to 4][component rows, K x (2 + D + D(D+1)/2)].
// The head is always staged in LDS; the component rows are staged with it
// when the whole image fits the LDS budget, and read from global memory
// (StepArgs::comp_global) otherwise.

// One particle's frame: component draw (pf2DRao.cpp:128), mean-reverting
// predict (KF_model.cpp:11-15) and the joint updates (pf2DRao.cpp:138-140);
// returns the particle's log-weight.  sm is the staged LDS model image.
// The shared constants come from a (a kernel argument, read on demand); the
// per-sequence values pi_total (component thresholds' total), comp_global
// (component rows outside LDS, or null) and comp_out (optional component
// record) are passed explicitly.
// Likelihood and update parameters of a joint that is not positive definite
// in FP64: zero weight (-inf) and no update.
template <int D>
__device__ __forceinline__ float joint_reject(float (&w0)[D], float (&w1)[D], float &v0, float &v1, float &m00,
                                              float &m01, float &m11)
{
#pragma unroll
    for (int c = 0; c < D; ++c) w0[c] = w1[c] = 0.f;
    v0 = v1 = m00 = m01 = m11 = 0.f;
    return -INFINITY;
}

// A joint whose innovation covariance FP32 cannot factor reliably
// (innovation_ill) is factored in FP64 from the particle's current state
// (joint_factor64, a per-lane branch taken only for those joints).
template <int D, int MODE, int SELJ>
__device__ __forceinline__ float particle_update(const StepArgs &a, int t, int j, uint32_t U1, const float *sm,
                                                 float (&x)[D], float (&P)[npacked(D)], unsigned long long pi_total,
                                                 const float *comp_global, int *comp_out, unsigned &flags)
{
    constexpr int NP = npacked(D);
    constexpr int CS = 2 + D + NP;
    const unsigned long long *sthr = reinterpret_cast<const unsigned long long *>(sm);
    const float *Hs = sm + 4 * a.K;
    const float *BHs = Hs + a.m * D;
    const float *Rs = BHs + a.m;
    const double *J64 = reinterpret_cast<const double *>(sm + image_j64_floats(a.K, a.m, D, a.J));
    const float *stab = comp_global ? comp_global : sm + image_head_floats(a.K, a.m, D, a.J);
    const int k = draw_component(sthr, a.K, (uint32_t)j, U1, pi_total);
    if (comp_out) comp_out[j] = k;
    const float *ck = stab + k * CS;
    const float g = ck[0], g2 = ck[1];
#pragma unroll
    for (int i = 0; i < D; ++i) x[i] = fmaf(g, x[i], ck[2 + i]);
#pragma unroll
    for (int p = 0; p < NP; ++p) P[p] = fmaf(g2, P[p], ck[2 + D + p]);

    float acc = 0.f;
    if constexpr (SELJ > 0) {
        // Candidate mode: every joint's staged detection count and
        // innovation constants read before the first joint (one LDS
        // latency); they are the same in every lane and are kept in scalar
        // registers.  Reference mode reads them at each joint.
        int jn[SELJ];
        float jc[SELJ][5];
        auto joint_consts = [&](int oi, int b) {
            jn[oi] = MODE == kModeCand ? staged_count(a, b, sm) : 1;
            jc[oi][0] = BHs[2 * b];
            jc[oi][1] = BHs[2 * b + 1];
            jc[oi][2] = Rs[3 * b];
            jc[oi][3] = Rs[3 * b + 1];
            jc[oi][4] = Rs[3 * b + 2];
        };
        if constexpr (MODE == kModeCand) {
            auto suni = [](float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); };
#pragma unroll
            for (int oi = 0; oi < SELJ; ++oi) joint_consts(oi, (int)((a.order_packed >> (4 * oi)) & 15u));
#pragma unroll
            for (int oi = 0; oi < SELJ; ++oi) {
                jn[oi] = __builtin_amdgcn_readfirstlane(jn[oi]);
#pragma unroll
                for (int e = 0; e < 5; ++e) jc[oi][e] = suni(jc[oi][e]);
            }
        }
#pragma unroll
        for (int oi = 0; oi < SELJ; ++oi) {
            const int b = (int)((a.order_packed >> (4 * oi)) & 15u);
            if constexpr (MODE != kModeCand) joint_consts(oi, b);
            int cnt;
            const float *cz;
            bool clutter;
            if (!joint_inputs_n<MODE>(a, b, jn[oi], cnt, cz, clutter, sm)) continue;
            float w0[D], w1[D], y0, y1, il00, l10, il11, logl;
            const bool ok = Joint<D>::prepare_sel(x, P, 2 * oi, jc[oi][0], jc[oi][1], jc[oi][2], jc[oi][3],
                                                  jc[oi][4], w0, w1, y0, y1, il00, l10, il11, logl);
            // States up to d = 16 apply the update inside each branch: on the
            // FP32 path the compiler may then start it before the likelihood
            // is complete (a certain joint's M is the identity), which keeps W
            // short-lived; the 24-dimensional state keeps one update after
            // the branches (fewer values to spill).
            float v0, v1, m00, m01, m11;
            if (ok) {
                acc += joint_measure<MODE>(a, t, j, b, clutter, cnt, cz, y0, y1, il00, l10, il11, logl, v0,
                                           v1, m00, m01, m11);
                if constexpr (D <= 16) Joint<D>::apply(x, P, w0, w1, v0, v1, m00, m01, m11);
            } else {
                Whiten64 wh;
                if (joint_factor64<D, true>(x, P, 2 * oi, nullptr, nullptr, J64 + 5 * b, wh, logl, w0, w1, flags)) {
                    acc += joint_measure<MODE>(a, t, j, b, clutter, cnt, cz, wh, logl, v0, v1, m00, m01, m11);
                    if constexpr (D <= 16) Joint<D>::apply(x, P, w0, w1, v0, v1, m00, m01, m11);
                } else {
                    acc += joint_reject<D>(w0, w1, v0, v1, m00, m01, m11);   // zero weight, no update
                }
            }
            if constexpr (D > 16) Joint<D>::apply(x, P, w0, w1, v0, v1, m00, m01, m11);
        }
    } else {
        for (int oi = 0; oi < a.J; ++oi) {
            const int b = (MODE == kModeRef) ? oi : (int)((a.order_packed >> (4 * oi)) & 15u);
            int cnt;
            const float *cz;
            bool clutter;
            if (!joint_inputs<MODE>(a, t, b, cnt, cz, clutter, sm)) continue;
            float w0[D], w1[D], y0, y1, il00, l10, il11, logl;
            const bool ok = Joint<D>::prepare(x, P, Hs + (2 * b) * D, Hs + (2 * b + 1) * D, BHs[2 * b],
                                              BHs[2 * b + 1], Rs[3 * b], Rs[3 * b + 1], Rs[3 * b + 2], w0, w1, y0,
                                              y1, il00, l10, il11, logl);
            float v0, v1, m00, m01, m11;
            if (ok) {
                acc += joint_measure<MODE>(a, t, j, b, clutter, cnt, cz, y0, y1, il00, l10, il11, logl, v0,
                                           v1, m00, m01, m11);
                if constexpr (D <= 16) Joint<D>::apply(x, P, w0, w1, v0, v1, m00, m01, m11);
            } else {
                Whiten64 wh;
                if (joint_factor64<D, false>(x, P, 0, Hs + (2 * b) * D, Hs + (2 * b + 1) * D, J64 + 5 * b, wh, logl,
                                             w0, w1, flags)) {
                    acc += joint_measure<MODE>(a, t, j, b, clutter, cnt, cz, wh, logl, v0, v1, m00, m01, m11);
                    if constexpr (D <= 16) Joint<D>::apply(x, P, w0, w1, v0, v1, m00, m01, m11);
                } else {
                    acc += joint_reject<D>(w0, w1, v0, v1, m00, m01, m11);   // zero weight, no update
                }
            }
            if constexpr (D > 16) Joint<D>::apply(x, P, w0, w1, v0, v1, m00, m01, m11);
        }
    }
    return acc;
}

// REC: the state buffers hold per-particle records (rows otherwise).
template <int D, int MODE, int SELJ, bool REC>
__device__ __forceinline__ void step_body(const StepArgs &ain)
{
    constexpr int NP = npacked(D);
    extern __shared__ float4 sm4[];
    int t;
    const StepArgs a = frame_args(ain, t);
    const uint32_t U1 = gload(a.uni + 2 * t);
    const int j0 = blockIdx.x * blockDim.x;
    const int j = j0 + threadIdx.x;
    const bool live = j < a.N;
    const float4 *table4 = reinterpret_cast<const float4 *>(a.table);
    const int *fcount = MODE == kModeCand ? a.count + (size_t)t * a.J : nullptr;
    const float *fcand = MODE == kModeCand ? a.cand + (size_t)t * a.J * a.cap * 2 : nullptr;
    // Records: the model image and the frame's detections are loaded beside
    // the resolver's first loads (one shared memory latency; their LDS area
    // is not the resolver's) and stored to LDS before the gather.
    TablePrefetch tpf;
    CandPrefetch cpf;
    auto prefetch = [&]() {
        if constexpr (REC) {
            table_prefetch(tpf, table4, a.table_floats >> 2);
            if (MODE == kModeCand) cand_prefetch(cpf, fcount, fcand, a.J, a.cap);
        }
    };
    int src;
    if (a.anc_mode == kAncFused) {
        // Consume the previous frame's resampling: a_j, then the frame's
        // statistics (block 0).
        unsigned long long *wsp = reinterpret_cast<unsigned long long *>(
            reinterpret_cast<float *>(sm4) + a.stage_floats);
        double es = 0.0;
        ResolveStats rs;
        // the previous frame's resampling uniform (its step recorded it as
        // pending too) and its largest block exponent (raised by its blocks)
        Pending pd = a.pend;
        if (a.w_cloc) pd.pow_max = &a.fs->pow_max[(t - 1) % 3];
        src = resolve_ancestor(pd, a.N, a.uni[2 * (t - 1) + 1], j0, wsp, blockIdx.x == 0, es, rs, prefetch);
        if (live) a.anc[j] = src;
        if (blockIdx.x == 0 && threadIdx.x == 0)
            finish_resampling(a.fs, a.parity ^ 1, es, rs, a.N, 0, a.lml ? a.lml + (t - 1) : nullptr);
    } else {
        prefetch();
        src = live ? a.anc[j] : 0;
    }
    // Rows: gathered state loads, then the model staging; the LDS barrier
    // waits for both, so they share one memory latency.
    const unsigned rowb = (unsigned)a.ld * 4u;
    const unsigned voff = (unsigned)src * 4u;
    float x[D], P[NP];
    // records: gathered by the wave through its LDS slice after the
    // resolver's workspace (free once every thread has its ancestor)
    const int j0w = j0 + (int)(threadIdx.x & ~63u);
    float4 *gw = reinterpret_cast<float4 *>(reinterpret_cast<float *>(sm4) + ((a.stage_floats + 3) & ~3)) +
                 (threadIdx.x >> 6) * kRecStageChunks;
    static_assert(!REC || rec_layout(D), "records are compiled for d <= 12");
    if constexpr (REC) {
        table_store(tpf, sm4, table4, a.table_floats >> 2);
        if (MODE == kModeCand) cand_store(cpf, reinterpret_cast<float *>(sm4) + a.table_floats, fcand, a.J, a.cap);
        if (a.anc_mode == kAncFused) __syncthreads();
        rec_gather<D>(a.xin, a.N, j0w, src, gw, x, P);
    } else {
        if (live) {
            const __amdgpu_buffer_rsrc_t rx = soa_rsrc(a.xin, rowb * D), rP = soa_rsrc(a.Pin, rowb * NP);
#pragma unroll
            for (int i = 0; i < D; ++i) x[i] = soa_ld(rx, voff, rowb * i);
#pragma unroll
            for (int p = 0; p < NP; ++p) P[p] = soa_ld(rP, voff, rowb * p);
        }
        if (MODE == kModeCand) cand_prefetch(cpf, fcount, fcand, a.J, a.cap);
        stage_table(sm4, table4, a.table_floats >> 2);
        if (MODE == kModeCand) cand_store(cpf, reinterpret_cast<float *>(sm4) + a.table_floats, fcand, a.J, a.cap);
    }
    __syncthreads();

    float lwj = -INFINITY;
    unsigned pflags = 0;
    if (live) {
        lwj = particle_update<D, MODE, SELJ>(a, t, j, U1, reinterpret_cast<const float *>(sm4), x, P, a.pi_total,
                                             a.comp_global, a.comp_out, pflags);
        if constexpr (!REC) {
            const unsigned doff = (unsigned)j * 4u;
            const __amdgpu_buffer_rsrc_t wx = soa_rsrc(a.xout, rowb * D), wP = soa_rsrc(a.Pout, rowb * NP);
#pragma unroll
            for (int i = 0; i < D; ++i) soa_st(wx, doff, rowb * i, x[i]);
#pragma unroll
            for (int p = 0; p < NP; ++p) soa_st(wP, doff, rowb * p, P[p]);
        }
        a.lw[j] = lwj;
    }
    if constexpr (REC) rec_store<D>(a.xout, a.N, j0w, gw, x, P);
    report_flags(pflags, &a.fs->status);
    if (a.w_cloc) {
        // this frame's resampler weights: the block is one resampler block;
        // its exponent raises slot t mod 3, and block 0 clears slot t + 1
        // (last read by frame t - 1's resolution, written next by frame t + 1)
        const int e = block_weights_log((double)lwj, live, a.w_cloc + (live ? j : 0), a.w_bsum + blockIdx.x,
                                        a.w_bexp + blockIdx.x, a.w_bpow + blockIdx.x);
        if (threadIdx.x == 0 && e != kPowEmpty) atomicMax(&a.fs->pow_max[t % 3], e);
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            a.fs->U_pending = a.uni[2 * t + 1];
            a.fs->pow_max[(t + 1) % 3] = kPowEmpty;
            if (a.frame_next) *a.frame_next = t + 1;
        }
    }
}

// At least two waves per SIMD for states up to d = 16 in the selection
// basis: the reference-mode kernels would otherwise take more than 256
// registers (their FP64 joint branch) and run one.
template <int D, int MODE, int SELJ, bool REC = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu((D > 16 || SELJ == 0) ? 1 : 2, 8))) k_step(StepArgs a)
{
    step_body<D, MODE, SELJ, REC>(a);
}

// Two filters of the same shape in one launch (grid.y = filter): the pose
// tracker advances its two arms together (state in rows: the paired launch
// has twice the blocks, where rows are the faster layout).
template <int D, int MODE, int SELJ>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu((D > 16 || SELJ == 0) ? 1 : 2, 8))) k_step_pair(StepArgs a0, StepArgs a1)
{
    step_body<D, MODE, SELJ, false>(blockIdx.y ? a1 : a0);
}

// A state buffer copied into another layout (the same particles).
template <int D>
__global__ void __launch_bounds__(256) k_relayout(const float *xs, const float *Ps, Lay Ls, float *xd, float *Pd,
                                                  Lay Ld, int N)
{
    constexpr int NP = npacked(D);
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
#pragma unroll
    for (int i = 0; i < D; ++i) xd[lidx(Ld, i, j)] = xs[lidx(Ls, i, j)];
#pragma unroll
    for (int p = 0; p < NP; ++p) Pd[lidx(Ld, p, j)] = Ps[lidx(Ls, p, j)];
}

// Weights of the resampler: q_i and per-block inclusive scans.
enum WeightKind { kLogF32 = 0, kLogF64 = 1, kLinF64 = 2 };

struct WeightArgs {
    const float *lw32;
    const double *w64;
    unsigned long long *cloc;     // L block-local inclusive prefix sums
    unsigned long long *bsum;     // nb block sums
    double *bexp;                 // nb block sums of the FP64 weights
    int *bpow;                    // nb block exponents (log-weights)
    FrameScratch *fs;
    const int *frame;             // frame counter read this frame
    int *frame_next;              // frame counter of the next frame (block 0 writes t + 1)
    const uint32_t *uni;          // U = uni[t * ustride + uoff]
    int L, parity, ustride, uoff;
};

template <int KIND>
__device__ __forceinline__ void weights_body(const WeightArgs &a)
{
    const int i = blockIdx.x * kWeightsBlock + threadIdx.x;
    if (KIND == kLogF32 || KIND == kLogF64) {
        const bool valid = i < a.L;
        const double x = !valid ? -INFINITY : (KIND == kLogF32 ? (double)a.lw32[i] : a.w64[i]);
        block_weights_log(x, valid, a.cloc + (valid ? i : 0), a.bsum + blockIdx.x, a.bexp + blockIdx.x,
                          a.bpow + blockIdx.x);
    } else {
        // linear weights relative to their maximum (slot `parity`)
        __shared__ unsigned long long s_wave[4];
        __shared__ double s_dwave[4];
        const unsigned long long key = a.fs->maxkey[a.parity];
        const bool degenerate = (key == 0ull);
        const double ref = degenerate ? 0.0 : key_value(key);
        unsigned long long q = 0;
        double v = 0.0;
        if (i < a.L) {
            const double x = a.w64[i];
            const bool ok = !degenerate && isfinite(x) && x > 0.0;
            v = ok ? x : 0.0;
            q = ok ? fixed32(x / ref) : 0;
        }
        const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
        unsigned long long s = q;
        s = wave_incl_scan(s);
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
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const int t = *a.frame;
        a.fs->U_pending = a.uni[(size_t)t * a.ustride + a.uoff];
        if (a.frame_next) *a.frame_next = t + 1;
    }
}

template <int KIND>
__global__ void __launch_bounds__(kWeightsBlock) k_weights(WeightArgs a)
{
    weights_body<KIND>(a);
}

template <int KIND>
__global__ void __launch_bounds__(kWeightsBlock) k_weights_pair(WeightArgs a0, WeightArgs a1)
{
    weights_body<KIND>(blockIdx.y ? a1 : a0);
}

// Materialises the pending resampling: anc[j] for every output j, and the
// statistics of the resampled frame (block 0).
struct AncArgs {
    Pending pend;
    FrameScratch *fs;
    int *anc;
    double *lml;     // optional: written with the frame's log mean weight
    const int *lml_frame;   // optional: the slot is lml[*lml_frame] (read at run time, graph replays)
    const int *pow_frame;   // optional: frame t whose step recorded the largest block exponent in
                            // fs->pow_max[t mod 3] (else the resolver reduces it)
    int N, parity, kind_linear;
};

// The pending resampling of a, with the recorded largest block exponent when
// its producer wrote one.
__device__ __forceinline__ Pending anc_pending(const AncArgs &a)
{
    Pending pd = a.pend;
    if (a.pow_frame && pd.bpow) pd.pow_max = &a.fs->pow_max[*a.pow_frame % 3];
    return pd;
}

static __global__ void __launch_bounds__(256) k_ancestors(AncArgs a)
{
    extern __shared__ unsigned long long wsp[];
    const int j0 = blockIdx.x * blockDim.x;
    const int j = j0 + threadIdx.x;
    double es = 0.0;
    ResolveStats rs;
    const int src = resolve_ancestor(anc_pending(a), a.N, a.fs->U_pending, j0, wsp, blockIdx.x == 0 && a.lml, es, rs);
    if (j < a.N) a.anc[j] = src;
    if (blockIdx.x == 0 && threadIdx.x == 0)
        finish_resampling(a.fs, a.parity, es, rs, a.N, a.kind_linear,
                          (a.lml && a.lml_frame) ? a.lml + *a.lml_frame : a.lml);
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
    if ((threadIdx.x & 63) == 0 && isfinite(v)) atomicMax(&fs->maxkey[0], order_key(v));
}

// ---------------------------------------------------------------- helpers
static __global__ void k_set_int(int *p, int v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

static __global__ void k_iota(int *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

// my_gmm::resetTracker (my_gmm.cpp:30-42) from the device copy of mu / Sigma.
template <int D>
__global__ void __launch_bounds__(256) k_reset(const int *bins, const float *mu, const float *sig,
                                               float *x, float *P, int N, Lay L)
{
    constexpr int NP = npacked(D);
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    const int b = bins[j];
#pragma unroll
    for (int i = 0; i < D; ++i) x[lidx(L, i, j)] = mu[b * D + i];
#pragma unroll
    for (int p = 0; p < NP; ++p) P[lidx(L, p, j)] = sig[b * NP + p];
}

// Tracks in: x' = T x, P' = T sym(P) T^T (FP64), stored FP32 packed in layout L.
template <int D>
__global__ void __launch_bounds__(256) k_import(const double *x, const double *P, const double *T, float *xs,
                                                float *Ps, int N, Lay L)
{
    __shared__ double sT[D * D];
    for (int i = threadIdx.x; i < D * D; i += blockDim.x) sT[i] = T[i];
    __syncthreads();
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    const double *xj = x + (size_t)j * D, *Pj = P + (size_t)j * D * D;
    for (int i = 0; i < D; ++i) {
        double acc = 0.0;
        for (int k = 0; k < D; ++k) acc = fma(sT[i * D + k], xj[k], acc);
        xs[lidx(L, i, j)] = (float)acc;
    }
    for (int i = 0; i < D; ++i) {
        double r[D];
        for (int k = 0; k < D; ++k) {
            double acc = 0.0;
            for (int l = 0; l < D; ++l) acc = fma(sT[i * D + l], 0.5 * (Pj[l * D + k] + Pj[k * D + l]), acc);
            r[k] = acc;
        }
        for (int q = i; q < D; ++q) {
            double acc = 0.0;
            for (int k = 0; k < D; ++k) acc = fma(r[k], sT[q * D + k], acc);
            Ps[lidx(L, pk(D, i, q), j)] = (float)acc;
        }
    }
}

// Tracks out: tracks[j] = updated[anc_j] (the copy of pf2DRao.cpp:153-156),
// back in the model basis: x = Ti x', P = Ti P' Ti^T (FP64, row-major).
template <int D>
__global__ void __launch_bounds__(256) k_export(const float *xs, const float *Ps, const int *anc,
                                                const double *Ti, double *x, double *P, int N, Lay L)
{
    constexpr int NP = npacked(D);
    __shared__ double sT[D * D];
    for (int i = threadIdx.x; i < D * D; i += blockDim.x) sT[i] = Ti[i];
    __syncthreads();
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    const int s = anc ? anc[j] : j;
    double xp[D], pp[NP];
    for (int i = 0; i < D; ++i) xp[i] = (double)xs[lidx(L, i, s)];
    for (int p = 0; p < NP; ++p) pp[p] = (double)Ps[lidx(L, p, s)];
    double *xo = x + (size_t)j * D, *Po = P + (size_t)j * D * D;
    for (int i = 0; i < D; ++i) {
        double acc = 0.0;
        for (int k = 0; k < D; ++k) acc = fma(sT[i * D + k], xp[k], acc);
        xo[i] = acc;
    }
    for (int i = 0; i < D; ++i) {
        double r[D];
        for (int k = 0; k < D; ++k) {
            double acc = 0.0;
            for (int l = 0; l < D; ++l) acc = fma(sT[i * D + l], pp[pk(D, l, k)], acc);
            r[k] = acc;
        }
        for (int q = i; q < D; ++q) {
            double acc = 0.0;
            for (int k = 0; k < D; ++k) acc = fma(r[k], sT[q * D + k], acc);
            Po[i * D + q] = acc;
            Po[q * D + i] = acc;
        }
    }
}

// Row-parallel forms of k_import / k_export for large states: one thread per
// (particle, row i), FP64, the packed state read from global memory.
template <int D>
__global__ void __launch_bounds__(256) k_import_rows(const double *x, const double *P, const double *T, float *xs,
                                                     float *Ps, int N, Lay L)
{
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long long)N * D) return;
    const int j = (int)(g / D), i = (int)(g % D);
    const double *xj = x + (size_t)j * D, *Pj = P + (size_t)j * D * D;
    double acc = 0.0;
    for (int k = 0; k < D; ++k) acc = fma(T[i * D + k], xj[k], acc);
    xs[lidx(L, i, j)] = (float)acc;
    double r[D];
    for (int k = 0; k < D; ++k) {
        double a = 0.0;
        for (int l = 0; l < D; ++l) a = fma(T[i * D + l], 0.5 * (Pj[l * D + k] + Pj[k * D + l]), a);
        r[k] = a;
    }
    for (int q = i; q < D; ++q) {
        double a = 0.0;
        for (int k = 0; k < D; ++k) a = fma(r[k], T[q * D + k], a);
        Ps[lidx(L, pk(D, i, q), j)] = (float)a;
    }
}

template <int D>
__global__ void __launch_bounds__(256) k_export_rows(const float *xs, const float *Ps, const int *anc,
                                                     const double *Ti, double *x, double *P, int N, Lay L)
{
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long long)N * D) return;
    const int j = (int)(g / D), i = (int)(g % D);
    const int s = anc ? anc[j] : j;
    double acc = 0.0;
    for (int k = 0; k < D; ++k) acc = fma(Ti[i * D + k], (double)xs[lidx(L, k, s)], acc);
    x[(size_t)j * D + i] = acc;
    double r[D];
    for (int k = 0; k < D; ++k) {
        double a = 0.0;
        for (int l = 0; l < D; ++l)
            a = fma(Ti[i * D + l], (double)Ps[lidx(L, pk(D, l < k ? l : k, l < k ? k : l), s)], a);
        r[k] = a;
    }
    double *Po = P + (size_t)j * D * D;
    for (int q = 0; q < D; ++q) {
        double a = 0.0;
        for (int k = 0; k < D; ++k) a = fma(r[k], Ti[q * D + k], a);
        Po[i * D + q] = a;
    }
}

// Sum over the 256 threads of a block of D values per thread, in a fixed
// order: the values go to LDS, lane (i, k) adds the 16 values of threads
// 16k .. 16k + 15 for dimension i in order, one lane per dimension adds the
// 16 run sums in order.  Result in s_out (all threads, after the call).
template <int D>
__device__ __forceinline__ void block_sum256(const double (&v)[D], double (*s_v)[16][17], double (*s_run)[16],
                                             double *s_out)
{
    const int t = threadIdx.x;
#pragma unroll
    for (int i = 0; i < D; ++i) s_v[i][t >> 4][t & 15] = v[i];
    __syncthreads();
    for (int q = t; q < D * 16; q += 256) {
        const int i = q >> 4, k = q & 15;
        double r = 0.0;
#pragma unroll
        for (int u = 0; u < 16; ++u) r += s_v[i][k][u];
        s_run[i][k] = r;
    }
    __syncthreads();
    if (t < D) {
        double r = 0.0;
#pragma unroll
        for (int k = 0; k < 16; ++k) r += s_run[t][k];
        s_out[t] = r;
    }
    __syncthreads();
}

// getEstimator (pf2DRao.cpp:23-31): xbar = sum_j x[anc_j] / N, FP64
// accumulation in a fixed order: block sums (block_sum256) into partial,
// then k_estimate_final adds them (thread t takes blocks t, t + 256, ... in
// order, then block_sum256).  256 threads per block.
// x of particle s as FP64 (zeros if !ok): 16-byte loads from a record
// (row stride 1, D a multiple of 4), else one load per row.
template <int D>
__device__ __forceinline__ void load_x64(const float *x, const Lay &L, int s, bool ok, double (&v)[D])
{
    if constexpr (D % 4 == 0) {
        if (L.row == 1) {
            const float4 *r = reinterpret_cast<const float4 *>(x + (size_t)s * L.part);
#pragma unroll
            for (int c = 0; c < D / 4; ++c) {
                const float4 q = ok ? r[c] : make_float4(0.f, 0.f, 0.f, 0.f);
                v[4 * c] = q.x;
                v[4 * c + 1] = q.y;
                v[4 * c + 2] = q.z;
                v[4 * c + 3] = q.w;
            }
            return;
        }
    }
#pragma unroll
    for (int i = 0; i < D; ++i) v[i] = ok ? (double)x[lidx(L, i, s)] : 0.0;
}

template <int D>
__global__ void __launch_bounds__(256) k_estimate(const float *x, const int *anc, int N, Lay L, double *partial)
{
    __shared__ double s_v[D][16][17];
    __shared__ double s_run[D][16];
    __shared__ double s_sum[D];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    double v[D];
    const int s = (j < N) ? (anc ? anc[j] : j) : 0;
    load_x64<D>(x, L, s, j < N, v);
    block_sum256<D>(v, s_v, s_run, s_sum);
    if (threadIdx.x < D) partial[(size_t)blockIdx.x * D + threadIdx.x] = s_sum[threadIdx.x];
}

// Sum of the nb block partials, / N, back to the model basis (xbar = Ti xbar').
template <int D>
__device__ __forceinline__ void estimate_final_body(const double *partial, int nb, int N, const double *Ti,
                                                    double *out)
{
    __shared__ double s_v[D][16][17];
    __shared__ double s_run[D][16];
    __shared__ double s_sum[D];
    double v[D];
#pragma unroll
    for (int i = 0; i < D; ++i) v[i] = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
#pragma unroll
        for (int i = 0; i < D; ++i) v[i] += partial[(size_t)b * D + i];
    }
    block_sum256<D>(v, s_v, s_run, s_sum);
    if (threadIdx.x < D) {
        double r = 0.0;
        for (int k = 0; k < D; ++k) r = fma(Ti[threadIdx.x * D + k], s_sum[k] / (double)N, r);
        out[threadIdx.x] = r;
    }
}

template <int D>
__global__ void __launch_bounds__(256) k_estimate_final(const double *partial, int nb, int N, const double *Ti,
                                                        double *out)
{
    estimate_final_body<D>(partial, nb, N, Ti, out);
}

// The estimates of two filters (block b = filter b).
struct EstFinalArgs {
    const double *partial;
    const double *Ti;
    double *out;
    int nb, N;
};

template <int D>
__global__ void __launch_bounds__(256) k_estimate_final_pair(EstFinalArgs e0, EstFinalArgs e1)
{
    const EstFinalArgs &e = blockIdx.x ? e1 : e0;
    estimate_final_body<D>(e.partial, e.nb, e.N, e.Ti, e.out);
}

// k_ancestors, then the block sums of getEstimator over the resampled state
// (x[anc_j], summed as k_estimate does; the resolver's LDS is reused for the
// reduction).  Followed by k_estimate_final.  Dynamic LDS:
// anc_estimate_lds_bytes(nb, D).
template <int D>
__device__ __forceinline__ void ancestors_estimate_body(const AncArgs &a, const float *x, Lay L, double *partial)
{
    extern __shared__ unsigned long long wsp[];
    const int j0 = blockIdx.x * blockDim.x;
    const int j = j0 + threadIdx.x;
    double es = 0.0;
    ResolveStats rs;
    const int src = resolve_ancestor(anc_pending(a), a.N, a.fs->U_pending, j0, wsp, blockIdx.x == 0 && a.lml, es, rs);
    if (j < a.N) a.anc[j] = src;
    if (blockIdx.x == 0 && threadIdx.x == 0)
        finish_resampling(a.fs, a.parity, es, rs, a.N, a.kind_linear,
                          (a.lml && a.lml_frame) ? a.lml + *a.lml_frame : a.lml);
    double v[D];
    load_x64<D>(x, L, src, j < a.N, v);
    __syncthreads();                     // every lane is done with the resolver's LDS
    double *ws = reinterpret_cast<double *>(wsp);
    block_sum256<D>(v, reinterpret_cast<double (*)[16][17]>(ws), reinterpret_cast<double (*)[16]>(ws + D * 16 * 17),
                    ws + D * 16 * 17 + D * 16);
    if (threadIdx.x < D) partial[(size_t)blockIdx.x * D + threadIdx.x] = ws[D * 16 * 17 + D * 16 + threadIdx.x];
}

template <int D>
__global__ void __launch_bounds__(256) k_ancestors_estimate(AncArgs a, const float *x, Lay L, double *partial)
{
    ancestors_estimate_body<D>(a, x, L, partial);
}

struct AncEstArgs {
    AncArgs an;
    const float *x;     // the resampled filter's updated state (layout lay)
    double *partial;    // block sums of the estimate
    Lay lay;
};

template <int D>
__global__ void __launch_bounds__(256) k_ancestors_estimate_pair(AncEstArgs e0, AncEstArgs e1)
{
    const AncEstArgs &e = blockIdx.y ? e1 : e0;
    ancestors_estimate_body<D>(e.an, e.x, e.lay, e.partial);
}

__host__ __device__ constexpr int anc_estimate_lds_bytes(int nb, int D)
{
    return 8 * resolve_lds_words(nb) > 8 * (D * 16 * 17 + D * 16 + D) ? 8 * resolve_lds_words(nb)
                                                                        : 8 * (D * 16 * 17 + D * 16 + D);
}

// Both estimate kernels on one stream.
template <int D>
inline void launch_estimate(hipStream_t st, const float *x, const int *anc, int N, Lay L, double *partial,
                            const double *Ti, double *out)
{
    const int nb = (N + 255) / 256;
    hipLaunchKernelGGL((k_estimate<D>), dim3(nb), dim3(256), 0, st, x, anc, N, L, partial);
    hipLaunchKernelGGL((k_estimate_final<D>), dim3(1), dim3(256), 0, st, (const double *)partial, nb, N, Ti, out);
}

// getSamples (pf2DRao.cpp:85-103): out = y[0:2] + 0.8 scale n, y = Hp xbar + M.
static __global__ void k_samples(const double *hp2, const double *m2, const double *xbar, int d,
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
static __global__ void k_sample_prob(const double *hp2, const double *m2, const double *xbar, int d,
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
static __global__ void k_kf_predict(int d, const double *F, const double *B, const double *Q, double *x,
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

static __global__ void k_kf_update(int d, int m, const double *H, const double *BH, const double *R,
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
