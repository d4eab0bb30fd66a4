// pf_batch_impl.hpp -- internals of the batched engine (pf_batch.hip) shared
// with its adjoint (pf_adjoint.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/pf_batch.h"
#include "pf_internal.hpp"
#include "pf_kernels.hpp"
#include "pf_model.hpp"

#define HIP_TRYB(expr)                                                                         \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return mkfbt::set_error(PF_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

namespace mkfbt {
namespace batch {

constexpr int kBD = 12;           // state dimension of the batched engine
constexpr int kSeqAlign = 512;    // particle offset alignment (largest tile)

struct SeqDesc {
    int pofs, N, T, tofs, bofs, nb, table_ofs, pad;
    unsigned long long pi_total;
};

struct BatchArgs {
    StepArgs base;                // constants shared by every sequence (K, m, J, cap, PDAF constants, image sizes)
    float *buf_x[2], *buf_P[2];
    int *anc;
    float *lw;
    unsigned long long *maxkey;   // 2 x S (parity-major)
    unsigned int *status;         // S
    // resampler weights of frame t in set t & 1 (frame t + 1 resolves them
    // while the weights of its own frame go to the other set): block-local
    // scans (ld), block sums, FP64 block sums and exponents (per block)
    unsigned long long *cloc, *bsum, *cloc1, *bsum1;
    double *bexp, *bexp1;
    int *bpow, *bpow1;
    const float *tables;
    const SeqDesc *seqs;
    const int *tile_seq, *tile_j0, *n_tiles;   // tiles of active sequences first
    const int *wb_seq, *wb_b, *n_wblocks;      // resampler blocks likewise
    const int *frame;
    int *frame_next;
    const uint32_t *uni;
    const int *count;
    const float *cand;
    double *lml;
    int comp_head;                // >= 0: component rows read from global memory at tables + table_ofs + comp_head
    int *rec_comp;                // optional record: components of frame t at rec_comp[t * ld + column]
    int *rec_anc;                 // optional record: ancestors of frame t's resampling, same layout
    int S, ld, K, m, J, cap, table_floats, stage_floats, parity, fused, nb_max, all_tiles;
    int wfused;                   // the step writes its frame's resampler weights (256-particle tiles, one per lane)
    unsigned int order_packed, clutter_mask;
    float pdl, b0, gate;
};

// A block-uniform descriptor read from global memory, moved to scalar
// registers (the loads are vector loads, so the values would otherwise
// occupy vector registers in every lane).
__device__ __forceinline__ int uni_int(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ SeqDesc uniform_seq(const SeqDesc *seqs, int s)
{
    const SeqDesc d = seqs[s];
    SeqDesc r;
    r.pofs = uni_int(d.pofs);
    r.N = uni_int(d.N);
    r.T = uni_int(d.T);
    r.tofs = uni_int(d.tofs);
    r.bofs = uni_int(d.bofs);
    r.nb = uni_int(d.nb);
    r.table_ofs = uni_int(d.table_ofs);
    r.pad = 0;
    const unsigned lo = (unsigned)uni_int((int)(unsigned)d.pi_total);
    const unsigned hi = (unsigned)uni_int((int)(unsigned)(d.pi_total >> 32));
    r.pi_total = ((unsigned long long)hi << 32) | lo;
    return r;
}

// Block index -> tile index so that each XCD works on a contiguous range of
// tiles.  Workgroups are dispatched to the 8 XCDs round robin (block b on XCD
// b mod 8); with this bijection of [0, gridDim.x) the blocks of one XCD take
// consecutive tiles, i.e. few sequences, whose model rows then stay in that
// XCD's L2 (the d = 24, K = 64 images are read from global memory).
__device__ __forceinline__ int xcd_tile()
{
    const int G = (int)gridDim.x, b = (int)blockIdx.x, q = G >> 3, r = G & 7, c = b & 7, i = b >> 3;
    return c < r ? c * (q + 1) + i : r * (q + 1) + (c - r) * q + i;
}

// The resampler weights of frame t (set t & 1).
struct WeightSet {
    unsigned long long *cloc, *bsum;
    double *bexp;
    int *bpow;
};
__device__ __forceinline__ WeightSet weight_set(const BatchArgs &a, int t)
{
    const bool o = (t & 1) != 0;     // selects, not indexes
    return {o ? a.cloc1 : a.cloc, o ? a.bsum1 : a.bsum, o ? a.bexp1 : a.bexp, o ? a.bpow1 : a.bpow};
}

__device__ __forceinline__ const float *seq_comp_global(const BatchArgs &a, const SeqDesc &sd)
{
    return a.comp_head >= 0 ? a.tables + sd.table_ofs + a.comp_head : nullptr;
}

// Per-sequence view of the arguments in the single-filter layout, so the
// per-particle update is shared with k_step.
__device__ __forceinline__ StepArgs seq_view(const BatchArgs &a, const SeqDesc &sd)
{
    StepArgs v{};
    v.count = a.count + (size_t)sd.tofs * a.J;
    v.cand = a.cand + (size_t)sd.tofs * a.J * a.cap * 2;
    v.pi_total = sd.pi_total;
    v.table_floats = a.table_floats;
    v.stage_floats = a.stage_floats;
    v.K = a.K;
    v.N = sd.N;
    v.m = a.m;
    v.J = a.J;
    v.cap = a.cap;
    v.order_packed = a.order_packed;
    v.clutter_mask = a.clutter_mask;
    v.pdl = a.pdl;
    v.b0 = a.b0;
    v.gate = a.gate;
    v.comp_out = nullptr;
    v.comp_global = a.comp_head >= 0 ? a.tables + sd.table_ofs + a.comp_head : nullptr;
    return v;
}


template <typename T>
hipError_t dallocb(T **p, size_t n)
{
    return hipMalloc((void **)p, sizeof(T) * std::max<size_t>(n, 1));
}

}  // namespace batch
}  // namespace mkfbt

struct pf_batch {
    int device = 0, S = 0, K = 0, D = 0, m = 0, J = 0, cap = 16, cus = 256;
    double pd = 0.95, gate = 9.210340371976184, lambda = 3.0 / (640.0 * 480.0);
    std::vector<int> N, T, order, rank;      // order[i] = user sequence of internal slot i; rank = inverse
    std::vector<int> pofs, tofs_user, tofs, bofs, nb;
    int ld = 0, Tmax = 0, nb_max = 0;
    std::vector<mkfbt::HostModel> hm;
    std::vector<char> loaded, sel_ok, fresh;   // fresh[s]: reset since the last run
    unsigned clutter_mask = 0, order_packed = 0;
    int table_floats = 0;          // floats of each sequence's image staged in LDS
    int comp_head = -1;            // >= 0: component rows read from global memory (large images)
    int stage_floats() const { return table_floats + mkfbt::cand_lds_floats(J, cap); }
    bool tracks = false;
    pf_batch_schedule sch{256, 1, 2};
    int variant = PF_BATCH_MFMA;   // kernel of the 24-dimensional state
    hipStream_t stream = nullptr;
    float *bx[2] = {nullptr, nullptr}, *bP[2] = {nullptr, nullptr};
    float *lw = nullptr, *tables = nullptr, *mus = nullptr, *sigs = nullptr;
    double *Tis = nullptr, *Ts = nullptr, *bexp = nullptr, *bexp1 = nullptr, *partial = nullptr, *est = nullptr;
    double *gbuf = nullptr;
    int *bpow = nullptr, *bpow1 = nullptr;
    size_t gbuf_bytes = 0;
    int *anc = nullptr, *frame = nullptr;
    unsigned long long *maxkey = nullptr, *cloc = nullptr, *bsum = nullptr, *cloc1 = nullptr, *bsum1 = nullptr;
    unsigned int *status = nullptr;
    mkfbt::batch::SeqDesc *seqs = nullptr;
    std::vector<mkfbt::batch::SeqDesc> h_seqs;
    int *tile_seq = nullptr, *tile_j0 = nullptr, *n_tiles = nullptr, all_tiles = 0;
    int *wb_seq = nullptr, *wb_b = nullptr, *n_wblocks = nullptr;
    mkfbt::FrameScratch *fs = nullptr;   // for the per-sequence estimate reduction
    float last_ms = 0.f;
    std::vector<int> h_ntiles;    // active tiles per frame (current schedule)
    // adjoint (pf_batch_lml_grad): records and checkpoints of the last recorded run
    int *rec_comp = nullptr, *rec_anc = nullptr, *rec_bins = nullptr;
    int rec_frames = 0, ckpt_every = 0, n_ckpt = 0, seg_slots = 0;
    float *ck_x = nullptr, *ck_P = nullptr;
    float *seg_x[2] = {nullptr, nullptr}, *seg_P[2] = {nullptr, nullptr};   // by segment parity
    hipStream_t stream2 = nullptr;            // the adjoint's replay stream (default priority)
    std::vector<hipEvent_t> ev_replay, ev_back;
    hipEvent_t ev_fwd = nullptr;
    float *adj_x[2] = {nullptr, nullptr}, *adj_P[2] = {nullptr, nullptr};
    int *first = nullptr;
    double *tacc = nullptr, *gsum = nullptr;
    int *t256_seq = nullptr, *t256_j0 = nullptr, *t256_first = nullptr;
    std::vector<int> h_t256_count;
    int t256_total = 0;
    bool recorded = false;
    double last_fwd_ms = 0.0, last_bwd_ms = 0.0;
    std::vector<size_t> adj_have = std::vector<size_t>(24, 0);   // element capacities of the buffers above
    std::vector<double> h_T;      // state bases per sequence (internal order), D x D each
};


namespace mkfbt {
namespace batch {

int bcheck(pf_batch *b);
int build_tiles(pf_batch *b);
BatchArgs batch_args(pf_batch *b, const int32_t *count, const float *cand, const uint32_t *uni, double *lml,
                     int parity, int fused);
int launch_batch_step(pf_batch *b, const BatchArgs &a);
// Whether the d = 12 step writes its frame's resampler weights itself: its
// tiles are resampler blocks (256 threads, one particle per lane); else
// k_batch_weights runs after it.
inline bool step_writes_weights(const pf_batch *b)
{
    return b->D == kBD && b->sch.block == kWeightsBlock && b->sch.ppl == 1;
}
int launch_batch_weights(pf_batch *b, const BatchArgs &a);
int launch_batch_finish(pf_batch *b, const BatchArgs &a);
// Frames [0, frames) of every sequence from its reset state.  ckpt: copy the
// input state of every frame t with t % ckpt_every == 0 into checkpoint
// t / ckpt_every (no graph capture then).
int run_all(pf_batch *b, const int32_t *count, const float *cand, const uint32_t *uni, double *lml, int frames,
            bool graph, bool record = false);

}  // namespace batch
}  // namespace mkfbt
