// pf_batch.hip -- many independent RBPF sequences advanced together on one
// gfx950 GPU (C ABI in include/pf_batch.h).
//
// Layout: all sequences share one structure-of-arrays state (x[D][ld],
// P[D(D+1)/2][ld], double buffered), sequence s owning particles
// [pofs_s, pofs_s + N_s) with pofs_s a multiple of 512.  Sequences are
// ordered by length (longest first), so at frame t the active sequences are
// a prefix and so are their tiles and resampler blocks; the persistent
// kernels read the active counts for frame t from device memory, which makes
// one captured frame pair replayable for the whole run.
//
// Per frame: k_batch_step (ancestors from the previous frame's weights, the
// same per-particle update as k_step) and the frame's per-sequence
// fixed-point weights and block scans: written by the step itself when its
// tiles are resampler blocks (256 threads, one particle each), else by
// k_batch_weights.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "../../include/pf_batch.h"
#include "pf_internal.hpp"
#include "pf_kernels.hpp"
#include "pf_model.hpp"
#include "pf_batch_impl.hpp"

using namespace mkfbt;
using namespace mkfbt::batch;

namespace {

// Ancestors of a tile's outputs from the resampling of frame tw's weights
// (set tw & 1, uniform uni[tw][1]); the sequence's first tile also writes
// the frame's log mean weight and the degenerate flag.
template <int PPL>
__device__ __forceinline__ void tile_ancestors(const BatchArgs &a, const SeqDesc &sd, int s, int tw, int jt0,
                                               unsigned long long *wsp, int (&src)[PPL], int *anc_out,
                                               int *anc_out2, double *lml_slot)
{
    const int bd = blockDim.x;
    const WeightSet ws = weight_set(a, tw);
    Pending pd;
    pd.cloc = ws.cloc + sd.pofs;
    pd.bsum = ws.bsum + sd.bofs;
    pd.bexp = ws.bexp + sd.bofs;
    pd.bpow = ws.bpow + sd.bofs;
    pd.L = sd.N;
    pd.nb = sd.nb;
    double es = 0.0;
    const bool first = (jt0 == 0);
    int E;
    const unsigned long long total = resolve_prefix(pd, wsp, first, es, E);
    const uint32_t U = a.uni[(size_t)(sd.tofs + tw) * 2 + 1];
#pragma unroll
    for (int r = 0; r < PPL; ++r) {
        src[r] = 0;
        if (jt0 + r * bd >= sd.N) continue;        // block-uniform: no barrier is skipped by part of the block
        src[r] = resolve_search(pd, sd.N, U, total, jt0 + r * bd, wsp);
        const int j = jt0 + r * bd + (int)threadIdx.x;
        if (j < sd.N) {
            if (anc_out) anc_out[sd.pofs + j] = src[r];
            if (anc_out2) anc_out2[sd.pofs + j] = src[r];
        }
    }
    if (first && threadIdx.x == 0) {
        if (total == 0ull) atomicOr(&a.status[s], kStatusDegenerate);
        if (lml_slot) *lml_slot = (total == 0ull) ? -INFINITY : pow_ref(E) + log(es) - log((double)sd.N);
    }
}

// PERSIST: a grid of blocks_per_cu x CUs walks the active tiles; otherwise
// one block per tile (grid = all tiles, the inactive ones exit at once),
// which keeps the 24-dimensional state's register footprint free of the
// loop's live values.
template <int D, int SELJ, int PPL, bool PERSIST = true>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu((D > 16 || PPL > 1) ? 1 : 2, 8)))
k_batch_step(BatchArgs a)
{
    constexpr int NP = npacked(D);
    extern __shared__ float4 sm4[];
    const int t = *a.frame;
    const int ntiles = a.n_tiles[t];
    const int bd = blockDim.x;
    const bool odd = (t & 1) != 0;      // selects, not indexes: a dynamic index would copy the array to scratch
    const float *xin = odd ? a.buf_x[1] : a.buf_x[0], *Pin = odd ? a.buf_P[1] : a.buf_P[0];
    float *xout = odd ? a.buf_x[0] : a.buf_x[1], *Pout = odd ? a.buf_P[0] : a.buf_P[1];
    const unsigned rowb = (unsigned)a.ld * 4u;
    unsigned long long *wsp = reinterpret_cast<unsigned long long *>(reinterpret_cast<float *>(sm4) + a.stage_floats);
    if (a.wfused && blockIdx.x == 0 && threadIdx.x == 0) *a.frame_next = t + 1;
    for (int tile = PERSIST ? (int)blockIdx.x : xcd_tile(); tile < ntiles; tile += PERSIST ? gridDim.x : ntiles) {
        const int s = uni_int(a.tile_seq[tile]);
        const int jt0 = uni_int(a.tile_j0[tile]);
        const SeqDesc sd = uniform_seq(a.seqs, s);
        int src[PPL];
        if (a.fused) {
            tile_ancestors<PPL>(a, sd, s, t - 1, jt0, wsp, src,
                                a.rec_anc ? a.rec_anc + (size_t)(t - 1) * a.ld : nullptr, nullptr,
                                a.lml + sd.tofs + (t - 1));
        } else {
#pragma unroll
            for (int r = 0; r < PPL; ++r) {
                const int j = jt0 + r * bd + (int)threadIdx.x;
                src[r] = j < sd.N ? a.anc[sd.pofs + j] : 0;
            }
        }
        // state loads of all PPL particles, then the model staging
        float x[PPL][D], P[PPL][NP];
        const __amdgpu_buffer_rsrc_t rx = soa_rsrc(xin, rowb * D), rP = soa_rsrc(Pin, rowb * NP);
#pragma unroll
        for (int r = 0; r < PPL; ++r) {
            const int j = jt0 + r * bd + (int)threadIdx.x;
            if (j < sd.N) {
                const unsigned voff = (unsigned)(sd.pofs + src[r]) * 4u;
                if constexpr (D > 16) {   // row offsets in the vector offset: no scalar register per row
#pragma unroll
                    for (int i = 0; i < D; ++i) x[r][i] = soa_ld(rx, voff + rowb * i, 0);
#pragma unroll
                    for (int p = 0; p < NP; ++p) P[r][p] = soa_ld(rP, voff + rowb * p, 0);
                } else {
#pragma unroll
                    for (int i = 0; i < D; ++i) x[r][i] = soa_ld(rx, voff, rowb * i);
#pragma unroll
                    for (int p = 0; p < NP; ++p) P[r][p] = soa_ld(rP, voff, rowb * p);
                }
            }
        }
        const float *fcand = a.cand + ((size_t)sd.tofs + t) * a.J * a.cap * 2;
        CandPrefetch cpf;
        cand_prefetch(cpf, a.count + ((size_t)sd.tofs + t) * a.J, fcand, a.J, a.cap);
        __syncthreads();          // the previous tile's LDS readers are done
        stage_table(sm4, reinterpret_cast<const float4 *>(a.tables + (size_t)sd.table_ofs), a.table_floats >> 2);
        cand_store(cpf, reinterpret_cast<float *>(sm4) + a.table_floats, fcand, a.J, a.cap);
        __syncthreads();
        int *comp_out = a.rec_comp ? a.rec_comp + (size_t)t * a.ld + sd.pofs : nullptr;
        const float *comp_global = seq_comp_global(a, sd);
        const uint32_t U1 = a.uni[(size_t)(sd.tofs + t) * 2];
        const __amdgpu_buffer_rsrc_t wx = soa_rsrc(xout, rowb * D), wP = soa_rsrc(Pout, rowb * NP);
        unsigned pflags = 0;
        float lw_last = -INFINITY;
#pragma unroll
        for (int r = 0; r < PPL; ++r) {
            const int j = jt0 + r * bd + (int)threadIdx.x;
            if (j < sd.N) {
                unsigned f = 0;
                float lwj = particle_update<D, kModeCand, SELJ>(a.base, t, j, U1, reinterpret_cast<const float *>(sm4),
                                                                x[r], P[r], sd.pi_total, comp_global, comp_out, f);
                const unsigned doff = (unsigned)(sd.pofs + j) * 4u;
                auto store = [&]() {
                    if constexpr (D > 16) {
#pragma unroll
                        for (int i = 0; i < D; ++i) soa_st(wx, doff + rowb * i, 0, x[r][i]);
#pragma unroll
                        for (int p = 0; p < NP; ++p) soa_st(wP, doff + rowb * p, 0, P[r][p]);
                    } else {
#pragma unroll
                        for (int i = 0; i < D; ++i) soa_st(wx, doff, rowb * i, x[r][i]);
#pragma unroll
                        for (int p = 0; p < NP; ++p) soa_st(wP, doff, rowb * p, P[r][p]);
                    }
                };
                store();
                pflags |= f;
                a.lw[sd.pofs + j] = lwj;
                lw_last = lwj;
            }
        }
        report_flags(pflags, &a.status[s]);
        if constexpr (PPL == 1) {
            if (a.wfused) {
                // the tile is resampler block jt0 / 256 of its sequence
                const int j = jt0 + (int)threadIdx.x;
                const bool valid = j < sd.N;
                const WeightSet ws = weight_set(a, t);
                const int wb = jt0 / kWeightsBlock;
                (void)block_weights_log((double)lw_last, valid, ws.cloc + sd.pofs + (valid ? j : 0),
                                        ws.bsum + sd.bofs + wb, ws.bexp + sd.bofs + wb, ws.bpow + sd.bofs + wb);
            }
        }
    }
}

// Fixed-point weights and block scans, per sequence (kWeightsBlock blocks):
// the block-scaled log-weights of the resampler (block_weights_log).
__global__ void __launch_bounds__(kWeightsBlock) k_batch_weights(BatchArgs a)
{
    const int t = *a.frame;
    const int nwb = a.n_wblocks[t];
    if (blockIdx.x == 0 && threadIdx.x == 0) *a.frame_next = t + 1;
    for (int wb = blockIdx.x; wb < nwb; wb += gridDim.x) {
        const int s = a.wb_seq[wb], b = a.wb_b[wb];
        const SeqDesc sd = a.seqs[s];
        const int i = b * kWeightsBlock + (int)threadIdx.x;
        const bool valid = i < sd.N;
        const double x = valid ? (double)a.lw[sd.pofs + i] : -INFINITY;
        const WeightSet ws = weight_set(a, t);
        block_weights_log(x, valid, ws.cloc + sd.pofs + (valid ? i : 0), ws.bsum + sd.bofs + b, ws.bexp + sd.bofs + b,
                          ws.bpow + sd.bofs + b);
    }
}

// Materialises every sequence's last resampling: ancestors and the log mean
// weight of its last frame.
template <int PPL>
__global__ void __launch_bounds__(256) k_batch_finish(BatchArgs a)
{
    extern __shared__ float4 sm4[];
    unsigned long long *wsp = reinterpret_cast<unsigned long long *>(sm4);
    for (int tile = blockIdx.x; tile < a.all_tiles; tile += gridDim.x) {
        const int s = uni_int(a.tile_seq[tile]);
        const int jt0 = uni_int(a.tile_j0[tile]);
        const SeqDesc sd = uniform_seq(a.seqs, s);
        int src[PPL];
        tile_ancestors<PPL>(a, sd, s, sd.T - 1, jt0, wsp, src, a.anc,
                            a.rec_anc ? a.rec_anc + (size_t)(sd.T - 1) * a.ld : nullptr, a.lml + sd.tofs + sd.T - 1);
    }
}

// Start of a run: per-sequence maxima and flags cleared (a kernel, so the
// clear may sit in a captured stream).
__global__ void k_batch_clear(unsigned long long *maxkey, unsigned int *status, int S)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 2 * S) maxkey[i] = 0ull;
    if (i < S) status[i] = 0u;
}

__global__ void k_identity_seq(int *anc, int pofs, int N)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < N) anc[pofs + j] = j;
}


}  // namespace

// the four-lanes-per-particle step of the 24-dimensional state (uses
// tile_ancestors above)
#include "pf_quad24.hpp"

namespace mkfbt {
namespace batch {

int bcheck(pf_batch *b)
{
    if (!b) return set_error(PF_ERR_ARG, "null batch");
    HIP_TRYB(hipSetDevice(b->device));
    return PF_OK;
}

// Host-side layout of a batch: sequences in rank order (longest first),
// their particle, block and frame offsets, and for a tile size the tile and
// weight-block tables with the active counts of every frame.
struct BatchLayout {
    std::vector<int> order;                 // rank -> sequence
    std::vector<SeqDesc> seqs;              // by rank
    int ld = 0, nb_max = 0, Tmax = 0, nblocks = 0;
    std::vector<int> ts, tj, wbs, wbb, nt, nw;
};

int make_layout(int S, const int32_t *N, const int32_t *T, BatchLayout &L)
{
    if (S <= 0 || S > (1 << 16)) return set_error(PF_ERR_ARG, "S=%d out of range", S);
    for (int s = 0; s < S; ++s)
        if (N[s] <= 0 || N[s] > kMaxScanBlocks * kWeightsBlock || T[s] <= 0)
            return set_error(PF_ERR_ARG, "sequence %d: bad size (N=%d, T=%d)", s, N[s], T[s]);
    L.order.resize(S);
    std::iota(L.order.begin(), L.order.end(), 0);
    std::stable_sort(L.order.begin(), L.order.end(), [&](int p, int q) { return T[p] > T[q]; });
    std::vector<int> tofs(S);
    for (int s = 0, acc = 0; s < S; ++s) { tofs[s] = acc; acc += T[s]; }
    L.seqs.assign(S, SeqDesc{});
    long long pofs = 0, bofs = 0;
    for (int i = 0; i < S; ++i) {
        const int s = L.order[i];
        SeqDesc &sd = L.seqs[i];
        sd.pofs = (int)pofs;
        sd.N = N[s];
        sd.T = T[s];
        sd.tofs = tofs[s];
        sd.nb = (sd.N + kWeightsBlock - 1) / kWeightsBlock;
        sd.bofs = (int)bofs;
        pofs += (sd.N + kSeqAlign - 1) / kSeqAlign * kSeqAlign;
        bofs += sd.nb;
        L.nb_max = std::max(L.nb_max, sd.nb);
        L.Tmax = std::max(L.Tmax, sd.T);
    }
    if (pofs > (1LL << 28)) return set_error(PF_ERR_ARG, "batch too large (%lld particles)", pofs);
    L.ld = (int)pofs;
    L.nblocks = (int)bofs;
    return PF_OK;
}

void make_tiles(BatchLayout &L, int tile)
{
    const int S = (int)L.seqs.size();
    L.ts.clear(); L.tj.clear(); L.wbs.clear(); L.wbb.clear();
    std::vector<int> first_tile(S + 1), first_wb(S + 1);
    for (int i = 0; i < S; ++i) {
        first_tile[i] = (int)L.ts.size();
        first_wb[i] = (int)L.wbs.size();
        const int n = L.seqs[i].N;
        for (int j0 = 0; j0 < n; j0 += tile) { L.ts.push_back(i); L.tj.push_back(j0); }
        for (int k = 0; k * kWeightsBlock < n; ++k) { L.wbs.push_back(i); L.wbb.push_back(k); }
    }
    first_tile[S] = (int)L.ts.size();
    first_wb[S] = (int)L.wbs.size();
    L.nt.assign((size_t)L.Tmax + 1, 0);
    L.nw.assign((size_t)L.Tmax + 1, 0);
    for (int t = 0; t <= L.Tmax; ++t) {
        int act = 0;                       // sequences are sorted by length, longest first
        while (act < S && L.seqs[act].T > t) ++act;
        L.nt[t] = first_tile[act];
        L.nw[t] = first_wb[act];
    }
}

// The invariants the kernels' indexing relies on (every index they form from
// these tables stays inside the buffers pf_batch_create allocates): disjoint,
// aligned particle ranges inside ld; disjoint block ranges; every tile and
// weight block inside its sequence, each particle and block covered once; at
// frame t only tiles and blocks of sequences with T > t are active.
int check_layout(const BatchLayout &L, int tile)
{
    const int S = (int)L.seqs.size();
    long long pend = 0, bend = 0, tsum = 0;
    for (int i = 0; i < S; ++i) {
        const SeqDesc &sd = L.seqs[i];
        if (sd.pofs % kSeqAlign || sd.pofs < pend || (long long)sd.pofs + sd.N > L.ld)
            return set_error(PF_ERR_STATE, "layout: particle range of rank %d", i);
        if (sd.bofs != bend || sd.nb != (sd.N + kWeightsBlock - 1) / kWeightsBlock || sd.nb > L.nb_max ||
            sd.bofs + sd.nb > L.nblocks)
            return set_error(PF_ERR_STATE, "layout: block range of rank %d", i);
        if (i > 0 && sd.T > L.seqs[i - 1].T) return set_error(PF_ERR_STATE, "layout: ranks not by length");
        pend = (long long)sd.pofs + sd.N;
        bend = (long long)sd.bofs + sd.nb;
        tsum += sd.T;
    }
    for (int i = 0; i < S; ++i)
        if ((long long)L.seqs[i].tofs + L.seqs[i].T > tsum) return set_error(PF_ERR_STATE, "layout: frame offsets");
    std::vector<long long> covered(S, 0), bcovered(S, 0);
    for (size_t k = 0; k < L.ts.size(); ++k) {
        const int i = L.ts[k], j0 = L.tj[k];
        if (i < 0 || i >= S || (k > 0 && i < L.ts[k - 1])) return set_error(PF_ERR_STATE, "layout: tile %zu sequence", k);
        if (j0 != covered[i] || j0 % tile || j0 >= L.seqs[i].N) return set_error(PF_ERR_STATE, "layout: tile %zu", k);
        covered[i] = std::min<long long>(L.seqs[i].N, (long long)j0 + tile);
    }
    for (size_t k = 0; k < L.wbs.size(); ++k) {
        const int i = L.wbs[k], bb = L.wbb[k];
        if (i < 0 || i >= S || bb != bcovered[i] || bb >= L.seqs[i].nb)
            return set_error(PF_ERR_STATE, "layout: weight block %zu", k);
        ++bcovered[i];
    }
    for (int i = 0; i < S; ++i)
        if (covered[i] != L.seqs[i].N || bcovered[i] != L.seqs[i].nb)
            return set_error(PF_ERR_STATE, "layout: rank %d not covered", i);
    for (int t = 0; t <= L.Tmax; ++t) {
        const int nt = L.nt[t], nw = L.nw[t];
        if (nt < 0 || nt > (int)L.ts.size() || nw < 0 || nw > (int)L.wbs.size() ||
            (t > 0 && (nt > L.nt[t - 1] || nw > L.nw[t - 1])))
            return set_error(PF_ERR_STATE, "layout: active counts at frame %d", t);
        if (nt > 0 && L.seqs[L.ts[nt - 1]].T <= t) return set_error(PF_ERR_STATE, "layout: inactive tile at %d", t);
        if (nw > 0 && L.seqs[L.wbs[nw - 1]].T <= t) return set_error(PF_ERR_STATE, "layout: inactive block at %d", t);
        if (nt < (int)L.ts.size() && L.seqs[L.ts[nt]].T > t) return set_error(PF_ERR_STATE, "layout: active tile skipped at %d", t);
        if (nw < (int)L.wbs.size() && L.seqs[L.wbs[nw]].T > t)
            return set_error(PF_ERR_STATE, "layout: active block skipped at %d", t);
    }
    return PF_OK;
}

// Tile tables for the current schedule (tile = block * ppl particles).
int build_tiles(pf_batch *b)
{
    const int tile = b->sch.block * b->sch.ppl;
    BatchLayout L;
    int st = make_layout(b->S, b->N.data(), b->T.data(), L);
    if (st) return st;
    make_tiles(L, tile);
    st = check_layout(L, tile);
    if (st) return st;
    const std::vector<int> &ts = L.ts, &tj = L.tj, &wbs = L.wbs, &wbb = L.wbb, &nt = L.nt, &nw = L.nw;
    b->all_tiles = (int)ts.size();
    b->h_ntiles = nt;
    (void)hipFree(b->tile_seq); (void)hipFree(b->tile_j0); (void)hipFree(b->n_tiles);
    (void)hipFree(b->wb_seq); (void)hipFree(b->wb_b); (void)hipFree(b->n_wblocks);
    HIP_TRYB(dallocb(&b->tile_seq, ts.size()));
    HIP_TRYB(dallocb(&b->tile_j0, tj.size()));
    HIP_TRYB(dallocb(&b->n_tiles, nt.size()));
    HIP_TRYB(dallocb(&b->wb_seq, wbs.size()));
    HIP_TRYB(dallocb(&b->wb_b, wbb.size()));
    HIP_TRYB(dallocb(&b->n_wblocks, nw.size()));
    HIP_TRYB(hipMemcpy(b->tile_seq, ts.data(), sizeof(int) * ts.size(), hipMemcpyHostToDevice));
    HIP_TRYB(hipMemcpy(b->tile_j0, tj.data(), sizeof(int) * tj.size(), hipMemcpyHostToDevice));
    HIP_TRYB(hipMemcpy(b->n_tiles, nt.data(), sizeof(int) * nt.size(), hipMemcpyHostToDevice));
    HIP_TRYB(hipMemcpy(b->wb_seq, wbs.data(), sizeof(int) * wbs.size(), hipMemcpyHostToDevice));
    HIP_TRYB(hipMemcpy(b->wb_b, wbb.data(), sizeof(int) * wbb.size(), hipMemcpyHostToDevice));
    HIP_TRYB(hipMemcpy(b->n_wblocks, nw.data(), sizeof(int) * nw.size(), hipMemcpyHostToDevice));
    return PF_OK;
}

// Model images of every sequence (same processing order for all).
int build_batch_tables(pf_batch *b)
{
    for (int i = 0; i < b->S; ++i)
        if (!b->loaded[b->order[i]]) return PF_OK;    // uploaded once every model is present
    std::vector<ModelImage> imgs((size_t)b->S);
    size_t tf = 0;
    for (int i = 0; i < b->S; ++i) {
        const int s = b->order[i];
        int st = make_model_image(b->hm[s], b->order_packed, b->N[s], imgs[i]);
        if (st) return st;
        tf = std::max(tf, imgs[i].blob.size());
        b->sel_ok[s] = imgs[i].sel_mode;
    }
    const int d = b->D, NP = npacked(d);
    b->table_floats = imgs[0].comp_in_lds ? (int)tf : imgs[0].head_floats;
    b->comp_head = imgs[0].comp_in_lds ? -1 : imgs[0].head_floats;
    std::vector<float> tabs(tf * b->S, 0.f), mus((size_t)b->S * b->K * d), sgs((size_t)b->S * b->K * NP);
    std::vector<double> Ts((size_t)b->S * d * d), Tis((size_t)b->S * d * d);
    for (int i = 0; i < b->S; ++i) {
        std::copy(imgs[i].blob.begin(), imgs[i].blob.end(), tabs.begin() + (size_t)i * tf);
        std::copy(imgs[i].mu.begin(), imgs[i].mu.end(), mus.begin() + (size_t)i * b->K * d);
        std::copy(imgs[i].sg.begin(), imgs[i].sg.end(), sgs.begin() + (size_t)i * b->K * NP);
        std::copy(imgs[i].T.begin(), imgs[i].T.end(), Ts.begin() + (size_t)i * d * d);
        std::copy(imgs[i].Ti.begin(), imgs[i].Ti.end(), Tis.begin() + (size_t)i * d * d);
        b->h_seqs[i].table_ofs = (int)(i * tf);
        b->h_seqs[i].pi_total = imgs[i].pi_total;
    }
    (void)hipFree(b->tables); (void)hipFree(b->mus); (void)hipFree(b->sigs); (void)hipFree(b->Ts); (void)hipFree(b->Tis);
    HIP_TRYB(dallocb(&b->tables, tabs.size()));
    HIP_TRYB(dallocb(&b->mus, mus.size()));
    HIP_TRYB(dallocb(&b->sigs, sgs.size()));
    HIP_TRYB(dallocb(&b->Ts, Ts.size()));
    HIP_TRYB(dallocb(&b->Tis, Tis.size()));
    HIP_TRYB(hipMemcpy(b->tables, tabs.data(), sizeof(float) * tabs.size(), hipMemcpyHostToDevice));
    HIP_TRYB(hipMemcpy(b->mus, mus.data(), sizeof(float) * mus.size(), hipMemcpyHostToDevice));
    HIP_TRYB(hipMemcpy(b->sigs, sgs.data(), sizeof(float) * sgs.size(), hipMemcpyHostToDevice));
    HIP_TRYB(hipMemcpy(b->Ts, Ts.data(), sizeof(double) * Ts.size(), hipMemcpyHostToDevice));
    HIP_TRYB(hipMemcpy(b->Tis, Tis.data(), sizeof(double) * Tis.size(), hipMemcpyHostToDevice));
    b->h_T = Ts;
    HIP_TRYB(hipMemcpy(b->seqs, b->h_seqs.data(), sizeof(SeqDesc) * b->S, hipMemcpyHostToDevice));
    return PF_OK;
}

BatchArgs batch_args(pf_batch *b, const int32_t *count, const float *cand, const uint32_t *uni, double *lml,
                     int parity, int fused)
{
    BatchArgs a{};
    a.buf_x[0] = b->bx[0]; a.buf_x[1] = b->bx[1];
    a.buf_P[0] = b->bP[0]; a.buf_P[1] = b->bP[1];
    a.anc = b->anc;
    a.lw = b->lw;
    a.maxkey = b->maxkey;
    a.status = b->status;
    a.cloc = b->cloc;
    a.bsum = b->bsum;
    a.bexp = b->bexp;
    a.bpow = b->bpow;
    a.cloc1 = b->cloc1;
    a.bsum1 = b->bsum1;
    a.bexp1 = b->bexp1;
    a.bpow1 = b->bpow1;
    a.wfused = step_writes_weights(b) ? 1 : 0;
    a.tables = b->tables;
    a.seqs = b->seqs;
    a.tile_seq = b->tile_seq;
    a.tile_j0 = b->tile_j0;
    a.n_tiles = b->n_tiles;
    a.wb_seq = b->wb_seq;
    a.wb_b = b->wb_b;
    a.n_wblocks = b->n_wblocks;
    a.frame = b->frame + parity;
    a.frame_next = b->frame + (1 - parity);
    a.uni = uni;
    a.count = count;
    a.cand = cand;
    a.lml = lml;
    a.rec_comp = nullptr;
    a.rec_anc = nullptr;
    a.S = b->S;
    a.ld = b->ld;
    a.K = b->K;
    a.m = b->m;
    a.J = b->J;
    a.cap = b->cap;
    a.table_floats = b->table_floats;
    a.comp_head = b->comp_head;
    a.stage_floats = b->stage_floats();
    a.base = StepArgs{};
    a.base.K = b->K;
    a.base.m = b->m;
    a.base.J = b->J;
    a.base.cap = b->cap;
    a.base.table_floats = b->table_floats;
    a.base.stage_floats = b->stage_floats();
    a.base.order_packed = b->order_packed;
    a.base.clutter_mask = b->clutter_mask;
    a.parity = parity;
    a.fused = fused;
    a.nb_max = b->nb_max;
    a.all_tiles = b->all_tiles;
    a.order_packed = b->order_packed;
    a.clutter_mask = b->clutter_mask;
    const double pg = 1.0 - std::exp(-0.5 * b->gate);
    a.pdl = (float)(b->pd / b->lambda);
    a.b0 = (float)(1.0 - b->pd * pg);
    a.gate = (float)b->gate;
    a.base.pdl = a.pdl;
    a.base.b0 = a.b0;
    a.base.gate = a.gate;
    return a;
}

template <int SELJ, int PPL>
void launch_step_t(pf_batch *b, const BatchArgs &a, size_t lds)
{
    const int grid = b->sch.blocks_per_cu * b->cus;
    hipLaunchKernelGGL((k_batch_step<kBD, SELJ, PPL>), dim3(grid), dim3(b->sch.block), lds, b->stream, a);
}

int launch_batch_step(pf_batch *b, const BatchArgs &a)
{
    const size_t lds = (size_t)step_lds_bytes(b->stage_floats(), b->nb_max);
    bool all_sel = true;
    for (int s = 0; s < b->S; ++s) all_sel = all_sel && b->sel_ok[s];
    const int selj = all_sel ? b->J : 0;
    const int ppl = b->sch.ppl;
    if (b->D == 24 && b->variant != PF_BATCH_VALU) {
        // four lanes per particle, covariance tiles on the matrix cores (or
        // the vector ALU); one 256-particle tile per block
        if (b->sch.block != 256)
            return set_error(PF_ERR_UNSUPPORTED, "the four-lane d = 24 step runs 256-thread tiles");
        const size_t qlds = lds + 256 * sizeof(int);
        const int grid = b->all_tiles;
        const bool mf = b->variant == PF_BATCH_MFMA;
        if (selj == 3 && mf)
            hipLaunchKernelGGL((k_batch_step_q24<3, true>), dim3(grid), dim3(256), qlds, b->stream, a);
        else if (selj == 3)
            hipLaunchKernelGGL((k_batch_step_q24<3, false>), dim3(grid), dim3(256), qlds, b->stream, a);
        else if (selj == 5 && mf)
            hipLaunchKernelGGL((k_batch_step_q24<5, true>), dim3(grid), dim3(256), qlds, b->stream, a);
        else if (selj == 5)
            hipLaunchKernelGGL((k_batch_step_q24<5, false>), dim3(grid), dim3(256), qlds, b->stream, a);
        else
            return set_error(PF_ERR_UNSUPPORTED, "batched d = 24 is compiled for 3 or 5 joints with a full-rank H");
        HIP_TRYB(hipGetLastError());
        return PF_OK;
    }
    if (b->D == 24) {              // scaled state: 3 or 5 joints in the selection basis, one particle per lane
        const int grid = b->all_tiles;
        if (selj == 3)
            hipLaunchKernelGGL((k_batch_step<24, 3, 1, false>), dim3(grid), dim3(b->sch.block), lds, b->stream, a);
        else if (selj == 5)
            hipLaunchKernelGGL((k_batch_step<24, 5, 1, false>), dim3(grid), dim3(b->sch.block), lds, b->stream, a);
        else
            return set_error(PF_ERR_UNSUPPORTED, "batched d = 24 is compiled for 3 or 5 joints with a full-rank H");
        HIP_TRYB(hipGetLastError());
        return PF_OK;
    }
#define PF_BATCH_CASE(J_)                                           \
    case J_:                                                        \
        if (ppl == 2) launch_step_t<J_, 2>(b, a, lds);             \
        else launch_step_t<J_, 1>(b, a, lds);                      \
        break;
    switch (selj) {
        PF_BATCH_CASE(0)
        PF_BATCH_CASE(1)
        PF_BATCH_CASE(2)
        PF_BATCH_CASE(3)
        PF_BATCH_CASE(4)
        PF_BATCH_CASE(5)
        PF_BATCH_CASE(6)
    default: return set_error(PF_ERR_UNSUPPORTED, "batched engine: %d joints", selj);
    }
#undef PF_BATCH_CASE
    HIP_TRYB(hipGetLastError());
    return PF_OK;
}

int launch_batch_weights(pf_batch *b, const BatchArgs &a)
{
    hipLaunchKernelGGL(k_batch_weights, dim3(4 * b->cus), dim3(kWeightsBlock), 0, b->stream, a);
    HIP_TRYB(hipGetLastError());
    return PF_OK;
}

int launch_batch_finish(pf_batch *b, const BatchArgs &a)
{
    const size_t lds = 8 * (size_t)resolve_lds_words(b->nb_max);
    if (b->sch.ppl == 2)
        hipLaunchKernelGGL((k_batch_finish<2>), dim3(2 * b->cus), dim3(b->sch.block), lds, b->stream, a);
    else
        hipLaunchKernelGGL((k_batch_finish<1>), dim3(2 * b->cus), dim3(b->sch.block), lds, b->stream, a);
    HIP_TRYB(hipGetLastError());
    return PF_OK;
}

int run_all(pf_batch *b, const int32_t *count, const float *cand, const uint32_t *uni, double *lml, int frames,
            bool graph, bool record)
{
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, b->stream, b->frame, 0);
    hipLaunchKernelGGL(k_batch_clear, dim3((2 * b->S + 255) / 256), dim3(256), 0, b->stream, b->maxkey, b->status,
                       b->S);
    if (record) graph = false;
    const size_t xw = (size_t)b->D * b->ld, Pw = (size_t)npacked(b->D) * b->ld;
    auto one = [&](int f) -> int {
        const int parity = f & 1;
        BatchArgs a = batch_args(b, count, cand, uni, lml, parity, f > 0);
        if (record) {
            a.rec_comp = b->rec_comp;
            a.rec_anc = b->rec_anc;
            // Checkpoint c holds the input state of frame c C: the reset state
            // is copied there; for c >= 1 frame c C - 1 writes its updated
            // states into the checkpoint and frame c C reads them from it.
            const int C = b->ckpt_every;
            if (f == 0) {
                HIP_TRYB(hipMemcpyAsync(b->ck_x, b->bx[0], sizeof(float) * xw, hipMemcpyDeviceToDevice, b->stream));
                HIP_TRYB(hipMemcpyAsync(b->ck_P, b->bP[0], sizeof(float) * Pw, hipMemcpyDeviceToDevice, b->stream));
            } else if (f % C == 0) {
                a.buf_x[f & 1] = b->ck_x + (size_t)(f / C) * xw;
                a.buf_P[f & 1] = b->ck_P + (size_t)(f / C) * Pw;
            }
            if ((f + 1) % C == 0 && f + 1 < frames) {
                a.buf_x[(f + 1) & 1] = b->ck_x + (size_t)((f + 1) / C) * xw;
                a.buf_P[(f + 1) & 1] = b->ck_P + (size_t)((f + 1) / C) * Pw;
            }
            // The frames of the last checkpoint segment write their updated
            // states into that segment's slots (slot f - t0 for frames t0 ..
            // frames - 2, where the backward's replay would put them) and
            // read them from there, so the backward starts without a replay.
            const int t0 = (b->n_ckpt - 1) * b->ckpt_every, q = (b->n_ckpt - 1) & 1;
            if (f - 1 >= t0 && f - 1 <= frames - 2) {
                a.buf_x[f & 1] = b->seg_x[q] + (size_t)(f - 1 - t0) * xw;
                a.buf_P[f & 1] = b->seg_P[q] + (size_t)(f - 1 - t0) * Pw;
            }
            if (f >= t0 && f <= frames - 2) {
                a.buf_x[(f + 1) & 1] = b->seg_x[q] + (size_t)(f - t0) * xw;
                a.buf_P[(f + 1) & 1] = b->seg_P[q] + (size_t)(f - t0) * Pw;
            }
        }
        int st = launch_batch_step(b, a);
        if (st) return st;
        return a.wfused ? PF_OK : launch_batch_weights(b, a);
    };
    int st = one(0);
    if (st) return st;
    const int pairs = (frames - 1) / 2;
    if (graph && pairs > 0) {
        hipGraph_t g;
        hipGraphExec_t ge;
        HIP_TRYB(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        int s1 = one(1);
        int s2 = s1 ? s1 : one(2);
        hipError_t ec = hipStreamEndCapture(b->stream, &g);
        if (s2) return s2;
        if (ec != hipSuccess) return set_error(PF_ERR_HIP, "graph capture failed: %s", hipGetErrorString(ec));
        HIP_TRYB(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
        for (int p = 0; p < pairs; ++p) HIP_TRYB(hipGraphLaunch(ge, b->stream));
        HIP_TRYB(hipStreamSynchronize(b->stream));
        (void)hipGraphExecDestroy(ge);
        (void)hipGraphDestroy(g);
        if ((frames - 1) & 1) {
            st = one(frames - 1);
            if (st) return st;
        }
    } else {
        for (int f = 1; f < frames; ++f) {
            st = one(f);
            if (st) return st;
        }
    }
    return PF_OK;
}

}  // namespace batch
}  // namespace mkfbt

using namespace mkfbt::batch;

namespace {

// The C ABI status of the last run: the per-sequence status words ORed.
int batch_status(pf_batch *b)
{
    std::vector<unsigned> w((size_t)b->S);
    HIP_TRYB(hipMemcpy(w.data(), b->status, sizeof(unsigned) * b->S, hipMemcpyDeviceToHost));
    unsigned word = 0;
    for (unsigned v : w) word |= v;
    return status_code(word);
}

}  // namespace

extern "C" {

int pf_batch_destroy(pf_batch *b)
{
    if (!b) return PF_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    void *ptrs[] = {b->bx[0], b->bx[1], b->bP[0], b->bP[1], b->lw, b->tables, b->mus, b->sigs, b->Tis, b->Ts,
                    b->bexp, b->bpow, b->partial, b->est, b->gbuf, b->anc, b->frame, b->maxkey, b->cloc, b->bsum,
                    b->cloc1, b->bsum1, b->bexp1, b->bpow1, b->status, b->seqs, b->tile_seq, b->tile_j0, b->n_tiles, b->wb_seq, b->wb_b,
                    b->n_wblocks, b->fs, b->rec_comp, b->rec_anc, b->rec_bins, b->ck_x, b->ck_P, b->seg_x[0],
                    b->seg_x[1], b->seg_P[0], b->seg_P[1], b->adj_x[0], b->adj_x[1], b->adj_P[0], b->adj_P[1], b->first, b->tacc, b->gsum,
                    b->t256_seq, b->t256_j0, b->t256_first};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    if (b->stream2) {
        (void)hipStreamSynchronize(b->stream2);
        (void)hipStreamDestroy(b->stream2);
    }
    for (hipEvent_t e : b->ev_replay) (void)hipEventDestroy(e);
    for (hipEvent_t e : b->ev_back) (void)hipEventDestroy(e);
    if (b->ev_fwd) (void)hipEventDestroy(b->ev_fwd);
    delete b;
    return PF_OK;
}

int pf_batch_create(pf_batch **out, int32_t S, const int32_t *n_particles, const int32_t *n_frames, int32_t K,
                    int32_t d, int32_t m, const pf_config *cfg)
{
    if (!out || !n_particles || !n_frames) return set_error(PF_ERR_ARG, "null argument");
    *out = nullptr;
    if (S <= 0 || S > (1 << 16)) return set_error(PF_ERR_ARG, "S=%d out of range", S);
    if (d != kBD && d != 24) return set_error(PF_ERR_UNSUPPORTED, "the batched engine runs d = 12 or 24");
    if (K <= 0 || K > 256 || m <= 0 || (m & 1) || m > 12) return set_error(PF_ERR_ARG, "bad K/m");
    pf_config c;
    pf_config_default(&c);
    if (cfg) c = *cfg;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return set_error(PF_ERR_HIP, "no HIP device available");
    if (c.device < 0 || c.device >= ndev) return set_error(PF_ERR_ARG, "device out of range");
    HIP_TRYB(hipSetDevice(c.device));
    pf_batch *b = new pf_batch();
    b->device = c.device;
    b->S = S; b->K = K; b->D = d; b->m = m; b->J = m / 2;
    b->cap = c.cand_capacity > 0 ? c.cand_capacity : 16;
    b->pd = c.p_detect; b->gate = c.gate_chi2; b->lambda = c.clutter_density;
    BatchLayout L;
    int lst = make_layout(S, n_particles, n_frames, L);
    if (lst) { delete b; return lst; }
    b->N.assign(n_particles, n_particles + S);
    b->T.assign(n_frames, n_frames + S);
    int cus = 0;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c.device);
    b->cus = cus > 0 ? cus : 256;
    b->order = L.order;
    b->rank.resize(S);
    for (int i = 0; i < S; ++i) b->rank[b->order[i]] = i;
    b->tofs_user.resize(S);
    for (int s = 0, acc = 0; s < S; ++s) { b->tofs_user[s] = acc; acc += b->T[s]; }
    b->h_seqs = L.seqs;
    b->nb_max = L.nb_max;
    b->Tmax = L.Tmax;
    const long long bofs = L.nblocks;
    b->ld = L.ld;
    b->hm.resize(S);
    b->loaded.assign(S, 0);
    b->sel_ok.assign(S, 0);
    b->fresh.assign(S, 0);
    const size_t NP = npacked(d);
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    for (int k = 0; k < 2 && e == hipSuccess; ++k) {
        e = dallocb(&b->bx[k], (size_t)d * b->ld);
        if (e == hipSuccess) e = dallocb(&b->bP[k], NP * b->ld);
    }
    e = e ? e : dallocb(&b->lw, (size_t)b->ld);
    e = e ? e : dallocb(&b->anc, (size_t)b->ld);
    e = e ? e : dallocb(&b->cloc, (size_t)b->ld);
    e = e ? e : dallocb(&b->bsum, (size_t)bofs);
    e = e ? e : dallocb(&b->bexp, (size_t)bofs);
    e = e ? e : dallocb(&b->bpow, (size_t)bofs);
    e = e ? e : dallocb(&b->cloc1, (size_t)b->ld);
    e = e ? e : dallocb(&b->bsum1, (size_t)bofs);
    e = e ? e : dallocb(&b->bexp1, (size_t)bofs);
    e = e ? e : dallocb(&b->bpow1, (size_t)bofs);
    e = e ? e : dallocb(&b->maxkey, 2 * (size_t)S);
    e = e ? e : dallocb(&b->status, (size_t)S);
    e = e ? e : dallocb(&b->seqs, (size_t)S);
    e = e ? e : dallocb(&b->frame, 2);
    e = e ? e : dallocb(&b->rec_bins, (size_t)b->ld);
    e = e ? e : dallocb(&b->partial, (size_t)kMaxScanBlocks * 64);
    e = e ? e : dallocb(&b->est, 64);
    e = e ? e : dallocb(&b->fs, 1);
    if (e == hipSuccess) e = hipMemset(b->maxkey, 0, sizeof(unsigned long long) * 2 * S);
    if (e == hipSuccess) e = hipMemset(b->status, 0, sizeof(unsigned) * S);
    if (e == hipSuccess) e = hipMemset(b->fs, 0, sizeof(FrameScratch));
    if (e == hipSuccess) e = hipMemcpy(b->seqs, b->h_seqs.data(), sizeof(SeqDesc) * S, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        pf_batch_destroy(b);
        return set_error(PF_ERR_HIP, "batch allocation failed: %s", hipGetErrorString(e));
    }
    b->order_packed = joint_order(0, b->J);
    int st = build_tiles(b);
    if (st) { pf_batch_destroy(b); return st; }
    *out = b;
    return PF_OK;
}

int pf_batch_plan(int32_t S, const int32_t *n_particles, const int32_t *n_frames, int32_t block, int32_t ppl,
                  int64_t *summary)
{
    if (!n_particles || !n_frames) return set_error(PF_ERR_ARG, "null argument");
    if ((block != 64 && block != 128 && block != 256) || (ppl != 1 && ppl != 2))
        return set_error(PF_ERR_ARG, "bad schedule");
    BatchLayout L;
    int st = make_layout(S, n_particles, n_frames, L);
    if (st) return st;
    make_tiles(L, block * ppl);
    st = check_layout(L, block * ppl);
    if (st) return st;
    if (summary) {
        summary[0] = L.ld;
        summary[1] = L.nblocks;
        summary[2] = (int64_t)L.ts.size();
        summary[3] = (int64_t)L.wbs.size();
        summary[4] = L.nb_max;
        summary[5] = L.Tmax;
    }
    return PF_OK;
}

int pf_batch_load_model(pf_batch *b, int32_t s, int32_t npose, const int32_t *sel, const double *means,
                        const double *covs, const double *weights, const double *gamma, const double *pca_proj,
                        const double *pca_mean, const double *R)
{
    int st = bcheck(b);
    if (st) return st;
    if (s < 0 || s >= b->S) return set_error(PF_ERR_ARG, "sequence %d out of range", s);
    HostModel hm;
    st = host_model_from_args(hm, b->K, b->D, npose, b->m, sel, means, covs, weights, gamma, pca_proj, pca_mean, R);
    if (st) return st;
    b->hm[s] = hm;
    b->loaded[s] = 1;
    b->fresh[s] = 0;                          // tracks live in the basis of the model image
    b->tracks = false;
    return build_batch_tables(b);
}

int pf_batch_set_joint_clutter(pf_batch *b, const int32_t *clutter)
{
    int st = bcheck(b);
    if (st) return st;
    if (std::any_of(b->fresh.begin(), b->fresh.end(), [](char c) { return c != 0; }))
        return set_error(PF_ERR_STATE, "joint kinds must be set before the tracks are initialised");
    unsigned mask = 0;
    for (int j = 0; j < b->J; ++j)
        if (clutter && clutter[j]) mask |= 1u << j;
    b->clutter_mask = mask;
    b->order_packed = joint_order(mask, b->J);
    return build_batch_tables(b);
}

int pf_batch_reset(pf_batch *b, int32_t s, const int32_t *bins)
{
    int st = bcheck(b);
    if (st) return st;
    if (s < 0 || s >= b->S || !bins) return set_error(PF_ERR_ARG, "bad reset arguments");
    for (int i = 0; i < b->S; ++i)
        if (!b->loaded[i]) return set_error(PF_ERR_STATE, "load every sequence's model first");
    const int i = b->rank[s];
    const SeqDesc &sd = b->h_seqs[i];
    for (int j = 0; j < sd.N; ++j)
        if (bins[j] < 0 || bins[j] >= b->K) return set_error(PF_ERR_ARG, "bins[%d] out of range", j);
    HIP_TRYB(hipMemcpyAsync(b->rec_bins + sd.pofs, bins, sizeof(int) * sd.N, hipMemcpyHostToDevice, b->stream));
    auto reset_kernel = b->D == 24 ? k_reset<24> : k_reset<kBD>;
    hipLaunchKernelGGL(reset_kernel, dim3((sd.N + 255) / 256), dim3(256), 0, b->stream, b->rec_bins + sd.pofs,
                       b->mus + (size_t)i * b->K * b->D, b->sigs + (size_t)i * b->K * npacked(b->D),
                       b->bx[0] + sd.pofs, b->bP[0] + sd.pofs, sd.N, Lay{b->ld, 1});
    hipLaunchKernelGGL(k_identity_seq, dim3((sd.N + 255) / 256), dim3(256), 0, b->stream, b->anc, sd.pofs, sd.N);
    HIP_TRYB(hipGetLastError());
    HIP_TRYB(hipStreamSynchronize(b->stream));
    b->fresh[s] = 1;
    b->tracks = std::all_of(b->fresh.begin(), b->fresh.end(), [](char c) { return c != 0; });
    return PF_OK;
}

int pf_batch_run_dev(pf_batch *b, const int32_t *count, const float *cand, const uint32_t *uniforms, double *lml,
                     int32_t flags)
{
    int st = bcheck(b);
    if (st) return st;
    if (!b->tracks) return set_error(PF_ERR_STATE, "reset every sequence before a run");
    if (!count || !cand || !uniforms || !lml) return set_error(PF_ERR_ARG, "null run argument");
    b->tracks = false;                        // a run consumes the initial tracks
    std::fill(b->fresh.begin(), b->fresh.end(), 0);
    if ((size_t)step_lds_bytes(b->stage_floats(), b->nb_max) > (size_t)kMaxLdsBytes)
        return set_error(PF_ERR_UNSUPPORTED, "model image and resampler workspace exceed the LDS budget");
    hipEvent_t e0, e1;
    HIP_TRYB(hipEventCreate(&e0));
    HIP_TRYB(hipEventCreate(&e1));
    HIP_TRYB(hipEventRecord(e0, b->stream));
    st = run_all(b, count, cand, uniforms, lml, b->Tmax, (flags & 1) != 0);
    if (st) return st;
    BatchArgs a = batch_args(b, count, cand, uniforms, lml, 0, 1);
    st = launch_batch_finish(b, a);
    if (st) return st;
    HIP_TRYB(hipEventRecord(e1, b->stream));
    HIP_TRYB(hipEventSynchronize(e1));
    HIP_TRYB(hipEventElapsedTime(&b->last_ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return batch_status(b);
}

int pf_batch_get_status(pf_batch *b, int32_t *status)
{
    int st = bcheck(b);
    if (st) return st;
    if (!status) return set_error(PF_ERR_ARG, "null argument");
    std::vector<unsigned> w((size_t)b->S);
    HIP_TRYB(hipMemcpy(w.data(), b->status, sizeof(unsigned) * b->S, hipMemcpyDeviceToHost));
    for (int s = 0; s < b->S; ++s) status[s] = (int32_t)(w[(size_t)b->rank[s]] & 7u);
    return PF_OK;
}

int pf_batch_last_run_ms(pf_batch *b, double *ms)
{
    if (!b || !ms) return set_error(PF_ERR_ARG, "null argument");
    *ms = b->last_ms;
    return PF_OK;
}

int pf_batch_get_tracks(pf_batch *b, int32_t s, double *x, double *P)
{
    int st = bcheck(b);
    if (st) return st;
    if (s < 0 || s >= b->S) return set_error(PF_ERR_ARG, "sequence out of range");
    const int i = b->rank[s], d = b->D;
    const SeqDesc &sd = b->h_seqs[i];
    const size_t nx = (size_t)sd.N * d, np = (size_t)sd.N * d * d;
    if (b->gbuf_bytes < sizeof(double) * (nx + np)) {
        (void)hipFree(b->gbuf);
        b->gbuf = nullptr;
        HIP_TRYB(hipMalloc((void **)&b->gbuf, sizeof(double) * (nx + np)));
        b->gbuf_bytes = sizeof(double) * (nx + np);
    }
    const int buf = sd.T & 1;   // frame f writes buffer (f + 1) & 1; the last frame is T - 1
    if (d == 24)
        hipLaunchKernelGGL((k_export_rows<24>), dim3((unsigned)(((long long)sd.N * 24 + 255) / 256)), dim3(256), 0,
                           b->stream, b->bx[buf] + sd.pofs, b->bP[buf] + sd.pofs, b->anc + sd.pofs,
                           b->Tis + (size_t)i * d * d, b->gbuf, b->gbuf + nx, sd.N, Lay{b->ld, 1});
    else
        hipLaunchKernelGGL((k_export<kBD>), dim3((sd.N + 255) / 256), dim3(256), 0, b->stream, b->bx[buf] + sd.pofs,
                           b->bP[buf] + sd.pofs, b->anc + sd.pofs, b->Tis + (size_t)i * d * d, b->gbuf,
                           b->gbuf + nx, sd.N, Lay{b->ld, 1});
    HIP_TRYB(hipGetLastError());
    if (x) HIP_TRYB(hipMemcpyAsync(x, b->gbuf, sizeof(double) * nx, hipMemcpyDeviceToHost, b->stream));
    if (P) HIP_TRYB(hipMemcpyAsync(P, b->gbuf + nx, sizeof(double) * np, hipMemcpyDeviceToHost, b->stream));
    HIP_TRYB(hipStreamSynchronize(b->stream));
    return PF_OK;
}

int pf_batch_set_tracks(pf_batch *b, int32_t s, const double *x, const double *P)
{
    int st = bcheck(b);
    if (st) return st;
    if (s < 0 || s >= b->S || !x || !P) return set_error(PF_ERR_ARG, "bad set_tracks arguments");
    for (int i = 0; i < b->S; ++i)
        if (!b->loaded[i]) return set_error(PF_ERR_STATE, "load every sequence's model first");
    const int i = b->rank[s], d = b->D;
    const SeqDesc &sd = b->h_seqs[i];
    const size_t nx = (size_t)sd.N * d, np = (size_t)sd.N * d * d;
    if (b->gbuf_bytes < sizeof(double) * (nx + np)) {
        (void)hipFree(b->gbuf);
        b->gbuf = nullptr;
        b->gbuf_bytes = 0;
        HIP_TRYB(hipMalloc((void **)&b->gbuf, sizeof(double) * (nx + np)));
        b->gbuf_bytes = sizeof(double) * (nx + np);
    }
    HIP_TRYB(hipMemcpyAsync(b->gbuf, x, sizeof(double) * nx, hipMemcpyHostToDevice, b->stream));
    HIP_TRYB(hipMemcpyAsync(b->gbuf + nx, P, sizeof(double) * np, hipMemcpyHostToDevice, b->stream));
    // x' = T x, P' = T P T^T into buffer 0 (frame 0 reads it), ancestors = identity
    if (d == 24)
        hipLaunchKernelGGL((k_import_rows<24>), dim3((unsigned)(((long long)sd.N * 24 + 255) / 256)), dim3(256), 0,
                           b->stream, (const double *)b->gbuf, (const double *)(b->gbuf + nx),
                           (const double *)(b->Ts + (size_t)i * d * d), b->bx[0] + sd.pofs, b->bP[0] + sd.pofs, sd.N,
                           Lay{b->ld, 1});
    else
        hipLaunchKernelGGL((k_import<kBD>), dim3((sd.N + 255) / 256), dim3(256), 0, b->stream,
                           (const double *)b->gbuf, (const double *)(b->gbuf + nx),
                           (const double *)(b->Ts + (size_t)i * d * d), b->bx[0] + sd.pofs, b->bP[0] + sd.pofs, sd.N,
                           Lay{b->ld, 1});
    hipLaunchKernelGGL(k_identity_seq, dim3((sd.N + 255) / 256), dim3(256), 0, b->stream, b->anc, sd.pofs, sd.N);
    HIP_TRYB(hipGetLastError());
    HIP_TRYB(hipStreamSynchronize(b->stream));
    b->fresh[s] = 1;
    b->tracks = std::all_of(b->fresh.begin(), b->fresh.end(), [](char c) { return c != 0; });
    return PF_OK;
}

int pf_batch_get_last_frame(pf_batch *b, int32_t s, float *lw, int32_t *anc)
{
    int st = bcheck(b);
    if (st) return st;
    if (s < 0 || s >= b->S) return set_error(PF_ERR_ARG, "sequence out of range");
    const SeqDesc &sd = b->h_seqs[b->rank[s]];
    if (lw) HIP_TRYB(hipMemcpyAsync(lw, b->lw + sd.pofs, sizeof(float) * sd.N, hipMemcpyDeviceToHost, b->stream));
    if (anc) HIP_TRYB(hipMemcpyAsync(anc, b->anc + sd.pofs, sizeof(int) * sd.N, hipMemcpyDeviceToHost, b->stream));
    HIP_TRYB(hipStreamSynchronize(b->stream));
    return PF_OK;
}

int pf_batch_estimate(pf_batch *b, int32_t s, double *xbar)
{
    int st = bcheck(b);
    if (st) return st;
    if (s < 0 || s >= b->S || !xbar) return set_error(PF_ERR_ARG, "bad estimate arguments");
    const int i = b->rank[s];
    const SeqDesc &sd = b->h_seqs[i];
    const int buf = sd.T & 1;
    if (b->D == 24)
        launch_estimate<24>(b->stream, b->bx[buf] + sd.pofs, b->anc + sd.pofs, sd.N, Lay{b->ld, 1}, b->partial,
                            b->Tis + (size_t)i * b->D * b->D, b->est);
    else
        launch_estimate<kBD>(b->stream, b->bx[buf] + sd.pofs, b->anc + sd.pofs, sd.N, Lay{b->ld, 1}, b->partial,
                             b->Tis + (size_t)i * b->D * b->D, b->est);
    HIP_TRYB(hipGetLastError());
    HIP_TRYB(hipMemcpyAsync(xbar, b->est, sizeof(double) * b->D, hipMemcpyDeviceToHost, b->stream));
    HIP_TRYB(hipStreamSynchronize(b->stream));
    return PF_OK;
}

int pf_batch_get_schedule(pf_batch *b, pf_batch_schedule *sch)
{
    if (!b || !sch) return set_error(PF_ERR_ARG, "null argument");
    *sch = b->sch;
    return PF_OK;
}

int pf_batch_set_schedule(pf_batch *b, const pf_batch_schedule *sch)
{
    int st = bcheck(b);
    if (st) return st;
    if (!sch || !(sch->block == 64 || sch->block == 128 || sch->block == 256) || !(sch->ppl == 1 || sch->ppl == 2) ||
        sch->blocks_per_cu < 1 || sch->blocks_per_cu > 8)
        return set_error(PF_ERR_ARG, "bad schedule");
    if (b->D == 24 && sch->ppl != 1) return set_error(PF_ERR_ARG, "d = 24 runs one particle per lane");
    if (b->D == 24 && b->variant != PF_BATCH_VALU && sch->block != 256)
        return set_error(PF_ERR_ARG, "the four-lane d = 24 variants run 256-thread tiles");
    b->sch = *sch;
    return build_tiles(b);
}

int pf_batch_set_variant(pf_batch *b, int32_t variant)
{
    int st = bcheck(b);
    if (st) return st;
    if (variant != PF_BATCH_VALU && variant != PF_BATCH_MFMA && variant != PF_BATCH_QUAD_VALU)
        return set_error(PF_ERR_ARG, "unknown variant %d", variant);
    if (variant != PF_BATCH_VALU && b->D != 24)
        return set_error(PF_ERR_UNSUPPORTED, "the four-lane variants are the d = 24 state's");
    if (variant != PF_BATCH_VALU && b->sch.block != 256) {
        const pf_batch_schedule s256{256, 1, b->sch.blocks_per_cu};
        st = pf_batch_set_schedule(b, &s256);
        if (st) return st;
    }
    b->variant = variant;
    return PF_OK;
}

int pf_batch_get_variant(pf_batch *b, int32_t *variant)
{
    if (!b || !variant) return set_error(PF_ERR_ARG, "null argument");
    *variant = b->D == 24 ? b->variant : PF_BATCH_VALU;
    return PF_OK;
}

int pf_batch_autotune(pf_batch *b, const int32_t *count, const float *cand, const uint32_t *uniforms, double *lml,
                      int32_t frames, pf_batch_schedule *best, double *best_ms, double *default_ms)
{
    int st = bcheck(b);
    if (st) return st;
    if (!b->tracks) return set_error(PF_ERR_STATE, "reset every sequence before tuning");
    if (!count || !cand || !uniforms || !lml) return set_error(PF_ERR_ARG, "null tuning argument");
    b->tracks = false;
    std::fill(b->fresh.begin(), b->fresh.end(), 0);
    frames = std::max(1, std::min(frames, b->Tmax));
    const pf_batch_schedule def{256, 1, 2};
    std::vector<pf_batch_schedule> cands;
    for (int blk : {64, 128, 256})
        for (int ppl : {1, 2})
            for (int bpc : {1, 2, 3, 4})
                if (b->D != 24 || (ppl == 1 && (b->variant == PF_BATCH_VALU || blk == 256)))
                    cands.push_back({blk, ppl, bpc});
    hipEvent_t e0, e1;
    HIP_TRYB(hipEventCreate(&e0));
    HIP_TRYB(hipEventCreate(&e1));
    double best_t = 1e30, def_t = -1.0;
    pf_batch_schedule best_s = def;
    for (const auto &c : cands) {
        st = pf_batch_set_schedule(b, &c);
        if (st) return st;
        // one untimed warm-up frame pair, then `frames` timed frames
        st = run_all(b, count, cand, uniforms, lml, 3, false);
        if (st) return st;
        HIP_TRYB(hipEventRecord(e0, b->stream));
        st = run_all(b, count, cand, uniforms, lml, frames, false);
        if (st) return st;
        HIP_TRYB(hipEventRecord(e1, b->stream));
        HIP_TRYB(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_TRYB(hipEventElapsedTime(&ms, e0, e1));
        if (c.block == def.block && c.ppl == def.ppl && c.blocks_per_cu == def.blocks_per_cu) def_t = ms;
        if (ms < best_t) { best_t = ms; best_s = c; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    st = pf_batch_set_schedule(b, &best_s);
    if (st) return st;
    if (best) *best = best_s;
    if (best_ms) *best_ms = best_t;
    if (default_ms) *default_ms = def_t;
    return PF_OK;
}

}  // extern "C"
