// pf_capi.hip -- C ABI of the gfx950 particle-filter hot path (include/pf_rbpf.h).
//
// The handle is one reference ParticleFilter (src/pf2DRao.h:13-31) with its
// my_gmm model tables resident in HBM.  Host code here only validates,
// converts layouts (row-major double <-> FP32 packed records or rows) and launches; every
// arithmetic step of the filter runs in the kernels of pf_kernels.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pf_rbpf.h"
#include "pf_internal.hpp"
#include "pf_engine.hpp"
#include "pf_kernels.hpp"
#include "pf_model.hpp"

using namespace mkfbt;

namespace mkfbt {

thread_local std::string g_last_error;

int set_error(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int status_code(unsigned word)
{
    if (word & kStatusNotSpd)
        return set_error(PF_ERR_NOT_SPD, "an innovation covariance was not positive definite even in FP64: those "
                                         "particles got zero weight and kept their predicted state");
    return (int)(word & (kStatusDegenerate | kStatusIllCond));
}

}  // namespace mkfbt

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return set_error(PF_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));       \
    } while (0)

struct pf_handle {
    pf_config cfg{};
    int N = 0, ld = 0;
    int D = 0, K = 0, m = 0, J = 0, npose = 0, cap = 16;
    bool model = false, tracks = false;
    unsigned int clutter_mask = 0, order_packed = 0;
    hipStream_t stream = nullptr;
    // state (FP32, packed covariance), double buffered: buffer q is one
    // allocation rec[q] holding x and P in layout lay (per-particle records
    // for d <= 12, rows otherwise); xA / PA view the buffer read by the next
    // step, xB / PB the other
    float *rec[2] = {nullptr, nullptr};
    Lay lay{0, 1};
    int layout_req = PF_LAYOUT_AUTO;   // pf_set_state_layout
    int cus = 0;                       // compute units of the device
    float *xA = nullptr, *PA = nullptr, *xB = nullptr, *PB = nullptr;
    int *anc = nullptr, *comp = nullptr, *bins = nullptr;
    float *lw = nullptr;
    // model
    float *table = nullptr;
    int table_floats = 0;          // floats of the image staged in LDS
    const float *comp_global = nullptr;   // component rows read from global memory (large images)
    unsigned long long pi_total = 0;
    float *mu = nullptr, *sig = nullptr;
    HostModel hm;
    // state basis x' = T x (T = I when H has no usable row-selection basis)
    bool sel_mode = false;
    double *Tdev = nullptr, *Tinv = nullptr;   // d x d each, device FP64
    // frame bookkeeping
    FrameScratch *fs = nullptr;
    int *frame = nullptr;
    uint32_t *uni = nullptr;   // 2 words for the host-driven step
    double *lml1 = nullptr;
    // resampler scratch (set 0), and set 1 of the device-resident runs, whose
    // step kernel writes frame t's weights while frame t - 1's are resolved
    unsigned long long *cloc = nullptr, *bsum = nullptr;
    double *bexp = nullptr;
    int *bpow = nullptr;
    int Lcap = 0;
    unsigned long long *cloc1 = nullptr, *bsum1 = nullptr;
    double *bexp1 = nullptr;
    int *bpow1 = nullptr;
    // staging for host-driven calls
    float *zbuf = nullptr;
    int *cntbuf = nullptr;
    float *candbuf = nullptr;
    double *est_partial = nullptr, *est_out = nullptr;
    double *gbuf = nullptr;
    size_t gbuf_bytes = 0;
    int *ibuf = nullptr;
    size_t ibuf_bytes = 0;
    // timing of the last device-resident run
    std::vector<hipEvent_t> ev;
    int timed_launches = 0;
    double last_step_ms = 0.0, last_frame_ms = 0.0;
    int graph_frames = 0;              // frames of the last run replayed from graphs
    // device-resident runs: the state buffers by identity (xA/xB swap roles),
    // the run's inputs and frame counters (frame t reads counter t & 1 and
    // writes t + 1 to the other), and the captured frame groups per mode and
    // parity of their first frame, group g holding 2^g frames (dropped when
    // anything they bake in changes: model tables, resampler scratch)
    float *xbuf[2] = {nullptr, nullptr}, *Pbuf[2] = {nullptr, nullptr};
    RunDesc *run_desc = nullptr;
    int *run_ctr = nullptr;
    static constexpr int kGroups = 7;  // 1, 2, 4, ..., 64 frames
    hipGraphExec_t groups[2][2][kGroups] = {};
};

namespace {

int check_handle(pf_handle *h)
{
    if (!h) return set_error(PF_ERR_ARG, "null handle");
    if (hipSetDevice(h->cfg.device) != hipSuccess) return set_error(PF_ERR_HIP, "hipSetDevice failed");
    return PF_OK;
}

template <typename T>
hipError_t dalloc(T **p, size_t n)
{
    return hipMalloc((void **)p, sizeof(T) * std::max<size_t>(n, 1));
}

// Destroys the captured frame groups (they bake in device addresses).
void drop_graphs(pf_handle *h)
{
    for (auto &per_mode : h->groups)
        for (auto &per_parity : per_mode)
            for (hipGraphExec_t &g : per_parity)
                if (g) {
                    (void)hipGraphExecDestroy(g);
                    g = nullptr;
                }
}

int ensure_resample_scratch(pf_handle *h, int L)
{
    if (L <= h->Lcap) return PF_OK;
    drop_graphs(h);
    (void)hipFree(h->cloc);
    (void)hipFree(h->bsum);
    (void)hipFree(h->bexp);
    (void)hipFree(h->bpow);
    h->cloc = nullptr; h->bsum = nullptr; h->bexp = nullptr; h->bpow = nullptr;
    const int nb = (L + kWeightsBlock - 1) / kWeightsBlock;
    HIP_TRY(dalloc(&h->cloc, (size_t)L));
    HIP_TRY(dalloc(&h->bsum, (size_t)nb));
    HIP_TRY(dalloc(&h->bexp, (size_t)nb));
    HIP_TRY(dalloc(&h->bpow, (size_t)nb));
    h->Lcap = L;
    return PF_OK;
}

// Set 1 of the resampler scratch, sized for the handle's N.
int ensure_run_scratch(pf_handle *h)
{
    if (h->cloc1) return PF_OK;
    const int nb = (h->N + kWeightsBlock - 1) / kWeightsBlock;
    HIP_TRY(dalloc(&h->cloc1, (size_t)h->N));
    HIP_TRY(dalloc(&h->bsum1, (size_t)nb));
    HIP_TRY(dalloc(&h->bexp1, (size_t)nb));
    HIP_TRY(dalloc(&h->bpow1, (size_t)nb));
    return PF_OK;
}

int ensure_gbuf(pf_handle *h, size_t bytes)
{
    if (bytes <= h->gbuf_bytes) return PF_OK;
    (void)hipFree(h->gbuf);
    h->gbuf = nullptr;
    HIP_TRY(hipMalloc((void **)&h->gbuf, bytes));
    h->gbuf_bytes = bytes;
    return PF_OK;
}

int ensure_ibuf(pf_handle *h, size_t bytes)
{
    if (bytes <= h->ibuf_bytes) return PF_OK;
    (void)hipFree(h->ibuf);
    h->ibuf = nullptr;
    HIP_TRY(hipMalloc((void **)&h->ibuf, bytes));
    h->ibuf_bytes = bytes;
    return PF_OK;
}

int check_launch(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(PF_ERR_HIP, "%s launch failed: %s", what, hipGetErrorString(e));
    return PF_OK;
}

// ---------------------------------------------------------- kernel dispatch
template <int D, int SELJ>
int launch_step_dj(pf_handle *h, const StepArgs &a, int mode, int block)
{
    const bool rec = h->lay.row == 1;
    const size_t lds = (size_t)step_lds_bytes_rec(rec, a.stage_floats, a.anc_mode == kAncFused ? a.pend.nb : 0);
    const int grid = (h->N + block - 1) / block;
    if constexpr (rec_layout(D)) {
        if (rec) {
            if (mode == kModeRef)
                hipLaunchKernelGGL((k_step<D, kModeRef, SELJ, true>), dim3(grid), dim3(block), lds, h->stream, a);
            else
                hipLaunchKernelGGL((k_step<D, kModeCand, SELJ, true>), dim3(grid), dim3(block), lds, h->stream, a);
            return check_launch("k_step");
        }
    }
    if (mode == kModeRef)
        hipLaunchKernelGGL((k_step<D, kModeRef, SELJ>), dim3(grid), dim3(block), lds, h->stream, a);
    else
        hipLaunchKernelGGL((k_step<D, kModeCand, SELJ>), dim3(grid), dim3(block), lds, h->stream, a);
    return check_launch("k_step");
}

// Selection-basis kernels exist for every joint count J with 2J <= D.
template <int D>
int launch_step_d(pf_handle *h, const StepArgs &a, int mode, int block)
{
    if (!h->sel_mode) return launch_step_dj<D, 0>(h, a, mode, block);
    switch (h->J) {
    case 1: return launch_step_dj<D, 1>(h, a, mode, block);
    case 2: return launch_step_dj<D, 2>(h, a, mode, block);
    case 3: return launch_step_dj<D, 3>(h, a, mode, block);
    case 4: return launch_step_dj<D, 4>(h, a, mode, block);
    case 5: if constexpr (D >= 10) return launch_step_dj<D, 5>(h, a, mode, block); break;
    case 6: if constexpr (D >= 12) return launch_step_dj<D, 6>(h, a, mode, block); break;
    case 7: if constexpr (D >= 14) return launch_step_dj<D, 7>(h, a, mode, block); break;
    case 8: if constexpr (D >= 16) return launch_step_dj<D, 8>(h, a, mode, block); break;
    default: break;
    }
    return launch_step_dj<D, 0>(h, a, mode, block);
}

// The 24-dimensional state is compiled for 3 and 5 joints in the selection
// basis (the scaled-state workloads); other joint counts report unsupported.
int launch_step_d24(pf_handle *h, const StepArgs &a, int mode, int block)
{
    if (h->sel_mode && h->J == 3) return launch_step_dj<24, 3>(h, a, mode, block);
    if (h->sel_mode && h->J == 5) return launch_step_dj<24, 5>(h, a, mode, block);
    return set_error(PF_ERR_UNSUPPORTED, "d = 24 is compiled for 3 or 5 joints with a full-rank H");
}

int launch_step_block(pf_handle *h, const StepArgs &a, int mode, int block)
{
    switch (h->D) {
    case 8: return launch_step_d<8>(h, a, mode, block);
    case 12: return launch_step_d<12>(h, a, mode, block);
    case 16: return launch_step_d<16>(h, a, mode, block);
    case 24: return launch_step_d24(h, a, mode, block);
    default: return set_error(PF_ERR_UNSUPPORTED, "state dimension %d not compiled", h->D);
    }
}

int launch_step(pf_handle *h, const StepArgs &a, int mode)
{
    return launch_step_block(h, a, mode, h->cfg.block_size > 0 ? h->cfg.block_size : 256);
}

template <template <int> class F, typename... Args>
int dispatch_d(int D, Args... args)
{
    switch (D) {
    case 8: F<8>::run(args...); return PF_OK;
    case 12: F<12>::run(args...); return PF_OK;
    case 16: F<16>::run(args...); return PF_OK;
    case 24: F<24>::run(args...); return PF_OK;
    default: return set_error(PF_ERR_UNSUPPORTED, "state dimension %d not compiled", D);
    }
}

template <int D>
struct ResetL {
    static void run(pf_handle *h)
    {
        hipLaunchKernelGGL((k_reset<D>), dim3((h->N + 255) / 256), dim3(256), 0, h->stream, h->bins,
                           h->mu, h->sig, h->xA, h->PA, h->N, h->lay);
    }
};

template <int D>
struct ExportL {
    static void run(pf_handle *h, double *x, double *P)
    {
        if constexpr (D > 16)
            hipLaunchKernelGGL((k_export_rows<D>), dim3((unsigned)(((long long)h->N * D + 255) / 256)), dim3(256), 0,
                               h->stream, h->xA, h->PA, h->anc, h->Tinv, x, P, h->N, h->lay);
        else
            hipLaunchKernelGGL((k_export<D>), dim3((h->N + 255) / 256), dim3(256), 0, h->stream, h->xA, h->PA,
                               h->anc, h->Tinv, x, P, h->N, h->lay);
    }
};

template <int D>
struct ImportL {
    static void run(pf_handle *h, const double *x, const double *P)
    {
        if constexpr (D > 16)
            hipLaunchKernelGGL((k_import_rows<D>), dim3((unsigned)(((long long)h->N * D + 255) / 256)), dim3(256), 0,
                               h->stream, x, P, h->Tdev, h->xA, h->PA, h->N, h->lay);
        else
            hipLaunchKernelGGL((k_import<D>), dim3((h->N + 255) / 256), dim3(256), 0, h->stream, x, P, h->Tdev,
                               h->xA, h->PA, h->N, h->lay);
    }
};

template <int D>
struct EstimateL {
    static void run(pf_handle *h)
    {
        launch_estimate<D>(h->stream, h->xA, h->anc, h->N, h->lay, h->est_partial, h->Tinv, h->est_out);
    }
};

// The pending resampling held in scratch set `set` (0, or 1 of the
// device-resident runs): of L log-weights (block exponents) or linear weights.
Pending pending_of(const pf_handle *h, int L, bool log_weights = true, int set = 0)
{
    Pending p{};
    p.cloc = set ? h->cloc1 : h->cloc;
    p.bsum = set ? h->bsum1 : h->bsum;
    p.bexp = set ? h->bexp1 : h->bexp;
    p.bpow = log_weights ? (set ? h->bpow1 : h->bpow) : nullptr;
    p.L = L;
    p.nb = (L + kWeightsBlock - 1) / kWeightsBlock;
    return p;
}

// Fixed-point weights and block scans of the step's FP32 log-weights.
// parity selects which of the two frame counters is read (the other
// receives t + 1) and which maximum slot holds the frame's maximum.
WeightArgs weight_args(pf_handle *h, const uint32_t *uni, int parity)
{
    WeightArgs w{};
    w.lw32 = h->lw;
    w.cloc = h->cloc;
    w.bsum = h->bsum;
    w.bexp = h->bexp;
    w.bpow = h->bpow;
    w.fs = h->fs;
    w.frame = h->frame + parity;
    w.frame_next = h->frame + (1 - parity);
    w.uni = uni;
    w.L = h->N;
    w.parity = parity;
    w.ustride = 2;
    w.uoff = 1;
    return w;
}

int launch_weights(pf_handle *h, const uint32_t *uni, int parity)
{
    const WeightArgs w = weight_args(h, uni, parity);
    const int nb = (h->N + kWeightsBlock - 1) / kWeightsBlock;
    hipLaunchKernelGGL((k_weights<kLogF32>), dim3(nb), dim3(kWeightsBlock), 0, h->stream, w);
    return check_launch("k_weights");
}

// Materialise the pending resampling of L weights into N ancestors.  The log
// mean weight goes to lml, or to lml[*lml_frame] when a frame counter is given.
int launch_ancestors(pf_handle *h, int L, int N, int *out, int parity, int kind_linear, double *lml,
                     const int *lml_frame = nullptr, int set = 0, const int *pow_frame = nullptr)
{
    AncArgs an{};
    an.pow_frame = pow_frame;
    an.pend = pending_of(h, L, !kind_linear, set);
    an.fs = h->fs;
    an.anc = out;
    an.lml = lml;
    an.lml_frame = lml_frame;
    an.N = N;
    an.parity = parity;
    an.kind_linear = kind_linear;
    const size_t lds = 8 * (size_t)resolve_lds_words(an.pend.nb);
    hipLaunchKernelGGL(k_ancestors, dim3((N + 255) / 256), dim3(256), lds, h->stream, an);
    return check_launch("k_ancestors");
}

StepArgs step_args(pf_handle *h, int mode, const uint32_t *uni, const float *Z, long long zstride,
                   int zld, const int *count, const float *cand, int *comp_out, int parity,
                   int anc_mode, double *lml)
{
    StepArgs a{};
    a.parity = parity;
    a.anc_mode = anc_mode;
    a.lml = lml;
    a.pend = pending_of(h, h->N);
    a.xin = h->xA;
    a.Pin = h->PA;
    a.xout = h->xB;
    a.Pout = h->PB;
    a.anc = h->anc;
    a.lw = h->lw;
    a.comp_out = comp_out;
    a.fs = h->fs;
    a.table = h->table;
    a.frame = h->frame + parity;
    a.uni = uni;
    a.Z = Z;
    a.z_frame_stride = zstride;
    a.zld = zld;
    a.count = count;
    a.cand = cand;
    a.pi_total = h->pi_total;
    a.table_floats = h->table_floats;
    a.comp_global = h->comp_global;
    a.stage_floats = h->table_floats + (mode == kModeCand ? cand_lds_floats(h->J, h->cap) : 0);
    a.K = h->K;
    a.N = h->N;
    a.ld = h->ld;
    a.m = h->m;
    a.J = h->J;
    a.cap = h->cap;
    a.order_packed = h->order_packed;
    a.clutter_mask = h->clutter_mask;
    const double pg = 1.0 - std::exp(-0.5 * h->cfg.gate_chi2);
    a.pdl = (float)(h->cfg.p_detect / h->cfg.clutter_density);
    a.b0 = (float)(1.0 - h->cfg.p_detect * pg);
    a.gate = (float)h->cfg.gate_chi2;
    return a;
}

void swap_state(pf_handle *h)
{
    std::swap(h->xA, h->xB);
    std::swap(h->PA, h->PB);
}

// The layout pf_set_state_layout asks for, resolved for this filter.
bool want_records(const pf_handle *h)
{
    if (!rec_layout(h->D) || h->layout_req == PF_LAYOUT_ROWS) return false;
    if (h->layout_req == PF_LAYOUT_RECORDS) return true;
    return (h->N + 255) / 256 <= 2 * h->cus;
}

// Views of the two state buffers in the given layout; buffer a (0 or 1)
// becomes xA / PA (the state the next step reads).
void set_views(pf_handle *h, bool records, int a)
{
    h->lay = records ? Lay{1, rec_floats(h->D)} : Lay{h->ld, 1};
    for (int q = 0; q < 2; ++q) {
        h->xbuf[q] = h->rec[q];
        h->Pbuf[q] = h->rec[q] + (records ? (size_t)h->D : (size_t)h->D * h->ld);
    }
    h->xA = h->xbuf[a]; h->PA = h->Pbuf[a];
    h->xB = h->xbuf[1 - a]; h->PB = h->Pbuf[1 - a];
}

template <int D>
struct RelayoutL {
    static void run(pf_handle *h, const float *xs, const float *Ps, Lay Ls)
    {
        hipLaunchKernelGGL((k_relayout<D>), dim3((h->N + 255) / 256), dim3(256), 0, h->stream, xs, Ps, Ls, h->xA,
                           h->PA, h->lay, h->N);
    }
};

// Waits for the handle's stream, then reads and clears its status word
// (kStatus* bits raised by the device work since the last read).
int read_status(pf_handle *h, unsigned *word)
{
    FrameScratch fs;
    HIP_TRY(hipMemcpyAsync(&fs, h->fs, sizeof(fs), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    *word = fs.status;
    if (fs.status) {
        const unsigned int zero = 0;
        HIP_TRY(hipMemcpyAsync(&h->fs->status, &zero, sizeof(zero), hipMemcpyHostToDevice, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
    }
    return PF_OK;
}


// Builds the LDS model image, the reset table and the basis matrices of the
// handle's model and uploads them.
int build_tables(pf_handle *h)
{
    drop_graphs(h);
    ModelImage img;
    int st = make_model_image(h->hm, h->order_packed, h->N, img);
    if (st) return st;
    h->sel_mode = img.sel_mode;
    h->pi_total = img.pi_total;
    (void)hipFree(h->table); (void)hipFree(h->mu); (void)hipFree(h->sig);
    (void)hipFree(h->Tdev); (void)hipFree(h->Tinv);
    h->table = nullptr; h->mu = nullptr; h->sig = nullptr; h->Tdev = nullptr; h->Tinv = nullptr;
    HIP_TRY(dalloc(&h->table, img.blob.size()));
    HIP_TRY(dalloc(&h->mu, img.mu.size()));
    HIP_TRY(dalloc(&h->sig, img.sg.size()));
    HIP_TRY(dalloc(&h->Tdev, img.T.size()));
    HIP_TRY(dalloc(&h->Tinv, img.Ti.size()));
    HIP_TRY(hipMemcpyAsync(h->table, img.blob.data(), sizeof(float) * img.blob.size(), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(h->mu, img.mu.data(), sizeof(float) * img.mu.size(), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(h->sig, img.sg.data(), sizeof(float) * img.sg.size(), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(h->Tdev, img.T.data(), sizeof(double) * img.T.size(), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(h->Tinv, img.Ti.data(), sizeof(double) * img.Ti.size(), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->table_floats = img.comp_in_lds ? (int)img.blob.size() : img.head_floats;
    h->comp_global = img.comp_in_lds ? nullptr : h->table + img.head_floats;
    return PF_OK;
}
}  // namespace

// ------------------------------------------- interface for chained handles
namespace mkfbt {
namespace eng {

int info(pf_handle *h, Info *out)
{
    int st = check_handle(h);
    if (st) return st;
    out->N = h->N;
    out->D = h->D;
    out->m = h->m;
    out->J = h->J;
    out->npose = h->npose;
    out->device = h->cfg.device;
    out->model = h->model;
    out->tracks = h->tracks;
    out->stream = h->stream;
    return PF_OK;
}

int reserve_resample(pf_handle *h, int L) { return ensure_resample_scratch(h, L); }

const double *estimate_ptr(pf_handle *h) { return h->est_out; }


int set_frame(pf_handle *h, int f)
{
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, h->stream, h->frame, f);
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, h->stream, h->frame + 1, f);
    // the exponent slot of frame f (a step that writes its own weights
    // raises it; later slots are cleared by the step before them)
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, h->stream, &h->fs->pow_max[f % 3], kPowEmpty);
    return check_launch("k_set_int");
}

int estimate(pf_handle *h)
{
    int st = dispatch_d<EstimateL>(h->D, h);
    if (st) return st;
    return check_launch("k_estimate");
}

Pending pending(pf_handle *h, int L) { return pending_of(h, L, false); }

FrameScratch *scratch(pf_handle *h) { return h->fs; }

template <int D>
struct AncEstimateL {
    static void run(pf_handle *h, const AncArgs &an)
    {
        const int nb = (h->N + 255) / 256;
        const size_t lds = (size_t)anc_estimate_lds_bytes(an.pend.nb, D);
        hipLaunchKernelGGL((k_ancestors_estimate<D>), dim3(nb), dim3(256), lds, h->stream, an, (const float *)h->xB,
                           h->lay, h->est_partial);
        hipLaunchKernelGGL((k_estimate_final<D>), dim3(1), dim3(256), 0, h->stream, (const double *)h->est_partial, nb,
                           h->N, (const double *)h->Tinv, h->est_out);
    }
};

// The frame's resampling of the filter's own weights (k_weights of the
// current frame), its log mean weight to lml[*frame].
AncArgs frame_anc_args(pf_handle *h, int p, double *lml)
{
    AncArgs an{};
    an.pend = pending_of(h, h->N);
    an.fs = h->fs;
    an.anc = h->anc;
    an.lml = lml;
    an.lml_frame = h->frame + p;
    an.N = h->N;
    an.parity = p;
    an.kind_linear = 0;
    return an;
}

// With blocks of kWeightsBlock particles the step kernel writes the frame's
// resampler weights itself (scratch set 0) and records their largest block
// exponent in fs->pow_max[t mod 3] (no weights launch); returns whether it does.
bool fuse_step_weights(pf_handle *h, StepArgs &a, int p)
{
    if ((h->cfg.block_size > 0 ? h->cfg.block_size : 256) != kWeightsBlock) return false;
    a.w_cloc = h->cloc;
    a.w_bsum = h->bsum;
    a.w_bexp = h->bexp;
    a.w_bpow = h->bpow;
    a.frame_next = h->frame + (1 - p);
    return true;
}

int step_ref(pf_handle *h, const float *Z, const uint32_t *uni, int parity, double *lml, bool with_estimate)
{
    const int p = parity & 1;
    StepArgs a = step_args(h, kModeRef, uni, Z, 0, h->N, nullptr, nullptr, nullptr, p, kAncArray, nullptr);
    const bool fused = fuse_step_weights(h, a, p);
    int st = launch_step(h, a, kModeRef);
    if (st) return st;
    if (!fused) st = launch_weights(h, uni, p);
    if (st) return st;
    if (with_estimate) {
        AncArgs an = frame_anc_args(h, p, lml);
        if (fused) an.pow_frame = h->frame + p;
        st = dispatch_d<AncEstimateL>(h->D, h, an);      // reads the new state, h->xB before the swap
        if (!st) st = check_launch("k_ancestors_estimate");
    } else {
        st = launch_ancestors(h, h->N, h->N, h->anc, p, 0, lml, h->frame + p, 0, fused ? h->frame + p : nullptr);
    }
    if (st) return st;
    swap_state(h);
    return PF_OK;
}

// The pair launches are compiled for the selection basis with 3 joints (the
// tracker's 6-row measurement) at every state dimension.
bool pair_compatible(pf_handle *h0, pf_handle *h1)
{
    return h0 && h1 && h0 != h1 && h0->model && h1->model && h0->D == h1->D && h0->N == h1->N &&
           h0->cfg.device == h1->cfg.device && h0->sel_mode && h1->sel_mode && h0->J == 3 && h1->J == 3 &&
           h0->m == h1->m && h0->table_floats == h1->table_floats && h0->cfg.block_size == h1->cfg.block_size &&
           (h0->D == 8 || h0->D == 12 || h0->D == 16 || h0->D == 24);
}

template <int D>
struct PairL {
    static void run(pf_handle *h0, pf_handle *h1, const StepArgs &a0, const StepArgs &a1, const WeightArgs &w0,
                    const WeightArgs &w1, const AncEstArgs &e0, const AncEstArgs &e1, const EstFinalArgs &f0,
                    const EstFinalArgs &f1, int block, bool fused, bool final_sums)
    {
        hipStream_t s = h0->stream;
        const dim3 sg((h0->N + block - 1) / block, 2);
        hipLaunchKernelGGL((k_step_pair<D, kModeRef, 3>), sg, dim3(block),
                           (size_t)step_lds_bytes(a0.stage_floats, 0), s, a0, a1);
        const dim3 wg((h0->N + kWeightsBlock - 1) / kWeightsBlock, 2);
        if (!fused) hipLaunchKernelGGL((k_weights_pair<kLogF32>), wg, dim3(kWeightsBlock), 0, s, w0, w1);
        const int nb = (h0->N + 255) / 256;
        hipLaunchKernelGGL((k_ancestors_estimate_pair<D>), dim3(nb, 2), dim3(256),
                           (size_t)anc_estimate_lds_bytes(e0.an.pend.nb, D), s, e0, e1);
        if (final_sums) hipLaunchKernelGGL((k_estimate_final_pair<D>), dim3(2), dim3(256), 0, s, f0, f1);
    }
};

EstFinalArgs estimate_final_args(pf_handle *h)
{
    EstFinalArgs f{};
    f.partial = h->est_partial;
    f.Ti = h->Tinv;
    f.out = h->est_out;
    f.nb = (h->N + 255) / 256;
    f.N = h->N;
    return f;
}

int step_ref_pair(pf_handle *h0, pf_handle *h1, const float *Z0, const float *Z1, const uint32_t *uni0,
                  const uint32_t *uni1, int parity, double *lml0, double *lml1, bool final_sums)
{
    if (!pair_compatible(h0, h1)) return set_error(PF_ERR_UNSUPPORTED, "the two filters cannot share launches");
    const int p = parity & 1;
    pf_handle *h[2] = {h0, h1};
    const float *Z[2] = {Z0, Z1};
    const uint32_t *uni[2] = {uni0, uni1};
    double *lml[2] = {lml0, lml1};
    StepArgs a[2];
    WeightArgs w[2];
    AncEstArgs e[2];
    EstFinalArgs f[2];
    bool fused = true;
    for (int i = 0; i < 2; ++i) {
        a[i] = step_args(h[i], kModeRef, uni[i], Z[i], 0, h[i]->N, nullptr, nullptr, nullptr, p, kAncArray, nullptr);
        fused = fuse_step_weights(h[i], a[i], p) && fused;
        w[i] = weight_args(h[i], uni[i], p);
        e[i].an = frame_anc_args(h[i], p, lml[i]);
        if (fused) e[i].an.pow_frame = h[i]->frame + p;
        e[i].x = h[i]->xB;                      // the updated state, before the swap
        e[i].partial = h[i]->est_partial;
        e[i].lay = h[i]->lay;
        f[i] = estimate_final_args(h[i]);
    }
    if (a[0].stage_floats != a[1].stage_floats)
        return set_error(PF_ERR_UNSUPPORTED, "the two filters stage different model images");
    const int block = h0->cfg.block_size > 0 ? h0->cfg.block_size : 256;
    int st = dispatch_d<PairL>(h0->D, h0, h1, a[0], a[1], w[0], w[1], e[0], e[1], f[0], f[1], block, fused,
                               final_sums);
    if (!st) st = check_launch("pair step");
    if (st) return st;
    swap_state(h0);
    swap_state(h1);
    return PF_OK;
}

int take_status(pf_handle *h, unsigned *word) { return read_status(h, word); }

}  // namespace eng
}  // namespace mkfbt

// =================================================================== C ABI
extern "C" {

void pf_config_default(pf_config *cfg)
{
    if (!cfg) return;
    cfg->n_particles = 500;
    cfg->device = 0;
    cfg->cand_capacity = 16;
    cfg->block_size = 0;
    cfg->p_detect = 0.95;
    cfg->gate_chi2 = 9.210340371976184;  // chi-square 2 dof, 99%
    cfg->clutter_density = 3.0 / (640.0 * 480.0);
}

const char *pf_last_error(void) { return g_last_error.c_str(); }

int pf_create(pf_handle **out, const pf_config *cfg)
{
    if (!out) return set_error(PF_ERR_ARG, "null output pointer");
    *out = nullptr;
    pf_config c;
    pf_config_default(&c);
    if (cfg) c = *cfg;
    if (c.n_particles <= 0 || c.n_particles > kMaxScanBlocks * kWeightsBlock)
        return set_error(PF_ERR_ARG, "n_particles %d out of range (1..%d)", c.n_particles,
                         kMaxScanBlocks * kWeightsBlock);
    if (c.cand_capacity <= 0) c.cand_capacity = 16;
    if (c.cand_capacity > 256) return set_error(PF_ERR_ARG, "cand_capacity %d > 256", c.cand_capacity);
    if (c.block_size != 0 && (c.block_size % 64 != 0 || c.block_size > 256 || c.block_size < 64))
        return set_error(PF_ERR_ARG, "block_size must be 64, 128, 192 or 256");
    if (!(c.p_detect > 0.0 && c.p_detect <= 1.0)) return set_error(PF_ERR_ARG, "p_detect must be in (0,1]");
    if (!(c.gate_chi2 > 0.0)) return set_error(PF_ERR_ARG, "gate_chi2 must be > 0");
    if (!(c.clutter_density > 0.0)) return set_error(PF_ERR_ARG, "clutter_density must be > 0");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return set_error(PF_ERR_HIP, "no HIP device available");
    if (c.device < 0 || c.device >= ndev) return set_error(PF_ERR_ARG, "device %d out of range", c.device);
    HIP_TRY(hipSetDevice(c.device));
    pf_handle *h = new pf_handle();
    h->cfg = c;
    h->N = c.n_particles;
    h->ld = (c.n_particles + 63) & ~63;
    h->cap = c.cand_capacity;
    if (hipDeviceGetAttribute(&h->cus, hipDeviceAttributeMultiprocessorCount, c.device) != hipSuccess ||
        h->cus <= 0)
        h->cus = 256;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return set_error(PF_ERR_HIP, "hipStreamCreate failed");
    }
    hipError_t e = hipSuccess;
    e = e ? e : dalloc(&h->fs, 1);
    e = e ? e : dalloc(&h->frame, 2);
    e = e ? e : dalloc(&h->uni, 2);
    e = e ? e : dalloc(&h->lml1, 1);
    e = e ? e : dalloc(&h->anc, (size_t)h->ld);
    e = e ? e : dalloc(&h->comp, (size_t)h->ld);
    e = e ? e : dalloc(&h->bins, (size_t)h->ld);
    e = e ? e : dalloc(&h->lw, (size_t)h->ld);
    e = e ? e : dalloc(&h->est_out, 64);
    e = e ? e : dalloc(&h->est_partial, (size_t)((h->N + 255) / 256) * 64);
    e = e ? e : dalloc(&h->run_desc, 1);
    e = e ? e : dalloc(&h->run_ctr, 2);
    if (e == hipSuccess) e = hipMemsetAsync(h->fs, 0, sizeof(FrameScratch), h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
        pf_destroy(h);
        return set_error(PF_ERR_HIP, "device allocation failed: %s", hipGetErrorString(e));
    }
    if (ensure_resample_scratch(h, h->N) != PF_OK) {
        pf_destroy(h);
        return PF_ERR_HIP;
    }
    *out = h;
    return PF_OK;
}

int pf_destroy(pf_handle *h)
{
    if (!h) return PF_OK;
    (void)hipSetDevice(h->cfg.device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    drop_graphs(h);
    void *ptrs[] = {h->run_desc, h->run_ctr, h->rec[0], h->rec[1], h->anc, h->comp, h->bins, h->lw, h->table,
                    h->mu, h->sig, h->fs, h->frame, h->uni, h->lml1, h->cloc, h->bsum, h->bexp, h->bpow,
                    h->cloc1, h->bsum1, h->bexp1, h->bpow1, h->zbuf, h->cntbuf, h->candbuf, h->est_partial, h->est_out, h->gbuf, h->ibuf,
                    h->Tdev, h->Tinv};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (hipEvent_t e : h->ev) (void)hipEventDestroy(e);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return PF_OK;
}

int pf_load_model(pf_handle *h, int32_t K, int32_t d, int32_t npose, int32_t m, const int32_t *sel,
                  const double *means, const double *covs, const double *weights,
                  const double *gamma, const double *pca_proj, const double *pca_mean,
                  const double *R)
{
    int st = check_handle(h);
    if (st) return st;
    HostModel hm;
    st = host_model_from_args(hm, K, d, npose, m, sel, means, covs, weights, gamma, pca_proj, pca_mean, R);
    if (st) return st;
    const int NP = npacked(d), J = m / 2;
    h->hm = hm;
    h->D = d; h->K = K; h->m = m; h->J = J; h->npose = npose;

    // (re)allocate device buffers sized by the model
    (void)hipFree(h->rec[0]); (void)hipFree(h->rec[1]);
    h->rec[0] = h->rec[1] = nullptr;
    h->xA = h->PA = h->xB = h->PB = nullptr;
    (void)hipFree(h->zbuf); (void)hipFree(h->cntbuf); (void)hipFree(h->candbuf);
    h->zbuf = nullptr; h->cntbuf = nullptr; h->candbuf = nullptr;
    h->model = false;
    h->tracks = false;
    // room for either layout (a record is padded to 16 bytes)
    const size_t per = rec_layout(d) ? (size_t)rec_floats(d) : (size_t)(d + NP);
    for (int q = 0; q < 2; ++q) HIP_TRY(dalloc(&h->rec[q], per * h->ld));
    set_views(h, want_records(h), 0);
    h->clutter_mask = 0;
    h->order_packed = 0;
    for (int b = 0; b < J; ++b) h->order_packed |= (unsigned)b << (4 * b);
    st = build_tables(h);
    if (st) return st;
    h->model = true;
    return PF_OK;
}

int pf_set_joint_clutter(pf_handle *h, const int32_t *clutter)
{
    int st = check_handle(h);
    if (st) return st;
    if (!h->model) return set_error(PF_ERR_STATE, "load a model first");
    unsigned mask = 0;
    for (int b = 0; b < h->J; ++b)
        if (clutter && clutter[b]) mask |= 1u << b;
    // certain joints first, then PDAF joints, each in index order
    const unsigned order = joint_order(mask, h->J);
    if (h->tracks && (mask != h->clutter_mask))
        return set_error(PF_ERR_STATE, "joint kinds must be set before the tracks are initialised");
    h->clutter_mask = mask;
    h->order_packed = order;
    return build_tables(h);
}

int pf_reset(pf_handle *h, const int32_t *bins)
{
    int st = check_handle(h);
    if (st) return st;
    if (!h->model) return set_error(PF_ERR_STATE, "load a model first");
    if (!bins) return set_error(PF_ERR_ARG, "null bins");
    for (int j = 0; j < h->N; ++j)
        if (bins[j] < 0 || bins[j] >= h->K) return set_error(PF_ERR_ARG, "bins[%d]=%d out of range", j, bins[j]);
    HIP_TRY(hipMemcpyAsync(h->bins, bins, sizeof(int) * h->N, hipMemcpyHostToDevice, h->stream));
    st = dispatch_d<ResetL>(h->D, h);
    if (st) return st;
    hipLaunchKernelGGL(k_iota, dim3((h->N + 255) / 256), dim3(256), 0, h->stream, h->anc, h->N);
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->tracks = true;
    return PF_OK;
}

int pf_set_tracks(pf_handle *h, const double *x, const double *P)
{
    int st = check_handle(h);
    if (st) return st;
    if (!h->model) return set_error(PF_ERR_STATE, "load a model first");
    if (!x || !P) return set_error(PF_ERR_ARG, "null tracks");
    const int D = h->D, N = h->N;
    const size_t nx = (size_t)N * D, np = (size_t)N * D * D;
    st = ensure_gbuf(h, sizeof(double) * (nx + np));
    if (st) return st;
    HIP_TRY(hipMemcpyAsync(h->gbuf, x, sizeof(double) * nx, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(h->gbuf + nx, P, sizeof(double) * np, hipMemcpyHostToDevice, h->stream));
    st = dispatch_d<ImportL>(D, h, (const double *)h->gbuf, (const double *)(h->gbuf + nx));
    if (st) return st;
    hipLaunchKernelGGL(k_iota, dim3((h->N + 255) / 256), dim3(256), 0, h->stream, h->anc, h->N);
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->tracks = true;
    return PF_OK;
}

int pf_get_tracks(pf_handle *h, double *x, double *P)
{
    int st = check_handle(h);
    if (st) return st;
    if (!h->tracks) return set_error(PF_ERR_STATE, "tracks not initialised");
    const int D = h->D, N = h->N;
    const size_t nx = (size_t)N * D, np = (size_t)N * D * D;
    st = ensure_gbuf(h, sizeof(double) * (nx + np));
    if (st) return st;
    st = dispatch_d<ExportL>(D, h, h->gbuf, h->gbuf + nx);
    if (st) return st;
    if (x) HIP_TRY(hipMemcpyAsync(x, h->gbuf, sizeof(double) * nx, hipMemcpyDeviceToHost, h->stream));
    if (P) HIP_TRY(hipMemcpyAsync(P, h->gbuf + nx, sizeof(double) * np, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return PF_OK;
}

static int finish_host_step(pf_handle *h, float *lw, int32_t *anc, int32_t *comp, double *lml)
{
    int st = launch_weights(h, h->uni, 0);
    if (st) return st;
    st = launch_ancestors(h, h->N, h->N, h->anc, 0, 0, h->lml1);
    if (st) return st;
    swap_state(h);
    if (lw) HIP_TRY(hipMemcpyAsync(lw, h->lw, sizeof(float) * h->N, hipMemcpyDeviceToHost, h->stream));
    if (anc) HIP_TRY(hipMemcpyAsync(anc, h->anc, sizeof(int) * h->N, hipMemcpyDeviceToHost, h->stream));
    if (comp) HIP_TRY(hipMemcpyAsync(comp, h->comp, sizeof(int) * h->N, hipMemcpyDeviceToHost, h->stream));
    if (lml) HIP_TRY(hipMemcpyAsync(lml, h->lml1, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    unsigned word = 0;
    st = read_status(h, &word);
    if (st) return st;
    return status_code(word);
}

int pf_step(pf_handle *h, const double *Z, uint32_t u_comp, uint32_t u_res, float *lw, int32_t *anc,
            int32_t *comp, double *lml)
{
    int st = check_handle(h);
    if (st) return st;
    if (!h->tracks) return set_error(PF_ERR_STATE, "tracks not initialised (pf_reset or pf_set_tracks)");
    if (!Z) return set_error(PF_ERR_ARG, "null measurement");
    const int m = h->m, N = h->N;
    if (!h->zbuf) HIP_TRY(dalloc(&h->zbuf, (size_t)m * N));
    std::vector<float> zf((size_t)m * N);
    for (size_t i = 0; i < zf.size(); ++i) zf[i] = (float)Z[i];
    const uint32_t u[2] = {u_comp, u_res};
    HIP_TRY(hipMemcpyAsync(h->zbuf, zf.data(), sizeof(float) * zf.size(), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(h->uni, u, sizeof(u), hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, h->stream, h->frame, 0);
    StepArgs a = step_args(h, kModeRef, h->uni, h->zbuf, 0, N, nullptr, nullptr, comp ? h->comp : nullptr, 0,
                           kAncArray, nullptr);
    st = launch_step(h, a, kModeRef);
    if (st) return st;
    return finish_host_step(h, lw, anc, comp, lml);
}

int pf_step_candidates(pf_handle *h, const int32_t *count, const double *cand, uint32_t u_comp,
                       uint32_t u_res, float *lw, int32_t *anc, int32_t *comp, double *lml)
{
    int st = check_handle(h);
    if (st) return st;
    if (!h->tracks) return set_error(PF_ERR_STATE, "tracks not initialised (pf_reset or pf_set_tracks)");
    if (!count || !cand) return set_error(PF_ERR_ARG, "null candidates");
    const int J = h->J, cap = h->cap;
    for (int b = 0; b < J; ++b)
        if (count[b] < 0 || count[b] > cap) return set_error(PF_ERR_ARG, "count[%d]=%d outside 0..%d", b, count[b], cap);
    if (!h->cntbuf) HIP_TRY(dalloc(&h->cntbuf, (size_t)J));
    if (!h->candbuf) HIP_TRY(dalloc(&h->candbuf, (size_t)J * cap * 2));
    std::vector<float> cf((size_t)J * cap * 2);
    for (size_t i = 0; i < cf.size(); ++i) cf[i] = (float)cand[i];
    const uint32_t u[2] = {u_comp, u_res};
    HIP_TRY(hipMemcpyAsync(h->cntbuf, count, sizeof(int) * J, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(h->candbuf, cf.data(), sizeof(float) * cf.size(), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(h->uni, u, sizeof(u), hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, h->stream, h->frame, 0);
    StepArgs a = step_args(h, kModeCand, h->uni, nullptr, 0, 0, h->cntbuf, h->candbuf, comp ? h->comp : nullptr, 0,
                           kAncArray, nullptr);
    st = launch_step(h, a, kModeCand);
    if (st) return st;
    return finish_host_step(h, lw, anc, comp, lml);
}

int pf_estimate(pf_handle *h, double *xbar)
{
    int st = check_handle(h);
    if (st) return st;
    if (!h->tracks) return set_error(PF_ERR_STATE, "tracks not initialised");
    if (!xbar) return set_error(PF_ERR_ARG, "null output");
    st = dispatch_d<EstimateL>(h->D, h);
    if (st) return st;
    HIP_TRY(hipMemcpyAsync(xbar, h->est_out, sizeof(double) * h->D, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return PF_OK;
}

static int resample_generic(pf_handle *h, const double *w, int L, int N, uint32_t u, int32_t *out, int kind)
{
    int st = check_handle(h);
    if (st) return st;
    if (!w || !out || L <= 0 || N <= 0 || L > kMaxScanBlocks * kWeightsBlock || N > (1 << 26))
        return set_error(PF_ERR_ARG, "bad resample arguments (L=%d, N=%d)", L, N);
    st = ensure_resample_scratch(h, L);
    if (st) return st;
    const size_t wbytes = sizeof(double) * (size_t)L;
    st = ensure_gbuf(h, wbytes);
    if (st) return st;
    st = ensure_ibuf(h, sizeof(int) * (size_t)N + 64);
    if (st) return st;
    int *dout = h->ibuf;
    uint32_t *duni = reinterpret_cast<uint32_t *>(h->ibuf + N);
    const uint32_t uu[2] = {u, u};
    HIP_TRY(hipMemcpyAsync(h->gbuf, w, wbytes, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(duni, uu, sizeof(uu), hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, h->stream, h->frame, 0);
    const int nbm = (L + 255) / 256;
    if (kind == kLinF64)      // linear weights are scaled by their maximum (log-weights per block)
        hipLaunchKernelGGL((k_blockmax64<kLinF64>), dim3(nbm), dim3(256), 0, h->stream, h->gbuf, L, h->fs);
    WeightArgs wa{};
    wa.w64 = h->gbuf;
    wa.cloc = h->cloc;
    wa.bsum = h->bsum;
    wa.bexp = h->bexp;
    wa.bpow = h->bpow;
    wa.fs = h->fs;
    wa.frame = h->frame;
    wa.frame_next = nullptr;
    wa.uni = duni;
    wa.L = L;
    wa.parity = 0;
    wa.ustride = 2;
    wa.uoff = 1;
    const int nb = (L + kWeightsBlock - 1) / kWeightsBlock;
    if (kind == kLogF64)
        hipLaunchKernelGGL((k_weights<kLogF64>), dim3(nb), dim3(kWeightsBlock), 0, h->stream, wa);
    else
        hipLaunchKernelGGL((k_weights<kLinF64>), dim3(nb), dim3(kWeightsBlock), 0, h->stream, wa);
    st = check_launch("k_weights");
    if (st) return st;
    st = launch_ancestors(h, L, N, dout, 0, kind == kLinF64, nullptr);
    if (st) return st;
    HIP_TRY(hipMemcpyAsync(out, dout, sizeof(int) * (size_t)N, hipMemcpyDeviceToHost, h->stream));
    unsigned word = 0;
    st = read_status(h, &word);
    if (st) return st;
    return status_code(word);
}

int pf_resample(pf_handle *h, const double *w, int32_t L, int32_t N, uint32_t u, int32_t *out)
{
    return resample_generic(h, w, L, N, u, out, kLinF64);
}

int pf_resample_log(pf_handle *h, const double *lw, int32_t L, int32_t N, uint32_t u, int32_t *out)
{
    return resample_generic(h, lw, L, N, u, out, kLogF64);
}

static int samples_common(pf_handle *h, const double *Hp, const double *M, int npose, int n)
{
    int st = check_handle(h);
    if (st) return st;
    if (!h->tracks) return set_error(PF_ERR_STATE, "tracks not initialised");
    if (!Hp || !M || npose < 2 || n <= 0) return set_error(PF_ERR_ARG, "bad sample arguments");
    return dispatch_d<EstimateL>(h->D, h);
}

int pf_get_samples(pf_handle *h, const double *Hp, const double *M, int32_t npose, int32_t n,
                   double scale, const double *normals, double *out)
{
    int st = samples_common(h, Hp, M, npose, n);
    if (st) return st;
    if (!normals || !out) return set_error(PF_ERR_ARG, "null normals/output");
    const int D = h->D;
    const size_t need = sizeof(double) * (size_t)(2 * D + 2 + 2 * (size_t)n + 2 * (size_t)n);
    st = ensure_gbuf(h, need);
    if (st) return st;
    double *dh = h->gbuf, *dm = dh + 2 * D, *dn = dm + 2, *dout = dn + 2 * (size_t)n;
    HIP_TRY(hipMemcpyAsync(dh, Hp, sizeof(double) * 2 * D, hipMemcpyHostToDevice, h->stream));  // rows 0,1
    HIP_TRY(hipMemcpyAsync(dm, M, sizeof(double) * 2, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(dn, normals, sizeof(double) * 2 * (size_t)n, hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_samples, dim3((n + 255) / 256), dim3(256), 0, h->stream, dh, dm, h->est_out, D, dn,
                       n, scale, dout);
    st = check_launch("k_samples");
    if (st) return st;
    HIP_TRY(hipMemcpyAsync(out, dout, sizeof(double) * 2 * (size_t)n, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return PF_OK;
}

int pf_get_sample_prob(pf_handle *h, const double *Hp, const double *M, int32_t npose,
                       const double *in1, const double *in2, int32_t n, double scale, double *w1,
                       double *w2)
{
    int st = samples_common(h, Hp, M, npose, n);
    if (st) return st;
    if (!in1 || !in2 || !w1 || !w2) return set_error(PF_ERR_ARG, "null inputs/outputs");
    const int D = h->D;
    const size_t need = sizeof(double) * (size_t)(2 * D + 2 + 6 * (size_t)n);
    st = ensure_gbuf(h, need);
    if (st) return st;
    double *dh = h->gbuf, *dm = dh + 2 * D, *d1 = dm + 2, *d2 = d1 + 2 * (size_t)n;
    double *o1 = d2 + 2 * (size_t)n, *o2 = o1 + n;
    HIP_TRY(hipMemcpyAsync(dh, Hp, sizeof(double) * 2 * D, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(dm, M, sizeof(double) * 2, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(d1, in1, sizeof(double) * 2 * (size_t)n, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(d2, in2, sizeof(double) * 2 * (size_t)n, hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_sample_prob, dim3((n + 255) / 256), dim3(256), 0, h->stream, dh, dm, h->est_out,
                       D, d1, d2, n, scale, o1, o2);
    st = check_launch("k_sample_prob");
    if (st) return st;
    HIP_TRY(hipMemcpyAsync(w1, o1, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipMemcpyAsync(w2, o2, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return PF_OK;
}

// ------------------------------------------------ device-resident sequences
}  // extern "C"

namespace {

// The step of a device-resident run for frames of parity p (the frame
// index read from counter p, everything else chosen in the kernel from it;
// see StepArgs::run).
StepArgs run_step_args(pf_handle *h, int mode, int p)
{
    StepArgs a = step_args(h, mode, nullptr, nullptr, (long long)h->m * h->N, h->N, nullptr, nullptr, nullptr, p,
                           kAncFused, nullptr);
    a.run = h->run_desc;
    a.frame = h->run_ctr + p;
    a.frame_next = h->run_ctr + (1 - p);
    a.xin = h->xbuf[0];
    a.Pin = h->Pbuf[0];
    a.xout = h->xbuf[1];
    a.Pout = h->Pbuf[1];
    a.pend = pending_of(h, h->N, true, 1);     // even frames resolve set 1
    a.pend1 = pending_of(h, h->N, true, 0);    // odd frames resolve set 0
    a.w_cloc = h->cloc;
    a.w_bsum = h->bsum;
    a.w_bexp = h->bexp;
    a.w_bpow = h->bpow;
    a.w1_cloc = h->cloc1;
    a.w1_bsum = h->bsum1;
    a.w1_bexp = h->bexp1;
    a.w1_bpow = h->bpow1;
    return a;
}

int launch_run_step(pf_handle *h, int mode, int p)
{
    return launch_step_block(h, run_step_args(h, mode, p), mode, kWeightsBlock);
}

// Group g of the run frames starting at a frame of parity p: 2^g step
// launches, captured once per handle, mode and parity.
int run_group(pf_handle *h, int mode, int p, int g, hipGraphExec_t *out)
{
    hipGraphExec_t &ge = h->groups[mode][p][g];
    if (!ge) {
        hipGraph_t graph;
        HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        int sc = 0;
        for (int r = 0; r < (1 << g) && !sc; ++r) sc = launch_run_step(h, mode, (p + r) & 1);
        hipError_t ec = hipStreamEndCapture(h->stream, &graph);
        if (sc) return sc;
        if (ec != hipSuccess) return set_error(PF_ERR_HIP, "graph capture failed: %s", hipGetErrorString(ec));
        ec = hipGraphInstantiate(&ge, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ec != hipSuccess) {
            ge = nullptr;
            return set_error(PF_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(ec));
        }
    }
    *out = ge;
    return PF_OK;
}

int ensure_events(pf_handle *h, size_t need)
{
    while (h->ev.size() < need) {
        hipEvent_t e;
        HIP_TRY(hipEventCreate(&e));
        h->ev.push_back(e);
    }
    return PF_OK;
}

// Frames whose model image and resolver workspace do not fit one workgroup's
// LDS: step, weights and ancestor launches per frame, launched one by one.
int run_frames_split(pf_handle *h, int T, int mode, const float *Z, const int *count, const float *cand,
                     const uint32_t *uniforms, double *lml, bool timing, bool run_timing)
{
    const long long zstride = (long long)h->m * h->N;
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, h->stream, h->frame, 0);
    int st = check_launch("k_set_int");
    if (st) return st;
    if (run_timing) HIP_TRY(hipEventRecord(h->ev[0], h->stream));
    if (timing) HIP_TRY(hipEventRecord(h->ev[2 * T], h->stream));
    for (int f = 0; f < T; ++f) {
        const int parity = f & 1;
        StepArgs a = step_args(h, mode, uniforms, Z, zstride, h->N, count, cand, nullptr, parity, kAncArray, lml);
        if (timing) HIP_TRY(hipEventRecord(h->ev[2 * f], h->stream));
        st = launch_step(h, a, mode);
        if (st) return st;
        if (timing) HIP_TRY(hipEventRecord(h->ev[2 * f + 1], h->stream));
        st = launch_weights(h, uniforms, parity);
        if (st) return st;
        // counter f & 1 still holds this frame's index (k_weights wrote the
        // next one into the other counter)
        st = launch_ancestors(h, h->N, h->N, h->anc, parity, 0, lml, h->frame + parity);
        if (st) return st;
        swap_state(h);
    }
    if (timing) HIP_TRY(hipEventRecord(h->ev[2 * T + 1], h->stream));
    if (run_timing) HIP_TRY(hipEventRecord(h->ev[1], h->stream));
    return PF_OK;
}

}  // namespace

extern "C" {

static int run_frames(pf_handle *h, int T, int mode, const float *Z, const int *count,
                      const float *cand, const uint32_t *uniforms, double *lml, int flags)
{
    int st = check_handle(h);
    if (st) return st;
    if (!h->tracks) return set_error(PF_ERR_STATE, "tracks not initialised (pf_reset or pf_set_tracks)");
    if (T <= 0 || !uniforms || !lml) return set_error(PF_ERR_ARG, "bad run arguments");
    const bool timing = (flags & 2) != 0;                 // an event pair around every launch (no graph)
    const bool run_timing = (flags & 4) != 0 && !timing;  // one event pair around all frames
    const bool graph = (flags & 1) != 0 && !timing;
    h->graph_frames = 0;
    st = ensure_events(h, timing ? 2 * (size_t)T + 2 : 2);
    if (st) return st;
    // A frame is one step launch when the step's LDS holds the model image,
    // the frame's detections and the resolver's workspace: it resolves its
    // ancestors from the previous frame's resampler weights and writes its
    // own.  Otherwise the frame runs as separate launches, without graphs.
    const int nbp = (h->N + kWeightsBlock - 1) / kWeightsBlock;
    const int stage = h->table_floats + (mode == kModeCand ? cand_lds_floats(h->J, h->cap) : 0);
    const bool fused = step_lds_bytes_rec(h->lay.row == 1, stage, nbp) <= kMaxLdsBytes;
    if (!fused) {
        st = run_frames_split(h, T, mode, Z, count, cand, uniforms, lml, timing, run_timing);
    } else {
        st = ensure_run_scratch(h);
        if (st) return st;
        RunDesc v{};
        v.Z = Z;
        v.count = count;
        v.cand = cand;
        v.uni = uniforms;
        v.lml = lml;
        v.flip = (h->xA == h->xbuf[1]) ? 1 : 0;
        hipLaunchKernelGGL(k_run_begin, dim3(1), dim3(64), 0, h->stream, h->run_desc, v, h->run_ctr, h->fs);
        st = check_launch("k_run_begin");
        if (st) return st;
        if (graph) {
            // every group is captured before the first frame is enqueued
            hipGraphExec_t ge[2][pf_handle::kGroups];
            for (int p = 0; p < 2; ++p)
                for (int g = 0; g < pf_handle::kGroups; ++g) {
                    st = run_group(h, mode, p, g, &ge[p][g]);
                    if (st) return st;
                }
            if (run_timing) HIP_TRY(hipEventRecord(h->ev[0], h->stream));
            int f = 0;
            for (int g = pf_handle::kGroups - 1; g >= 0; --g)
                while (T - f >= (1 << g)) {
                    HIP_TRY(hipGraphLaunch(ge[f & 1][g], h->stream));
                    f += 1 << g;
                }
            h->graph_frames = T;
            if (run_timing) HIP_TRY(hipEventRecord(h->ev[1], h->stream));
        } else {
            if (run_timing) HIP_TRY(hipEventRecord(h->ev[0], h->stream));
            if (timing) HIP_TRY(hipEventRecord(h->ev[2 * T], h->stream));
            for (int f = 0; f < T; ++f) {
                if (timing) HIP_TRY(hipEventRecord(h->ev[2 * f], h->stream));
                st = launch_run_step(h, mode, f & 1);
                if (st) return st;
                if (timing) HIP_TRY(hipEventRecord(h->ev[2 * f + 1], h->stream));
            }
            if (timing) HIP_TRY(hipEventRecord(h->ev[2 * T + 1], h->stream));
            if (run_timing) HIP_TRY(hipEventRecord(h->ev[1], h->stream));
        }
        // frame t wrote buffer 1 - ((t + flip) & 1): the tracks now live in
        // buffer (T + flip) & 1
        if (T & 1) swap_state(h);
        // the last frame's ancestors and log mean weight
        st = launch_ancestors(h, h->N, h->N, h->anc, (T - 1) & 1, 0, lml + (T - 1), nullptr, (T - 1) & 1);
    }
    if (st) return st;
    if (run_timing) {
        HIP_TRY(hipEventSynchronize(h->ev[1]));
        float tot = 0.f;
        HIP_TRY(hipEventElapsedTime(&tot, h->ev[0], h->ev[1]));
        h->last_step_ms = tot;            // the step launches back to back when a frame is one launch
        h->last_frame_ms = tot;
        h->timed_launches = T;
    }
    if (timing) {
        HIP_TRY(hipEventSynchronize(h->ev[2 * T + 1]));
        double sum = 0.0;
        for (int f = 0; f < T; ++f) {
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, h->ev[2 * f], h->ev[2 * f + 1]));
            sum += ms;
        }
        float tot = 0.f;
        HIP_TRY(hipEventElapsedTime(&tot, h->ev[2 * T], h->ev[2 * T + 1]));
        h->last_step_ms = sum;
        h->last_frame_ms = tot;
        h->timed_launches = T;
    }
    return PF_OK;
}

int pf_run_ref_dev(pf_handle *h, int32_t T, const float *Z, const uint32_t *uniforms, double *lml,
                   int32_t use_graph)
{
    if (!Z) return set_error(PF_ERR_ARG, "null Z");
    return run_frames(h, T, kModeRef, Z, nullptr, nullptr, uniforms, lml, use_graph);
}

int pf_run_candidates_dev(pf_handle *h, int32_t T, const int32_t *count, const float *cand,
                          const uint32_t *uniforms, double *lml, int32_t use_graph)
{
    if (!count || !cand) return set_error(PF_ERR_ARG, "null candidates");
    return run_frames(h, T, kModeCand, nullptr, count, cand, uniforms, lml, use_graph);
}

int pf_synchronize(pf_handle *h)
{
    int st = check_handle(h);
    if (st) return st;
    unsigned word = 0;
    st = read_status(h, &word);
    if (st) return st;
    return status_code(word);
}

int pf_get_stream(pf_handle *h, void **stream)
{
    if (!h || !stream) return set_error(PF_ERR_ARG, "null argument");
    *stream = (void *)h->stream;
    return PF_OK;
}

int pf_last_run_kernel_time(pf_handle *h, double *step_ms, double *frame_ms, int32_t *launches)
{
    if (!h) return set_error(PF_ERR_ARG, "null handle");
    if (step_ms) *step_ms = h->last_step_ms;
    if (frame_ms) *frame_ms = h->last_frame_ms;
    if (launches) *launches = h->timed_launches;
    return PF_OK;
}

int pf_last_run_graph_frames(pf_handle *h, int32_t *frames)
{
    if (!h || !frames) return set_error(PF_ERR_ARG, "null argument");
    *frames = h->graph_frames;
    return PF_OK;
}

int pf_set_state_layout(pf_handle *h, int32_t layout)
{
    int st = check_handle(h);
    if (st) return st;
    if (layout != PF_LAYOUT_AUTO && layout != PF_LAYOUT_ROWS && layout != PF_LAYOUT_RECORDS)
        return set_error(PF_ERR_ARG, "unknown state layout %d", layout);
    if (layout == PF_LAYOUT_RECORDS && h->model && !rec_layout(h->D))
        return set_error(PF_ERR_UNSUPPORTED, "per-particle records are compiled for d <= 12");
    h->layout_req = layout;
    if (!h->model) return PF_OK;
    const bool rec = want_records(h);
    if (rec == (h->lay.row == 1)) return PF_OK;
    drop_graphs(h);                   // they bake in the step kernel of the layout
    const int a = h->xA == h->xbuf[1] ? 1 : 0;
    if (!h->tracks) {
        set_views(h, rec, a);
        return PF_OK;
    }
    // the current state copied into the other buffer in the new layout, which
    // becomes the state the next step reads
    const float *xs = h->xA, *Ps = h->PA;
    const Lay Ls = h->lay;
    set_views(h, rec, 1 - a);
    st = dispatch_d<RelayoutL>(h->D, h, xs, Ps, Ls);
    if (st) return st;
    st = check_launch("k_relayout");
    if (st) return st;
    HIP_TRY(hipStreamSynchronize(h->stream));
    return PF_OK;
}

int pf_get_state_layout(pf_handle *h, int32_t *layout)
{
    int st = check_handle(h);
    if (st) return st;
    if (!layout) return set_error(PF_ERR_ARG, "null argument");
    *layout = (h->model ? h->lay.row == 1 : want_records(h)) ? PF_LAYOUT_RECORDS : PF_LAYOUT_ROWS;
    return PF_OK;
}

// -------------------------------------------------------------- KF_model
int pf_kf_predict(pf_handle *h, int32_t n, int32_t d, const double *F, const double *B,
                  const double *Q, double *x, double *P)
{
    int st = check_handle(h);
    if (st) return st;
    if (n <= 0 || d <= 0 || d > 32 || !F || !B || !Q || !x || !P)
        return set_error(PF_ERR_ARG, "bad KF predict arguments");
    const size_t mats = (size_t)(2 * d * d + d);
    const size_t states = (size_t)n * (d + d * d);
    st = ensure_gbuf(h, sizeof(double) * (mats + states));
    if (st) return st;
    double *dF = h->gbuf, *dQ = dF + d * d, *dB = dQ + d * d, *dx = dB + d, *dP = dx + (size_t)n * d;
    HIP_TRY(hipMemcpyAsync(dF, F, sizeof(double) * d * d, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(dQ, Q, sizeof(double) * d * d, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(dB, B, sizeof(double) * d, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(dx, x, sizeof(double) * (size_t)n * d, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(dP, P, sizeof(double) * (size_t)n * d * d, hipMemcpyHostToDevice, h->stream));
    const size_t lds = sizeof(double) * (size_t)(2 * d * d + d);
    hipLaunchKernelGGL(k_kf_predict, dim3(n), dim3(256), lds, h->stream, d, dF, dB, dQ, dx, dP);
    st = check_launch("k_kf_predict");
    if (st) return st;
    HIP_TRY(hipMemcpyAsync(x, dx, sizeof(double) * (size_t)n * d, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipMemcpyAsync(P, dP, sizeof(double) * (size_t)n * d * d, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return PF_OK;
}

int pf_kf_update(pf_handle *h, int32_t n, int32_t d, int32_t m, const double *H, const double *BH,
                 const double *R, const double *z, double *x, double *P)
{
    int st = check_handle(h);
    if (st) return st;
    if (n <= 0 || d <= 0 || d > 32 || m <= 0 || m > 16 || !H || !BH || !R || !z || !x || !P)
        return set_error(PF_ERR_ARG, "bad KF update arguments");
    const size_t mats = (size_t)(m * d + m + m * m);
    const size_t states = (size_t)n * (m + d + d * d);
    st = ensure_gbuf(h, sizeof(double) * (mats + states));
    if (st) return st;
    st = ensure_ibuf(h, sizeof(int) * 16);
    if (st) return st;
    double *dH = h->gbuf, *dBH = dH + m * d, *dR = dBH + m, *dz = dR + m * m, *dx = dz + (size_t)n * m;
    double *dP = dx + (size_t)n * d;
    HIP_TRY(hipMemcpyAsync(dH, H, sizeof(double) * m * d, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(dBH, BH, sizeof(double) * m, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(dR, R, sizeof(double) * m * m, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(dz, z, sizeof(double) * (size_t)n * m, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(dx, x, sizeof(double) * (size_t)n * d, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(dP, P, sizeof(double) * (size_t)n * d * d, hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, h->stream, h->ibuf, 0);
    const size_t lds = sizeof(double) * (size_t)(d * d + 2 * m * d + m * m + d + m);
    hipLaunchKernelGGL(k_kf_update, dim3(n), dim3(256), lds, h->stream, d, m, dH, dBH, dR, dz, dx, dP, h->ibuf);
    st = check_launch("k_kf_update");
    if (st) return st;
    int flag = 0;
    HIP_TRY(hipMemcpyAsync(&flag, h->ibuf, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipMemcpyAsync(x, dx, sizeof(double) * (size_t)n * d, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipMemcpyAsync(P, dP, sizeof(double) * (size_t)n * d * d, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (flag) return set_error(PF_ERR_NOT_SPD, "innovation covariance not positive definite");
    return PF_OK;
}

void pf_philox_u32(uint64_t seed, uint32_t stream, uint64_t first, int64_t count, uint32_t *out)
{
    philox_u32(seed, stream, first, count, out);
}

}  // extern "C"
