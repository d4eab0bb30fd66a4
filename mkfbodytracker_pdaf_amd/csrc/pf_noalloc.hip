// pf_noalloc.hip -- systematic resampling of log-weights that allocates no
// memory: the variant of pf_resample_log for callers that cannot call
// hipMalloc (code under HIP graph capture, tight serving loops).
//
// The resampler's scratch (block-local scans, block sums, block exponents,
// the frame record and the uniform) lives in the file-scope device array
// `scratch` below, whose device address is taken once per device with
// hipGetSymbolAddress and used for every kernel argument.  Every call on a
// device shares that device's array and leaves the last call's contents;
// each call first clears what must start at zero (with a kernel, so the call
// can be captured into a graph).  Calls that overlap on one device race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/pf_rbpf.h"
#include "pf_internal.hpp"
#include "pf_kernels.hpp"

using namespace mkfbt;

namespace {

// Largest weight vector: 2^19 log-weights (2048 resampler blocks; the
// resolver's LDS workspace is then 49 KB).
constexpr int kMaxL = PF_NOALLOC_MAX_WEIGHTS;
constexpr int kMaxNb = (kMaxL + kWeightsBlock - 1) / kWeightsBlock;
constexpr int kMaxOut = 1 << 26;
constexpr int kMaxDevices = 64;

// Word layout (8-byte words): cloc [kMaxL], bsum [kMaxNb], bexp [kMaxNb],
// bpow [kMaxNb int32 = kMaxNb / 2 words], frame record, counters.
constexpr size_t kOffCloc = 0;
constexpr size_t kOffBsum = kOffCloc + kMaxL;
constexpr size_t kOffBexp = kOffBsum + kMaxNb;
constexpr size_t kOffBpow = kOffBexp + kMaxNb;
constexpr size_t kOffFs = kOffBpow + (kMaxNb + 1) / 2;
constexpr size_t kFsWords = (sizeof(FrameScratch) + 7) / 8;
constexpr size_t kOffMisc = kOffFs + kFsWords;   // frame counter (int32 x 2), uniforms (uint32 x 2)
constexpr size_t kWords = kOffMisc + 2;
static_assert(sizeof(FrameScratch) % 8 == 0, "frame record is a whole number of words");

}  // namespace

// 4.1 MiB per device, held until the process ends.
static __device__ __align__(256) unsigned long long scratch[kWords];

namespace {

std::mutex g_mu;
unsigned long long *g_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device (taken once).
int scratch_ptr(int dev, unsigned long long **out)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(PF_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        g_ptr[dev] = static_cast<unsigned long long *>(p);
    }
    *out = g_ptr[dev];
    return PF_OK;
}

// Clears the frame record, sets the frame counter to 0 and both uniform
// slots to u (k_weights reads uni[2 t + 1], k_ancestors the recorded one).
__global__ void k_noalloc_begin(FrameScratch *fs, int *frame, uint32_t *uni, uint32_t u)
{
    if (threadIdx.x == 0) {
        fs->maxkey[0] = 0ull;
        fs->maxkey[1] = 0ull;
        fs->spare = 0u;
        fs->status = 0u;
        fs->U_pending = 0u;
        fs->pad = 0u;
        fs->pow_max[0] = kPowEmpty;
        fs->pow_max[1] = kPowEmpty;
        fs->pow_max[2] = kPowEmpty;
        fs->pad2 = 0;
        frame[0] = 0;
        frame[1] = 0;
        uni[0] = u;
        uni[1] = u;
    }
}

__global__ void k_noalloc_status(const FrameScratch *fs, int32_t *status)
{
    if (threadIdx.x == 0) *status = (fs->status & 1u) ? PF_WARN_DEGENERATE : PF_OK;
}

}  // namespace

extern "C" int pf_resample_log_noalloc(const double *lw, int32_t L, int32_t N, uint32_t u, int32_t *anc,
                                       int32_t *status, void *stream)
{
    if (!lw || !anc || L <= 0 || N <= 0)
        return set_error(PF_ERR_ARG, "bad resample arguments (L=%d, N=%d, null pointer)", L, N);
    if (L > kMaxL) return set_error(PF_ERR_ARG, "L=%d exceeds the scratch maximum of %d weights", L, kMaxL);
    if (N > kMaxOut) return set_error(PF_ERR_ARG, "N=%d exceeds %d outputs", N, kMaxOut);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
        return set_error(PF_ERR_HIP, "no current HIP device");
    unsigned long long *base = nullptr;
    int st = scratch_ptr(dev, &base);
    if (st) return st;
    hipStream_t s = static_cast<hipStream_t>(stream);
    FrameScratch *fs = reinterpret_cast<FrameScratch *>(base + kOffFs);
    int *frame = reinterpret_cast<int *>(base + kOffMisc);
    uint32_t *uni = reinterpret_cast<uint32_t *>(base + kOffMisc + 1);
    hipLaunchKernelGGL(k_noalloc_begin, dim3(1), dim3(64), 0, s, fs, frame, uni, u);

    const int nb = (L + kWeightsBlock - 1) / kWeightsBlock;
    WeightArgs wa{};
    wa.w64 = lw;
    wa.cloc = base + kOffCloc;
    wa.bsum = base + kOffBsum;
    wa.bexp = reinterpret_cast<double *>(base + kOffBexp);
    wa.bpow = reinterpret_cast<int *>(base + kOffBpow);
    wa.fs = fs;
    wa.frame = frame;
    wa.frame_next = nullptr;
    wa.uni = uni;
    wa.L = L;
    wa.parity = 0;
    wa.ustride = 2;
    wa.uoff = 1;
    hipLaunchKernelGGL((k_weights<kLogF64>), dim3(nb), dim3(kWeightsBlock), 0, s, wa);

    AncArgs an{};
    an.pend.cloc = wa.cloc;
    an.pend.bsum = wa.bsum;
    an.pend.bexp = wa.bexp;
    an.pend.bpow = wa.bpow;
    an.pend.L = L;
    an.pend.nb = nb;
    an.fs = fs;
    an.anc = anc;
    an.N = N;
    an.parity = 0;
    an.kind_linear = 0;
    const size_t lds = 8 * (size_t)resolve_lds_words(nb);
    hipLaunchKernelGGL(k_ancestors, dim3((N + 255) / 256), dim3(256), lds, s, an);
    if (status) hipLaunchKernelGGL(k_noalloc_status, dim3(1), dim3(64), 0, s, (const FrameScratch *)fs, status);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(PF_ERR_HIP, "resample launch failed: %s", hipGetErrorString(e));
    return PF_OK;
}
