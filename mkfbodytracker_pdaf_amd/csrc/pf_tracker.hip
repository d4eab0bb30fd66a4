// pf_tracker.hip -- the pose tracker's per-frame pipeline (include/pf_tracker.h)
// on gfx950: PFTracker::getMeasurementProposal (src/pfPose.cpp:210-329), the
// two arm updates it ends with, and the pose output of callback
// (pfPose.cpp:347-348) with get3Dpose (pfPose.cpp:93-127).
//
// Kernels, per frame:
//   k_blur        separable Gaussian of the likelihood image, LDS-tiled
//   k_propose     2 proposals per lane for both arms: Philox words ->
//                 position, association weight against the blurred image and
//                 both arms' proposal densities, and the resampler's
//                 fixed-point weights and block scans (the weights' bound,
//                 not their maximum, sets the fixed-point scale, so no second
//                 pass over the weights is needed)
//   k_prop_select the systematic resampling of the proposals into N indices
//                 and the 6 x N measurement; the chosen proposals are
//                 recomputed from their Philox words rather than stored
//                 (nothing but the weights reaches HBM)
//   (filter)      update, weights, ancestors fused with the estimate's block
//                 sums, the estimate (own stream)
//   k_pose        e = pca_proj^T xbar + pca_mean and get3Dpose for both arms
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/pf_tracker.h"
#include "pf_common.hpp"
#include "pf_engine.hpp"
#include "pf_internal.hpp"
#include "pf_kernels.hpp"
#include "pf_rng.hpp"

using namespace mkfbt;

#define TR_TRY(expr)                                                                           \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return set_error(PF_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));       \
    } while (0)

namespace {

constexpr uint32_t kStreamProposal = 16;  // + arm
constexpr uint32_t kStreamFrame = 24;     // 8 words per frame
constexpr int kMaxRadius = 15;
constexpr int kBlurW = 32, kBlurH = 16, kBlurThreads = 128;
constexpr int kProposeBlock = 256;
constexpr int kMaxArmD = 24;              // k_propose stages 3 D + 2 projection operands per arm
constexpr int kTrGraphFrames = 16;        // frames per captured graph of a device-resident run (even)
constexpr int kPropShift = 10;            // the proposal resampler's input blocks: 1024 proposals
constexpr int kMaxProposals = 1 << 20;    // resampler scan limit

// Frame record written by k_propose: what k_prop_select and the inspection
// call need to recompute a proposal.
struct FrameRec {
    double y[4];     // projected hand mean of arm 0 (u, v), arm 1 (u, v)
    int roi[4];
    long long g;     // global frame index
    int uniform;     // first-frame (uniform) proposals
    int pad;
};

struct ArmDev {
    const double *xbar;       // estimate, model basis (D)
    const double *hp;         // npose x D
    const double *pm;         // npose
    double *w;                // n_c association weights (kept for inspection)
    unsigned long long *cloc; // n_c block-local inclusive scans of the fixed-point weights
    unsigned long long *bsum; // their block sums (kWeightsBlock inputs per block)
    FrameScratch *fs;         // the arm handle's resampler scratch (pending uniform, status)
    int D;
};

// ------------------------------------------------------------------ blur
struct BlurArgs {
    const uint8_t *src;
    uint8_t *dst;
    const float *taps;        // 2 radius + 1
    const int *frame;         // source frame = *frame (nullptr: 0)
    long long frame_stride;
    int rows, cols;
};

__device__ __forceinline__ int reflect101(int i, int n)
{
    i = i < 0 ? -i : i;
    i = i >= n ? 2 * n - 2 - i : i;
    return min(max(i, 0), n - 1);
}

// GaussianBlur(ksize, sigma, BORDER_REFLECT_101) (pfPose.cpp:213): row pass
// then column pass, each an FP32 fma chain over the taps in index order,
// round to nearest even and saturate to uint8.  One 32 x 16 output tile per
// block (1200 blocks for 640 x 480), source tile and row-pass results in LDS.
template <int R>
__global__ void __launch_bounds__(kBlurThreads) k_blur(BlurArgs a)
{
    constexpr int IW = kBlurW + 2 * R, IH = kBlurH + 2 * R, NT = 2 * R + 1;
    __shared__ uint8_t s_in[IH * IW];
    __shared__ float s_row[IH * kBlurW];
    __shared__ float s_tap[NT];
    const int x0 = blockIdx.x * kBlurW, y0 = blockIdx.y * kBlurH;
    const long long f = a.frame ? (long long)*a.frame : 0;
    const uint8_t *src = a.src + f * a.frame_stride;
    if ((int)threadIdx.x < NT) s_tap[threadIdx.x] = a.taps[threadIdx.x];
    // all of a thread's source loads are issued before its first LDS store,
    // so the tile costs one memory latency rather than one per element
    constexpr int kPer = (IH * IW + kBlurThreads - 1) / kBlurThreads;
    uint8_t v[kPer];
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
        const int e = (int)threadIdx.x + r * kBlurThreads;
        const int iy = e / IW, ix = e - iy * IW;
        const int gy = reflect101(y0 + iy - R, a.rows), gx = reflect101(x0 + ix - R, a.cols);
        v[r] = e < IH * IW ? src[(size_t)gy * a.cols + gx] : (uint8_t)0;
    }
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
        const int e = (int)threadIdx.x + r * kBlurThreads;
        if (e < IH * IW) s_in[e] = v[r];
    }
    __syncthreads();
    // both passes unrolled over a thread's outputs, so their independent fma
    // chains interleave
    constexpr int kRow = (IH * kBlurW + kBlurThreads - 1) / kBlurThreads;
#pragma unroll
    for (int r = 0; r < kRow; ++r) {
        const int e = (int)threadIdx.x + r * kBlurThreads;
        if (e < IH * kBlurW) {
            const int iy = e / kBlurW, ox = e - iy * kBlurW;
            const uint8_t *p = s_in + iy * IW + ox;
            float acc = 0.f;
#pragma unroll
            for (int k = 0; k < NT; ++k) acc = fmaf(s_tap[k], (float)p[k], acc);
            s_row[e] = acc;
        }
    }
    __syncthreads();
    constexpr int kCol = (kBlurH * kBlurW + kBlurThreads - 1) / kBlurThreads;
#pragma unroll
    for (int r = 0; r < kCol; ++r) {
        const int e = (int)threadIdx.x + r * kBlurThreads;
        const int oy = e / kBlurW, ox = e - oy * kBlurW;
        const int gy = y0 + oy, gx = x0 + ox;
        if (e < kBlurH * kBlurW && gy < a.rows && gx < a.cols) {
            const float *p = s_row + oy * kBlurW + ox;
            float acc = 0.f;
#pragma unroll
            for (int k = 0; k < NT; ++k) acc = fmaf(s_tap[k], p[k * kBlurW], acc);
            a.dst[(size_t)gy * a.cols + gx] = (uint8_t)fminf(fmaxf(rintf(acc), 0.f), 255.f);
        }
    }
}

// --------------------------------------------------------------- proposals
struct PropParams {
    uint64_t seed;
    long long g0;             // global index of frame t = 0 of this run
    long long uniform_frame;  // global index of the uniform-proposal frame (-1: none)
    const int *roi;           // frames x 4
    const int *frame;         // t
    const uint32_t *u_prop;   // frames x 2: the proposal resampling uniform of each arm
    int n_c, rows, cols;
    double spread, pa, zfloor;
};

// y = Hp[r] . xbar + M[r], products accumulated in index order.
__device__ inline double project_row(const double *hp, const double *pm, const double *xbar, int D, int r)
{
#pragma clang fp contract(off)
    double s = 0.0;
    for (int k = 0; k < D; ++k) s = s + hp[r * D + k] * xbar[k];
    return s + pm[r];
}

// Philox block of proposals 2b and 2b + 1 of arm `arm` in frame g.
__device__ inline void proposal_block(const PropParams &p, int arm, long long g, long long b, uint32_t wd[4])
{
    philox_stream_block(p.seed, kStreamProposal + arm, (uint64_t)g * (uint64_t)(p.n_c / 2) + (uint64_t)b, wd);
}

// A proposal from its two words: getSamples (pf2DRao.cpp:85-103) or the
// uniform first-frame box (pfPose.cpp:222-235).
__device__ inline void proposal_at(const PropParams &p, bool uniform, const int *roi, double y0, double y1,
                                   uint32_t w0, uint32_t w1, double &px, double &py)
{
    if (uniform) {
        const long long xo = roi[0], yo = roi[1], wi = roi[2], he = roi[3];
        const long long x_lo = xo - 4 * wi, x_hi = xo + 5 * wi, y_lo = yo + he, y_hi = yo + 7 * he;
        const double xmin = (double)(x_lo > 0 ? x_lo : 0), xmax = (double)(x_hi < p.cols ? x_hi : p.cols);
        const double ymin = (double)(y_lo < p.rows ? y_lo : p.rows), ymax = (double)(y_hi < p.rows ? y_hi : p.rows);
        px = fma(xmax - xmin, unit_interval(w0), xmin);
        py = fma(ymax - ymin, unit_interval(w1), ymin);
    } else {
        const double sd = p.spread * (double)roi[2];
        double z0, z1;
        normal_pair(w0, w1, z0, z1);
        px = fma(sd, z0, y0);
        py = fma(sd, z1, y1);
    }
}

// Proposal c of arm `arm` in frame g (words 2c, 2c + 1 of the arm's stream).
__device__ inline void proposal(const PropParams &p, int arm, long long g, bool uniform, const int *roi, double y0,
                                double y1, long long c, double &px, double &py)
{
    uint32_t wd[4];
    proposal_block(p, arm, g, c >> 1, wd);
    proposal_at(p, uniform, roi, y0, y1, wd[2 * (c & 1)], wd[2 * (c & 1) + 1], px, py);
}

// N((px, py); y, var I) (pf2DRao.cpp:69-83,105-122): getSampleProb takes
// spread x width as the variance, while getSamples draws with it as the
// standard deviation (pf2DRao.cpp:90 vs :114).
__device__ inline double density(double px, double py, double y0, double y1, double nhiv, double inv2piv)
{
#pragma clang fp contract(off)
    const double d0 = px - y0, d1 = py - y1;
    const double q = fma(d0, d0, d1 * d1);
    return inv2piv * det_exp(q * nhiv);
}

// Hand rows 0, 1 of both arms' projections y = Hp xbar + M, summed as
// project_row does; the operands (Hp rows 0, 1, xbar, M[0..1] of each arm,
// 3 D + 2 doubles per arm) are first staged in LDS by one load per thread so
// that the block pays one memory latency.  Every thread of the block calls it.
__device__ inline void hand_projections(const ArmDev &a0, const ArmDev &a1, double *stage, double *y)
{
#pragma clang fp contract(off)
    const int n0 = 3 * a0.D + 2, n1 = 3 * a1.D + 2;
    for (int e = threadIdx.x; e < n0 + n1; e += blockDim.x) {
        const ArmDev &A = e < n0 ? a0 : a1;
        const int i = e < n0 ? e : e - n0;
        const int D = A.D;
        stage[e] = i < 2 * D ? A.hp[i] : (i < 3 * D ? A.xbar[i - 2 * D] : A.pm[i - 3 * D]);
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        const int arm = threadIdx.x >> 1, r = threadIdx.x & 1;
        const int D = arm ? a1.D : a0.D;
        const double *st = stage + (arm ? n0 : 0);
        double v = 0.0;
        for (int k = 0; k < D; ++k) v = v + st[r * D + k] * st[2 * D + k];
        y[threadIdx.x] = v + st[3 * D + r];
    }
    __syncthreads();
}

// Proposals of both arms (grid.y = arm).  A chunk of 512 proposal pairs
// (1024 proposals) is one input block of the proposal resampler; a block
// walks the chunks c = blockIdx.x, + gridDim.x, ... (one resident round of
// blocks, so the frame prologue is paid once per block), each lane taking
// pairs 512 c + tid and 512 c + 256 + tid (one Philox block per pair).  It
// writes the weights and the block-local inclusive scans and block sums of
// q = fixed32(w z): with z = z_floor (1 - 2 Pa) > 0 every weight is at most
// 1 / z, so w z is in [0, 1] and the maximum is not needed
// (resample(weights, N), pfPose.cpp:300-301).
__global__ void __launch_bounds__(kProposeBlock) k_propose(PropParams p, ArmDev a0, ArmDev a1, const uint8_t *img,
                                                          FrameRec *rec)
{
    static_assert(kProposeBlock == 256 && (1 << kPropShift) == 4 * kProposeBlock,
                  "a chunk is two pairs per lane, one resampler block");
    __shared__ double s_stage[2 * (3 * kMaxArmD + 2)];
    __shared__ double s_y[4];
    __shared__ unsigned long long s_wave[2][kProposeBlock / 64];
    const int t = *p.frame;
    const long long g = p.g0 + t;
    const bool uniform = (g == p.uniform_frame);
    const int *roi = p.roi + 4 * (size_t)t;
    const int arm = blockIdx.y;
    const ArmDev &A = arm ? a1 : a0;
    hand_projections(a0, a1, s_stage, s_y);
    if (blockIdx.x == 0 && threadIdx.x == 0) A.fs->U_pending = p.u_prop[2 * (size_t)t + arm];
    if (blockIdx.x == 0 && arm == 0 && threadIdx.x < 4) {
        rec->y[threadIdx.x] = s_y[threadIdx.x];
        rec->roi[threadIdx.x] = roi[threadIdx.x];
        if (threadIdx.x == 0) {
            rec->g = g;
            rec->uniform = uniform ? 1 : 0;
        }
    }
    const double yo0 = s_y[2 * arm], yo1 = s_y[2 * arm + 1];
    const double yt0 = s_y[2 * (1 - arm)], yt1 = s_y[2 * (1 - arm) + 1];
    const double var = p.spread * (double)roi[2];
    const double inv2piv = 1.0 / (6.283185307179586 * var), nhiv = -0.5 / var;
    const long long pairs = p.n_c / 2;
    const int nchunks = (p.n_c + (1 << kPropShift) - 1) >> kPropShift;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int c = blockIdx.x; c < nchunks; c += gridDim.x) {
        unsigned long long q[4] = {0ull, 0ull, 0ull, 0ull};
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const long long k = ((long long)c << (kPropShift - 1)) + h * kProposeBlock + threadIdx.x;
            if (k < pairs) {
                uint32_t wd[4];
                proposal_block(p, arm, g, k, wd);     // one Philox block serves both proposals
                double wv[2] = {0.0, 0.0};
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    double px, py;
                    proposal_at(p, uniform, roi, yo0, yo1, wd[2 * i], wd[2 * i + 1], px, py);
                    // pfPose.cpp:251-266 (arm 1) and :274-289 (arm 2)
                    if (py > 0.0 && py < (double)p.rows && px > 0.0 && px < (double)p.cols) {
                        const unsigned v = img[(size_t)(int)py * p.cols + (int)px];
                        if (v != 0u) {     // (v / 255) / Z with one division
                            const double po = density(px, py, yo0, yo1, nhiv, inv2piv);
                            const double pt = density(px, py, yt0, yt1, nhiv, inv2piv);
                            wv[i] = (double)v / (255.0 * fma(p.pa, po, fma(p.pa, pt, p.zfloor)));
                        }
                    }
                    q[2 * h + i] = (isfinite(wv[i]) && wv[i] > 0.0) ? fixed32(wv[i] * p.zfloor) : 0ull;
                }
                *reinterpret_cast<double2 *>(A.w + 2 * k) = make_double2(wv[0], wv[1]);
            }
        }
        // block-local inclusive scan over the chunk: pairs of the first half
        // (lane order), then those of the second half
        unsigned long long s1 = wave_incl_scan(q[0] + q[1]), s2 = wave_incl_scan(q[2] + q[3]);
        if (lane == 63) {
            s_wave[0][wid] = s1;
            s_wave[1][wid] = s2;
        }
        __syncthreads();
        unsigned long long first = 0;
#pragma unroll
        for (int w = 0; w < kProposeBlock / 64; ++w) {
            first += s_wave[0][w];
            if (w < wid) {
                s1 += s_wave[0][w];
                s2 += s_wave[1][w];
            }
        }
        s2 += first;
        const long long k1 = ((long long)c << (kPropShift - 1)) + threadIdx.x, k2 = k1 + kProposeBlock;
        if (k1 < pairs) {
            A.cloc[2 * k1] = s1 - q[1];
            A.cloc[2 * k1 + 1] = s1;
        }
        if (k2 < pairs) {
            A.cloc[2 * k2] = s2 - q[3];
            A.cloc[2 * k2 + 1] = s2;
        }
        if (threadIdx.x == kProposeBlock - 1) A.bsum[c] = s2;
        __syncthreads();                     // s_wave is reused by the next chunk
    }
}

// One arm's proposal selection: its resampler scans, frame scratch, and the
// chosen indices and measurement rows it writes.
struct SelectArgs {
    Pending pend;
    FrameScratch *fs;
    int *bins;
    float *Z;
};

__device__ __forceinline__ void prop_select_body(const PropParams &p, int arm, const SelectArgs &sa,
                                                 const FrameRec *rec, int N)
{
    extern __shared__ unsigned long long wsp[];
    const int j0 = blockIdx.x * blockDim.x, j = j0 + threadIdx.x;
    // the frame record is loaded before the resolution so that its latency
    // overlaps the block-sum loads
    const int roi[4] = {rec->roi[0], rec->roi[1], rec->roi[2], rec->roi[3]};
    const long long g = rec->g;
    const bool uniform = rec->uniform != 0;
    const double y0 = rec->y[2 * arm], y1 = rec->y[2 * arm + 1];
    double es = 0.0;
    int pw;
    const unsigned long long total = resolve_prefix(sa.pend, wsp, false, es, pw);
    const int c = resolve_search(sa.pend, N, sa.fs->U_pending, total, j0, wsp);
    if (blockIdx.x == 0 && threadIdx.x == 0 && total == 0ull) sa.fs->status |= 1u;
    if (j >= N) return;
    sa.bins[j] = c;
    double px, py;
    proposal(p, arm, g, uniform, roi, y0, y1, c, px, py);
    float *Z = sa.Z;
    const double fx = (double)roi[0] + (double)roi[2] / 2.0;
    Z[j] = (float)fx;
    Z[(size_t)N + j] = (float)((double)roi[1] + 0.5 * (double)roi[3]);
    Z[2 * (size_t)N + j] = (float)px;
    Z[3 * (size_t)N + j] = (float)py;
    Z[4 * (size_t)N + j] = (float)fx;
    Z[5 * (size_t)N + j] = (float)((double)roi[1] + 1.65 * (double)roi[3]);
}

__global__ void __launch_bounds__(256) k_prop_select(PropParams p, int arm, SelectArgs sa, const FrameRec *rec, int N)
{
    prop_select_body(p, arm, sa, rec, N);
}

// Both arms in one launch (grid.y = arm).
__global__ void __launch_bounds__(256) k_prop_select_pair(PropParams p, SelectArgs s0, SelectArgs s1,
                                                          const FrameRec *rec, int N)
{
    prop_select_body(p, (int)blockIdx.y, blockIdx.y ? s1 : s0, rec, N);
}

// All proposals of the recorded frame (inspection only): out 2 x 2 x n_c.
__global__ void __launch_bounds__(256) k_dump_proposals(PropParams p, const FrameRec *rec, double *out)
{
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int arm = blockIdx.y;
    if (c >= p.n_c) return;
    double px, py;
    proposal(p, arm, rec->g, rec->uniform != 0, rec->roi, rec->y[2 * arm], rec->y[2 * arm + 1], c, px, py);
    out[(size_t)arm * 2 * p.n_c + c] = px;
    out[(size_t)arm * 2 * p.n_c + p.n_c + c] = py;
}

// -------------------------------------------------------------------- pose
struct PoseArgs {
    ArmDev arm[2];
    int *frame;               // t (incremented at the end)
    double *pose;             // frames x 2 x npose (nullable)
    double *pos3d;            // frames x 2 x 15 (nullable)
    double *tf;               // frames x 24 (nullable)
    double *pose2d;           // frames x 16 (nullable)
    int npose;
    double fx, fy, cx, cy;
};

// get3Dpose (pfPose.cpp:93-127): joint k of e at pixel (u, v) and depth d,
// X = (K R)^-1 ((u d, v d, d) - K t) = R^T ((d (u - cx)/fx, d (v - cy)/fy, d) - t)
// with R = rpy(e16, e17, e15) = R3(yaw) R2(pitch) R1(roll) (pfPose.cpp:84-91)
// and t = e18..20.  Row r of R^T q is column r of R dotted with q.
__device__ double pose3d_entry(const double *e, const double *tg, int r, int k, double fx, double fy, double cx,
                               double cy)
{
    const double cr = tg[0], sr = tg[1], cp = tg[2], sp = tg[3], cyw = tg[4], syw = tg[5];
    double c0, c1, c2;     // column r of R
    if (r == 0) {
        c0 = cyw * cp; c1 = syw * cp; c2 = -sp;
    } else if (r == 1) {
        c0 = cyw * sp * sr - syw * cr; c1 = syw * sp * sr + cyw * cr; c2 = cp * sr;
    } else {
        c0 = cyw * sp * cr + syw * sr; c1 = syw * sp * cr - cyw * sr; c2 = cp * cr;
    }
    const double u = e[3 * k], v = e[3 * k + 1], d = e[3 * k + 2];
    const double q0 = d * (u - cx) / fx - e[18], q1 = d * (v - cy) / fy - e[19], q2 = d - e[20];
    return c0 * q0 + c1 * q1 + c2 * q2;
}

// Entry i of publishTFtree's joints relative to its Head frame (pfPose.cpp:
// 129-171 as extractTFdata.cpp:87-103 records them): joint i / 3 of Head,
// Left_Hand, Right_Hand, Left_Elbow, Right_Elbow, Right_Shoulder,
// Left_Shoulder, Neck (arm 0 = Right_*, arm 1 = Left_*); the published
// offsets map a camera-frame vector v to (v.x, v.z, -v.y), every rotation is
// the identity, so a joint relative to Head is that map of (joint - head),
// head and neck being the means of the two arms' head and neck columns.
__device__ double tf_entry(const double (*p)[15], int i)
{
    const int jt = i / 3, c = i % 3;
    const int ax = c == 0 ? 0 : (c == 1 ? 2 : 1);
    const double head = (p[0][ax * 5 + 3] + p[1][ax * 5 + 3]) / 2.0;
    double v;
    switch (jt) {
    case 0: return 0.0;
    case 1: v = p[1][ax * 5 + 0]; break;
    case 2: v = p[0][ax * 5 + 0]; break;
    case 3: v = p[1][ax * 5 + 1]; break;
    case 4: v = p[0][ax * 5 + 1]; break;
    case 5: v = p[0][ax * 5 + 2]; break;
    case 6: v = p[1][ax * 5 + 2]; break;
    default: v = (p[0][ax * 5 + 4] + p[1][ax * 5 + 4]) / 2.0; break;
    }
    return c == 2 ? head - v : v - head;
}

// Entry i of publish2Dpos's points (pfPose.cpp:173-208): hand 1, hand 2,
// head, neck (means of the arms), elbow 1, elbow 2, shoulder 1, shoulder 2.
__device__ double pose2d_entry(const double (*e)[64], int i)
{
    const int pt = i >> 1, c = i & 1;
    switch (pt) {
    case 0: return e[0][c];
    case 1: return e[1][c];
    case 2: return 0.5 * (e[0][9 + c] + e[1][9 + c]);
    case 3: return 0.5 * (e[0][12 + c] + e[1][12 + c]);
    case 4: return e[0][3 + c];
    case 5: return e[1][3 + c];
    case 6: return e[0][6 + c];
    default: return e[1][6 + c];
    }
}

// Pose output of both arms: e (one lane per row), the six angle functions
// of each arm (one lane each), the 3 x 5 joint positions (one lane each),
// then the TF-tree and 2-D point lists.
// x0, x1: the arms' estimates (model basis).  Every thread of the block
// calls it.
__device__ inline void pose_body(const PoseArgs &a, const double *x0, const double *x1)
{
    __shared__ double s_e[2][64];
    __shared__ double s_tg[2][6];
    __shared__ double s_p[2][15];
    const int t = *a.frame;
    for (int i = threadIdx.x; i < 2 * a.npose; i += blockDim.x) {
        const int arm = i / a.npose, r = i - arm * a.npose;
        const ArmDev &A = arm ? a.arm[1] : a.arm[0];
        s_e[arm][r] = project_row(A.hp, A.pm, arm ? x1 : x0, A.D, r);
    }
    __syncthreads();
    if (threadIdx.x < 12) {
        const int arm = threadIdx.x / 6, f = threadIdx.x % 6;
        const double ang = s_e[arm][f < 2 ? 16 : (f < 4 ? 17 : 15)];   // roll, pitch, yaw
        s_tg[arm][f] = (f & 1) ? sin(ang) : cos(ang);
    }
    __syncthreads();
    if (threadIdx.x < 30) {
        const int arm = threadIdx.x / 15, i = threadIdx.x % 15, r = i / 5, k = i % 5;
        const double v = pose3d_entry(s_e[arm], s_tg[arm], r, k, a.fx, a.fy, a.cx, a.cy);
        s_p[arm][i] = v;
        if (a.pos3d) a.pos3d[((size_t)t * 2 + arm) * 15 + i] = v;
    }
    __syncthreads();
    if (a.tf && threadIdx.x < 24) a.tf[(size_t)t * 24 + threadIdx.x] = tf_entry(s_p, threadIdx.x);
    if (a.pose2d && threadIdx.x < 16) a.pose2d[(size_t)t * 16 + threadIdx.x] = pose2d_entry(s_e, threadIdx.x);
    if (a.pose)
        for (int i = threadIdx.x; i < 2 * a.npose; i += blockDim.x) {
            const int arm = i / a.npose, r = i - arm * a.npose;
            a.pose[((size_t)t * 2 + arm) * a.npose + r] = s_e[arm][r];
        }
    __syncthreads();
    if (threadIdx.x == 0) *a.frame = t + 1;
}

__global__ void __launch_bounds__(64) k_pose(PoseArgs a)
{
    pose_body(a, a.arm[0].xbar, a.arm[1].xbar);
}

// The arms' estimates from their block sums, then the pose: threads
// 256 h .. 256 h + 255 add arm h's sums exactly as k_estimate_final does
// (thread t takes blocks t, t + 256, ... in order, then the fixed tree of
// block_sum256, then Ti / N), write the estimate and keep it in LDS for the
// projection.  For arms of the same state dimension D (one launch instead of
// the estimate launch and the pose launch).
template <int D>
__global__ void __launch_bounds__(512) k_pose_est(PoseArgs a, EstFinalArgs e0, EstFinalArgs e1)
{
    __shared__ double s_v[2][D][16][17];
    __shared__ double s_run[2][D][16];
    __shared__ double s_sum[2][D];
    __shared__ double s_x[2][D];
    const int h = threadIdx.x >> 8, t = threadIdx.x & 255;
    const EstFinalArgs &e = h ? e1 : e0;
    double v[D];
#pragma unroll
    for (int i = 0; i < D; ++i) v[i] = 0.0;
    for (int b = t; b < e.nb; b += 256) {
#pragma unroll
        for (int i = 0; i < D; ++i) v[i] += e.partial[(size_t)b * D + i];
    }
#pragma unroll
    for (int i = 0; i < D; ++i) s_v[h][i][t >> 4][t & 15] = v[i];
    __syncthreads();
    for (int q = t; q < D * 16; q += 256) {
        const int i = q >> 4, k = q & 15;
        double r = 0.0;
#pragma unroll
        for (int u = 0; u < 16; ++u) r += s_v[h][i][k][u];
        s_run[h][i][k] = r;
    }
    __syncthreads();
    if (t < D) {
        double r = 0.0;
#pragma unroll
        for (int k = 0; k < 16; ++k) r += s_run[h][t][k];
        s_sum[h][t] = r;
    }
    __syncthreads();
    if (t < D) {
        double r = 0.0;
        for (int k = 0; k < D; ++k) r = fma(e.Ti[t * D + k], s_sum[h][k] / (double)e.N, r);
        e.out[t] = r;
        s_x[h][t] = r;
    }
    __syncthreads();
    pose_body(a, s_x[0], s_x[1]);
}

__global__ void k_tr_set(int *p, int v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

template <typename T>
hipError_t talloc(T **p, size_t n)
{
    return hipMalloc((void **)p, sizeof(T) * std::max<size_t>(n, 1));
}

}  // namespace

struct pf_tracker {
    pf_tracker_config cfg{};
    pf_handle *arm[2] = {nullptr, nullptr};
    hipStream_t s[2] = {nullptr, nullptr};
    int N = 0, D[2] = {0, 0}, npose = 0, n_c = 0, device = 0, radius = 0;
    int prop_blocks = 1;              // k_propose blocks per arm: one resident round over the device
    double *hp[2] = {nullptr, nullptr}, *pm[2] = {nullptr, nullptr};
    float *taps = nullptr;
    uint8_t *img_stage = nullptr, *blur = nullptr;
    int *roi_stage = nullptr;
    double *w[2] = {nullptr, nullptr};
    int *bins[2] = {nullptr, nullptr};
    float *Z[2] = {nullptr, nullptr};
    FrameRec *rec = nullptr;
    int *fctr = nullptr, *fctr_aux = nullptr;
    uint32_t *uni = nullptr;          // [arm 0: cap x 2][arm 1: cap x 2][proposals: cap x 2]
    int uni_cap = 0;
    std::vector<uint32_t> h_uni;
    pf_tracker_outputs stage_out{};     // device outputs of one host-driven frame
    hipEvent_t ev_fork[2] = {nullptr, nullptr}, ev_join[2] = {nullptr, nullptr}, ev_pre = nullptr;
    hipEvent_t ev_t0 = nullptr, ev_t1 = nullptr;
    long long frames_done = 0, uniform_frame = 0;
    bool ran = false;
};

namespace {

int tr_check(pf_tracker *tr)
{
    if (!tr) return set_error(PF_ERR_ARG, "null tracker");
    if (hipSetDevice(tr->device) != hipSuccess) return set_error(PF_ERR_HIP, "hipSetDevice failed");
    return PF_OK;
}

// The arm handle's resampler scratch as the pending resampling of its n_c
// proposal weights, in k_propose's input blocks of 1024.
Pending proposal_pending(pf_tracker *tr, int a)
{
    Pending pend = eng::pending(tr->arm[a], tr->n_c);
    pend.bshift = kPropShift;
    pend.nb = (tr->n_c + (1 << kPropShift) - 1) >> kPropShift;
    return pend;
}

ArmDev arm_dev(pf_tracker *tr, int a)
{
    ArmDev d{};
    d.xbar = eng::estimate_ptr(tr->arm[a]);
    d.hp = tr->hp[a];
    d.pm = tr->pm[a];
    d.w = tr->w[a];
    const Pending pend = proposal_pending(tr, a);
    d.cloc = const_cast<unsigned long long *>(pend.cloc);
    d.bsum = const_cast<unsigned long long *>(pend.bsum);
    d.fs = eng::scratch(tr->arm[a]);
    d.D = tr->D[a];
    return d;
}

PropParams prop_params(pf_tracker *tr, const int *roi, const int *frame)
{
    PropParams p{};
    p.seed = tr->cfg.seed;
    p.g0 = tr->frames_done;
    p.uniform_frame = tr->uniform_frame;
    p.roi = roi;
    p.frame = frame;
    p.u_prop = tr->uni + 4 * (size_t)tr->uni_cap;
    p.n_c = tr->n_c;
    p.rows = tr->cfg.rows;
    p.cols = tr->cfg.cols;
    p.spread = tr->cfg.spread;
    p.pa = tr->cfg.p_assoc;
    p.zfloor = tr->cfg.z_floor * (1.0 - 2.0 * tr->cfg.p_assoc);
    return p;
}

int launch_check(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(PF_ERR_HIP, "%s launch failed: %s", what, hipGetErrorString(e));
    return PF_OK;
}

int launch_blur(pf_tracker *tr, const uint8_t *src, const int *frame, uint8_t *dst)
{
    BlurArgs b{};
    b.src = src;
    b.dst = dst;
    b.taps = tr->taps;
    b.frame = frame;
    b.frame_stride = (long long)tr->cfg.rows * tr->cfg.cols;
    b.rows = tr->cfg.rows;
    b.cols = tr->cfg.cols;
    const dim3 grid((tr->cfg.cols + kBlurW - 1) / kBlurW, (tr->cfg.rows + kBlurH - 1) / kBlurH);
    using K = void (*)(BlurArgs);
    static const K kernels[kMaxRadius + 1] = {k_blur<0>, k_blur<1>, k_blur<2>, k_blur<3>, k_blur<4>, k_blur<5>,
                                              k_blur<6>, k_blur<7>, k_blur<8>, k_blur<9>, k_blur<10>, k_blur<11>,
                                              k_blur<12>, k_blur<13>, k_blur<14>, k_blur<15>};
    hipLaunchKernelGGL(kernels[tr->radius], grid, dim3(kBlurThreads), 0, tr->s[0], b);
    return launch_check("k_blur");
}

PoseArgs pose_args(pf_tracker *tr, int *frame, const pf_tracker_outputs &out)
{
    PoseArgs pa{};
    pa.arm[0] = arm_dev(tr, 0);
    pa.arm[1] = arm_dev(tr, 1);
    pa.frame = frame;
    pa.pose = out.pose;
    pa.pos3d = out.pos3d;
    pa.tf = out.tf;
    pa.pose2d = out.pose2d;
    pa.npose = tr->npose;
    pa.fx = tr->cfg.fx;
    pa.fy = tr->cfg.fy;
    pa.cx = tr->cfg.cx;
    pa.cy = tr->cfg.cy;
    return pa;
}

int launch_pose(pf_tracker *tr, int *frame, const pf_tracker_outputs &out)
{
    hipLaunchKernelGGL(k_pose, dim3(1), dim3(64), 0, tr->s[0], pose_args(tr, frame, out));
    return launch_check("k_pose");
}

// Whether the pose kernel adds the estimates' block sums itself (paired arms
// of a state dimension whose reduction fits its LDS).
bool pose_adds_estimates(pf_tracker *tr) { return tr->D[0] == tr->D[1] && (tr->D[0] == 8 || tr->D[0] == 12); }

int launch_pose_est(pf_tracker *tr, int *frame, const pf_tracker_outputs &out)
{
    const PoseArgs pa = pose_args(tr, frame, out);
    const EstFinalArgs e0 = eng::estimate_final_args(tr->arm[0]), e1 = eng::estimate_final_args(tr->arm[1]);
    if (tr->D[0] == 8)
        hipLaunchKernelGGL((k_pose_est<8>), dim3(1), dim3(512), 0, tr->s[0], pa, e0, e1);
    else
        hipLaunchKernelGGL((k_pose_est<12>), dim3(1), dim3(512), 0, tr->s[0], pa, e0, e1);
    return launch_check("k_pose_est");
}

// One frame of the pipeline; `parity` alternates from frame to frame (it
// selects the handles' frame counter and maximum slot).
int enqueue_frame(pf_tracker *tr, int parity, int T, const uint8_t *lik, const int *roi, const pf_tracker_outputs &out)
{
    int st = launch_blur(tr, lik, tr->fctr, tr->blur);
    if (st) return st;
    const PropParams p = prop_params(tr, roi, tr->fctr);
    const ArmDev a0 = arm_dev(tr, 0), a1 = arm_dev(tr, 1);
    const int pairs = tr->n_c / 2;
    const int pblocks = std::min((pairs + kProposeBlock - 1) / kProposeBlock, tr->prop_blocks);
    hipLaunchKernelGGL(k_propose, dim3(pblocks, 2), dim3(kProposeBlock), 0,
                       tr->s[0], p, a0, a1, (const uint8_t *)tr->blur, tr->rec);
    st = launch_check("k_propose");
    if (st) return st;
    SelectArgs sa[2];
    for (int a = 0; a < 2; ++a) {
        sa[a].pend = proposal_pending(tr, a);
        sa[a].fs = eng::scratch(tr->arm[a]);
        sa[a].bins = tr->bins[a];
        sa[a].Z = tr->Z[a];
    }
    const size_t sel_lds = 8 * (size_t)resolve_lds_words(sa[0].pend.nb);
    const uint32_t *u_arm[2] = {tr->uni, tr->uni + 2 * (size_t)tr->uni_cap};
    double *lml[2] = {out.lml, out.lml ? out.lml + T : nullptr};
    if (eng::pair_compatible(tr->arm[0], tr->arm[1])) {
        // both arms advance together on stream 0, one launch per stage
        hipLaunchKernelGGL(k_prop_select_pair, dim3((tr->N + 255) / 256, 2), dim3(256), sel_lds, tr->s[0], p, sa[0],
                           sa[1], (const FrameRec *)tr->rec, tr->N);
        st = launch_check("k_prop_select_pair");
        const bool fold = pose_adds_estimates(tr);
        if (!st) st = eng::step_ref_pair(tr->arm[0], tr->arm[1], tr->Z[0], tr->Z[1], u_arm[0], u_arm[1], parity,
                                         lml[0], lml[1], !fold);
        if (st) return st;
        return fold ? launch_pose_est(tr, tr->fctr, out) : launch_pose(tr, tr->fctr, out);
    }
    // otherwise each arm on its own stream, forked here and joined before the pose
    TR_TRY(hipEventRecord(tr->ev_fork[parity], tr->s[0]));
    TR_TRY(hipStreamWaitEvent(tr->s[1], tr->ev_fork[parity], 0));
    for (int a = 0; a < 2; ++a) {
        pf_handle *h = tr->arm[a];
        hipLaunchKernelGGL(k_prop_select, dim3((tr->N + 255) / 256), dim3(256), sel_lds, tr->s[a], p, a, sa[a],
                           (const FrameRec *)tr->rec, tr->N);
        st = launch_check("k_prop_select");
        if (st) return st;
        st = eng::step_ref(h, tr->Z[a], u_arm[a], parity, lml[a], true);
        if (st) return st;
    }
    TR_TRY(hipEventRecord(tr->ev_join[parity], tr->s[1]));
    TR_TRY(hipStreamWaitEvent(tr->s[0], tr->ev_join[parity], 0));
    return launch_pose(tr, tr->fctr, out);
}

// Frame uniforms of global frames [g0, g0 + T) into the device table.
int upload_uniforms(pf_tracker *tr, int T)
{
    if (T > tr->uni_cap) {
        (void)hipFree(tr->uni);
        tr->uni = nullptr;
        const int cap = std::max(T, 64);
        TR_TRY(talloc(&tr->uni, (size_t)6 * cap));
        tr->uni_cap = cap;
    }
    const size_t cap = (size_t)tr->uni_cap;
    tr->h_uni.assign(6 * cap, 0u);
    std::vector<uint32_t> w(8 * (size_t)T);
    philox_u32(tr->cfg.seed, kStreamFrame, 8 * (uint64_t)tr->frames_done, 8 * (int64_t)T, w.data());
    for (int t = 0; t < T; ++t) {
        const uint32_t *f = w.data() + 8 * (size_t)t;
        tr->h_uni[2 * (size_t)t] = f[0];
        tr->h_uni[2 * (size_t)t + 1] = f[1];
        tr->h_uni[2 * cap + 2 * (size_t)t] = f[2];
        tr->h_uni[2 * cap + 2 * (size_t)t + 1] = f[3];
        tr->h_uni[4 * cap + 2 * (size_t)t] = f[4];
        tr->h_uni[4 * cap + 2 * (size_t)t + 1] = f[5];
    }
    TR_TRY(hipMemcpyAsync(tr->uni, tr->h_uni.data(), sizeof(uint32_t) * tr->h_uni.size(), hipMemcpyHostToDevice,
                          tr->s[0]));
    TR_TRY(hipStreamSynchronize(tr->s[0]));
    return PF_OK;
}

// Estimates of both arms' current tracks, stream 1 joined into stream 0.
int prepare_estimates(pf_tracker *tr)
{
    for (int a = 0; a < 2; ++a) {
        int st = eng::estimate(tr->arm[a]);
        if (st) return st;
    }
    TR_TRY(hipEventRecord(tr->ev_pre, tr->s[1]));
    TR_TRY(hipStreamWaitEvent(tr->s[0], tr->ev_pre, 0));
    return PF_OK;
}

// Copies the staged device outputs of one frame to the host pointers of o
// (each may be NULL) and waits for both streams.
int copy_stage_out(pf_tracker *tr, const pf_tracker_outputs &o)
{
    const pf_tracker_outputs &d = tr->stage_out;
    hipStream_t s = tr->s[0];
    if (o.lml) TR_TRY(hipMemcpyAsync(o.lml, d.lml, sizeof(double) * 2, hipMemcpyDeviceToHost, s));
    if (o.pose) TR_TRY(hipMemcpyAsync(o.pose, d.pose, sizeof(double) * 2 * tr->npose, hipMemcpyDeviceToHost, s));
    if (o.pos3d) TR_TRY(hipMemcpyAsync(o.pos3d, d.pos3d, sizeof(double) * 30, hipMemcpyDeviceToHost, s));
    if (o.tf) TR_TRY(hipMemcpyAsync(o.tf, d.tf, sizeof(double) * 24, hipMemcpyDeviceToHost, s));
    if (o.pose2d) TR_TRY(hipMemcpyAsync(o.pose2d, d.pose2d, sizeof(double) * 16, hipMemcpyDeviceToHost, s));
    TR_TRY(hipStreamSynchronize(tr->s[0]));
    TR_TRY(hipStreamSynchronize(tr->s[1]));
    return PF_OK;
}

int status_of_arms(pf_tracker *tr)
{
    unsigned word = 0;
    for (int a = 0; a < 2; ++a) {
        unsigned w = 0;
        int st = eng::take_status(tr->arm[a], &w);
        if (st) return st;
        word |= w;
    }
    return status_code(word);
}

}  // namespace

// =================================================================== C ABI
extern "C" {

void pf_tracker_config_default(pf_tracker_config *cfg)
{
    if (!cfg) return;
    cfg->rows = 480;
    cfg->cols = 640;
    cfg->proposals_per_particle = 10;
    cfg->p_assoc = 0.05;
    cfg->z_floor = 1e-4;
    cfg->spread = 0.8;
    cfg->blur_ksize = 15;
    cfg->blur_sigma = 3.0;
    cfg->fx = 525.0;
    cfg->fy = 525.0;
    cfg->cx = 319.5;
    cfg->cy = 239.5;
    cfg->seed = 1;
}

int pf_tracker_destroy(pf_tracker *tr)
{
    if (!tr) return PF_OK;
    (void)hipSetDevice(tr->device);
    for (int a = 0; a < 2; ++a) {
        (void)hipFree(tr->hp[a]);
        (void)hipFree(tr->pm[a]);
        (void)hipFree(tr->w[a]);
        (void)hipFree(tr->bins[a]);
        (void)hipFree(tr->Z[a]);
        if (tr->ev_fork[a]) (void)hipEventDestroy(tr->ev_fork[a]);
        if (tr->ev_join[a]) (void)hipEventDestroy(tr->ev_join[a]);
    }
    (void)hipFree(tr->taps);
    (void)hipFree(tr->img_stage);
    (void)hipFree(tr->blur);
    (void)hipFree(tr->roi_stage);
    (void)hipFree(tr->rec);
    (void)hipFree(tr->fctr);
    (void)hipFree(tr->fctr_aux);
    (void)hipFree(tr->uni);
    (void)hipFree(tr->stage_out.lml);
    (void)hipFree(tr->stage_out.pose);
    (void)hipFree(tr->stage_out.pos3d);
    (void)hipFree(tr->stage_out.tf);
    (void)hipFree(tr->stage_out.pose2d);
    if (tr->ev_pre) (void)hipEventDestroy(tr->ev_pre);
    if (tr->ev_t0) (void)hipEventDestroy(tr->ev_t0);
    if (tr->ev_t1) (void)hipEventDestroy(tr->ev_t1);
    delete tr;
    return PF_OK;
}

int pf_tracker_create(pf_tracker **out, const pf_tracker_config *cfg, pf_handle *arm1, pf_handle *arm2,
                      int32_t npose, const double *pca_proj1, const double *pca_mean1, const double *pca_proj2,
                      const double *pca_mean2)
{
    if (!out) return set_error(PF_ERR_ARG, "null output");
    *out = nullptr;
    if (!arm1 || !arm2 || arm1 == arm2) return set_error(PF_ERR_ARG, "the tracker needs two distinct arm handles");
    if (!pca_proj1 || !pca_mean1 || !pca_proj2 || !pca_mean2) return set_error(PF_ERR_ARG, "null PCA matrices");
    pf_tracker_config c;
    pf_tracker_config_default(&c);
    if (cfg) c = *cfg;
    eng::Info in[2];
    for (int a = 0; a < 2; ++a) {
        int st = eng::info(a ? arm2 : arm1, &in[a]);
        if (st) return st;
        if (!in[a].model) return set_error(PF_ERR_STATE, "arm %d has no model", a + 1);
        if (in[a].m != 6) return set_error(PF_ERR_ARG, "arm %d measures %d rows; the tracker's measurement has 6", a + 1, in[a].m);
        if (in[a].D > kMaxArmD) return set_error(PF_ERR_UNSUPPORTED, "arm %d state dimension %d > %d", a + 1, in[a].D, kMaxArmD);
    }
    if (in[0].N != in[1].N) return set_error(PF_ERR_ARG, "arms with different particle counts (%d, %d)", in[0].N, in[1].N);
    if (in[0].device != in[1].device) return set_error(PF_ERR_ARG, "arms on different devices");
    if (npose < 21) return set_error(PF_ERR_ARG, "npose = %d: get3Dpose reads 21 pose entries", npose);
    if (npose > 64) return set_error(PF_ERR_ARG, "npose = %d > 64", npose);
    if (c.blur_ksize < 1 || c.blur_ksize > 2 * kMaxRadius + 1 || !(c.blur_ksize & 1))
        return set_error(PF_ERR_ARG, "blur_ksize %d: odd, at most %d", c.blur_ksize, 2 * kMaxRadius + 1);
    if (!(c.blur_sigma > 0.0)) return set_error(PF_ERR_ARG, "blur_sigma must be positive");
    const int R = c.blur_ksize / 2;
    if (c.rows <= R || c.cols <= R) return set_error(PF_ERR_ARG, "image %d x %d smaller than the blur radius", c.rows, c.cols);
    if (c.proposals_per_particle < 1) return set_error(PF_ERR_ARG, "proposals_per_particle must be >= 1");
    const long long nc = (long long)c.proposals_per_particle * in[0].N;
    if (nc > kMaxProposals) return set_error(PF_ERR_ARG, "%lld proposals per arm exceed %d", nc, kMaxProposals);
    if (nc & 1) return set_error(PF_ERR_ARG, "the proposal count per arm (%lld) must be even", nc);
    if (!(c.p_assoc >= 0.0 && c.p_assoc < 0.5) || !(c.z_floor > 0.0) || !(c.spread > 0.0))
        return set_error(PF_ERR_ARG, "bad association constants");
    if (!(c.fx != 0.0 && c.fy != 0.0)) return set_error(PF_ERR_ARG, "zero focal length");
    // the arms' state in rows: their paired step launches have twice the
    // blocks of one filter, where rows are the faster layout
    for (pf_handle *arm : {arm1, arm2}) {
        const int st = pf_set_state_layout(arm, PF_LAYOUT_ROWS);
        if (st) return st;
    }

    pf_tracker *tr = new pf_tracker();
    tr->cfg = c;
    tr->arm[0] = arm1;
    tr->arm[1] = arm2;
    tr->s[0] = in[0].stream;
    tr->s[1] = in[1].stream;
    tr->N = in[0].N;
    tr->D[0] = in[0].D;
    tr->D[1] = in[1].D;
    tr->npose = npose;
    tr->n_c = (int)nc;
    tr->device = in[0].device;
    tr->radius = R;
    auto fail = [&](int st) {
        pf_tracker_destroy(tr);
        return st;
    };
    auto hip_ok = [&](hipError_t e, const char *what) {
        return e == hipSuccess ? PF_OK : set_error(PF_ERR_HIP, "%s failed: %s", what, hipGetErrorString(e));
    };
    int st = tr_check(tr);
    if (st) return fail(st);
    {
        // k_propose's grid-stride loop: as many blocks as are resident at once
        // (both arms together), so no block waits for a slot
        int cus = 0, per_cu = 0;
        st = hip_ok(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, tr->device),
                    "hipDeviceGetAttribute");
        if (!st)
            st = hip_ok(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_propose, kProposeBlock, 0),
                        "hipOccupancyMaxActiveBlocksPerMultiprocessor");
        if (st) return fail(st);
        tr->prop_blocks = std::max(1, (cus * per_cu) / 2);
    }
    // Hp = pca_proj^T (npose x D), pca_mean as given
    for (int a = 0; a < 2; ++a) {
        const double *P = a ? pca_proj2 : pca_proj1;
        const int D = tr->D[a];
        std::vector<double> hp((size_t)npose * D);
        for (int r = 0; r < npose; ++r)
            for (int k = 0; k < D; ++k) hp[(size_t)r * D + k] = P[(size_t)k * npose + r];
        st = hip_ok(talloc(&tr->hp[a], hp.size()), "hipMalloc");
        if (!st) st = hip_ok(talloc(&tr->pm[a], (size_t)npose), "hipMalloc");
        if (!st) st = hip_ok(hipMemcpy(tr->hp[a], hp.data(), sizeof(double) * hp.size(), hipMemcpyHostToDevice), "hipMemcpy");
        if (!st) st = hip_ok(hipMemcpy(tr->pm[a], a ? pca_mean2 : pca_mean1, sizeof(double) * npose, hipMemcpyHostToDevice), "hipMemcpy");
        if (!st) st = hip_ok(talloc(&tr->w[a], (size_t)tr->n_c), "hipMalloc");
        if (!st) st = hip_ok(talloc(&tr->bins[a], (size_t)tr->N), "hipMalloc");
        if (!st) st = hip_ok(talloc(&tr->Z[a], (size_t)6 * tr->N), "hipMalloc");
        if (!st) st = eng::reserve_resample(a ? arm2 : arm1, tr->n_c);
        if (!st) st = hip_ok(hipEventCreateWithFlags(&tr->ev_fork[a], hipEventDisableTiming), "hipEventCreate");
        if (!st) st = hip_ok(hipEventCreateWithFlags(&tr->ev_join[a], hipEventDisableTiming), "hipEventCreate");
        if (st) return fail(st);
    }
    // Gaussian taps exp(-x^2 / (2 sigma^2)), x = -R..R, normalised by their sum
    std::vector<double> g(2 * R + 1);
    double sum = 0.0;
    const double s2 = -0.5 / (c.blur_sigma * c.blur_sigma);
    for (int i = -R; i <= R; ++i) {
        g[i + R] = det_exp((double)(i * i) * s2);
        sum += g[i + R];
    }
    std::vector<float> taps(2 * R + 1);
    for (int i = 0; i <= 2 * R; ++i) taps[i] = (float)(g[i] / sum);
    const size_t px = (size_t)c.rows * c.cols;
    st = hip_ok(talloc(&tr->taps, taps.size()), "hipMalloc");
    if (!st) st = hip_ok(hipMemcpy(tr->taps, taps.data(), sizeof(float) * taps.size(), hipMemcpyHostToDevice), "hipMemcpy");
    if (!st) st = hip_ok(talloc(&tr->img_stage, px), "hipMalloc");
    if (!st) st = hip_ok(talloc(&tr->blur, px), "hipMalloc");
    if (!st) st = hip_ok(talloc(&tr->roi_stage, 4), "hipMalloc");
    if (!st) st = hip_ok(talloc(&tr->rec, 1), "hipMalloc");
    if (!st) st = hip_ok(hipMemset(tr->rec, 0, sizeof(FrameRec)), "hipMemset");
    if (!st) st = hip_ok(talloc(&tr->fctr, 1), "hipMalloc");
    if (!st) st = hip_ok(talloc(&tr->fctr_aux, 1), "hipMalloc");
    if (!st) st = hip_ok(talloc(&tr->stage_out.lml, 2), "hipMalloc");
    if (!st) st = hip_ok(talloc(&tr->stage_out.pose, (size_t)2 * npose), "hipMalloc");
    if (!st) st = hip_ok(talloc(&tr->stage_out.pos3d, 30), "hipMalloc");
    if (!st) st = hip_ok(talloc(&tr->stage_out.tf, 24), "hipMalloc");
    if (!st) st = hip_ok(talloc(&tr->stage_out.pose2d, 16), "hipMalloc");
    if (!st) st = hip_ok(hipEventCreateWithFlags(&tr->ev_pre, hipEventDisableTiming), "hipEventCreate");
    if (!st) st = hip_ok(hipEventCreate(&tr->ev_t0), "hipEventCreate");
    if (!st) st = hip_ok(hipEventCreate(&tr->ev_t1), "hipEventCreate");
    if (st) return fail(st);
    tr->frames_done = 0;
    tr->uniform_frame = 0;
    *out = tr;
    return PF_OK;
}

int pf_tracker_reset(pf_tracker *tr)
{
    int st = tr_check(tr);
    if (st) return st;
    tr->frames_done = 0;
    tr->uniform_frame = 0;
    return PF_OK;
}

int pf_tracker_num_proposals(pf_tracker *tr, int32_t *n_c)
{
    if (!tr || !n_c) return set_error(PF_ERR_ARG, "null argument");
    *n_c = tr->n_c;
    return PF_OK;
}

int pf_tracker_run_dev(pf_tracker *tr, int32_t T, const uint8_t *likelihood, const int32_t *roi,
                       const pf_tracker_outputs *outputs, int32_t flags)
{
    int st = tr_check(tr);
    if (st) return st;
    if (T <= 0 || !likelihood || !roi) return set_error(PF_ERR_ARG, "bad run arguments");
    pf_tracker_outputs out{};
    if (outputs) out = *outputs;
    for (int a = 0; a < 2; ++a) {
        eng::Info in;
        st = eng::info(tr->arm[a], &in);
        if (st) return st;
        if (!in.tracks) return set_error(PF_ERR_STATE, "arm %d tracks not initialised", a + 1);
        if (in.N != tr->N) return set_error(PF_ERR_STATE, "arm %d particle count changed", a + 1);
    }
    const bool graph = (flags & 1) != 0;
    st = upload_uniforms(tr, T);
    if (st) return st;
    TR_TRY(hipEventRecord(tr->ev_t0, tr->s[0]));
    hipLaunchKernelGGL(k_tr_set, dim3(1), dim3(64), 0, tr->s[0], tr->fctr, 0);
    st = launch_check("k_tr_set");
    for (int a = 0; !st && a < 2; ++a) st = eng::set_frame(tr->arm[a], 0);
    if (!st) st = prepare_estimates(tr);
    if (st) return st;
    auto frame = [&](int f) { return enqueue_frame(tr, f & 1, T, likelihood, roi, out); };
    st = frame(0);
    if (st) return st;
    // kTrGraphFrames consecutive frames (an even count: frame f uses the
    // buffers and counters of parity f & 1) captured once and replayed
    const int reps = graph ? (T - 1) / kTrGraphFrames : 0;
    if (reps > 0) {
        hipGraph_t gr;
        hipGraphExec_t ge;
        TR_TRY(hipStreamBeginCapture(tr->s[0], hipStreamCaptureModeThreadLocal));
        int sc = 0;
        for (int f = 1; f <= kTrGraphFrames && !sc; ++f) sc = frame(f);
        hipError_t ec = hipStreamEndCapture(tr->s[0], &gr);
        if (sc) return sc;
        if (ec != hipSuccess) return set_error(PF_ERR_HIP, "graph capture failed: %s", hipGetErrorString(ec));
        TR_TRY(hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0));
        for (int q = 0; q < reps; ++q) TR_TRY(hipGraphLaunch(ge, tr->s[0]));
        TR_TRY(hipStreamSynchronize(tr->s[0]));
        (void)hipGraphExecDestroy(ge);
        (void)hipGraphDestroy(gr);
    }
    for (int f = 1 + reps * kTrGraphFrames; f < T; ++f) {
        st = frame(f);
        if (st) return st;
    }
    TR_TRY(hipEventRecord(tr->ev_t1, tr->s[0]));
    // arm 2's stream orders its later work after the run (the frames may
    // have advanced arm 2 on stream 0)
    TR_TRY(hipStreamWaitEvent(tr->s[1], tr->ev_t1, 0));
    tr->frames_done += T;
    tr->ran = true;
    return PF_OK;
}

int pf_tracker_synchronize(pf_tracker *tr)
{
    int st = tr_check(tr);
    if (st) return st;
    TR_TRY(hipStreamSynchronize(tr->s[0]));
    TR_TRY(hipStreamSynchronize(tr->s[1]));
    return PF_OK;
}

int pf_tracker_last_run_ms(pf_tracker *tr, double *ms)
{
    int st = tr_check(tr);
    if (st) return st;
    if (!ms) return set_error(PF_ERR_ARG, "null output");
    if (!tr->ran) return set_error(PF_ERR_STATE, "no run yet");
    float v = 0.f;
    TR_TRY(hipEventSynchronize(tr->ev_t1));
    TR_TRY(hipEventElapsedTime(&v, tr->ev_t0, tr->ev_t1));
    *ms = v;
    return PF_OK;
}

int pf_tracker_frame(pf_tracker *tr, const uint8_t *likelihood, const int32_t *roi, const pf_tracker_outputs *out)
{
    int st = tr_check(tr);
    if (st) return st;
    pf_tracker_outputs o{};
    if (out) o = *out;
    if (!roi) {            // no face: pfPose.cpp:370-377
        tr->uniform_frame = tr->frames_done;
        if (o.pose2d) std::memset(o.pose2d, 0, sizeof(double) * 16);
        return PF_OK;
    }
    if (!likelihood) return set_error(PF_ERR_ARG, "null image");
    if (roi[2] <= 0 || roi[3] <= 0) return set_error(PF_ERR_ARG, "empty face ROI");
    const size_t px = (size_t)tr->cfg.rows * tr->cfg.cols;
    TR_TRY(hipMemcpyAsync(tr->img_stage, likelihood, px, hipMemcpyHostToDevice, tr->s[0]));
    TR_TRY(hipMemcpyAsync(tr->roi_stage, roi, sizeof(int32_t) * 4, hipMemcpyHostToDevice, tr->s[0]));
    TR_TRY(hipStreamSynchronize(tr->s[0]));
    st = pf_tracker_run_dev(tr, 1, tr->img_stage, tr->roi_stage, &tr->stage_out, 0);
    if (st) return st;
    st = copy_stage_out(tr, o);
    if (st) return st;
    return status_of_arms(tr);
}

int pf_tracker_get_last(pf_tracker *tr, uint8_t *blurred, double *proposals, double *weights, int32_t *bins,
                        float *measurements)
{
    int st = pf_tracker_synchronize(tr);
    if (st) return st;
    if (!tr->ran) return set_error(PF_ERR_STATE, "no frame yet");
    hipStream_t s = tr->s[0];
    const size_t px = (size_t)tr->cfg.rows * tr->cfg.cols;
    if (blurred) TR_TRY(hipMemcpyAsync(blurred, tr->blur, px, hipMemcpyDeviceToHost, s));
    if (proposals) {
        double *d = nullptr;
        TR_TRY(talloc(&d, (size_t)4 * tr->n_c));
        PropParams p = prop_params(tr, nullptr, nullptr);
        hipLaunchKernelGGL(k_dump_proposals, dim3((tr->n_c + 255) / 256, 2), dim3(256), 0, s, p,
                           (const FrameRec *)tr->rec, d);
        st = launch_check("k_dump_proposals");
        hipError_t e = hipSuccess;
        if (!st) e = hipMemcpyAsync(proposals, d, sizeof(double) * 4 * tr->n_c, hipMemcpyDeviceToHost, s);
        if (!st && e == hipSuccess) e = hipStreamSynchronize(s);
        (void)hipFree(d);
        if (st) return st;
        if (e != hipSuccess) return set_error(PF_ERR_HIP, "proposal copy failed: %s", hipGetErrorString(e));
    }
    for (int a = 0; a < 2; ++a) {
        if (weights)
            TR_TRY(hipMemcpyAsync(weights + (size_t)a * tr->n_c, tr->w[a], sizeof(double) * tr->n_c,
                                  hipMemcpyDeviceToHost, s));
        if (bins)
            TR_TRY(hipMemcpyAsync(bins + (size_t)a * tr->N, tr->bins[a], sizeof(int) * tr->N, hipMemcpyDeviceToHost, s));
        if (measurements)
            TR_TRY(hipMemcpyAsync(measurements + (size_t)a * 6 * tr->N, tr->Z[a], sizeof(float) * 6 * tr->N,
                                  hipMemcpyDeviceToHost, s));
    }
    TR_TRY(hipStreamSynchronize(s));
    return PF_OK;
}

int pf_tracker_pose(pf_tracker *tr, const pf_tracker_outputs *out)
{
    int st = tr_check(tr);
    if (st) return st;
    for (int a = 0; a < 2; ++a) {
        eng::Info in;
        st = eng::info(tr->arm[a], &in);
        if (st) return st;
        if (!in.tracks) return set_error(PF_ERR_STATE, "arm %d tracks not initialised", a + 1);
    }
    st = prepare_estimates(tr);
    if (st) return st;
    hipLaunchKernelGGL(k_tr_set, dim3(1), dim3(64), 0, tr->s[0], tr->fctr_aux, 0);
    st = launch_check("k_tr_set");
    if (!st) st = launch_pose(tr, tr->fctr_aux, tr->stage_out);
    if (st) return st;
    pf_tracker_outputs o{};
    if (out) o = *out;
    o.lml = nullptr;
    return copy_stage_out(tr, o);
}

int pf_tracker_blur(pf_tracker *tr, const uint8_t *src, uint8_t *dst)
{
    int st = tr_check(tr);
    if (st) return st;
    if (!src || !dst) return set_error(PF_ERR_ARG, "null image");
    const size_t px = (size_t)tr->cfg.rows * tr->cfg.cols;
    uint8_t *d = nullptr;
    TR_TRY(talloc(&d, px));
    hipError_t e = hipMemcpyAsync(tr->img_stage, src, px, hipMemcpyHostToDevice, tr->s[0]);
    if (e == hipSuccess) {
        st = launch_blur(tr, tr->img_stage, nullptr, d);
        if (!st) e = hipMemcpyAsync(dst, d, px, hipMemcpyDeviceToHost, tr->s[0]);
        if (!st && e == hipSuccess) e = hipStreamSynchronize(tr->s[0]);
    }
    (void)hipFree(d);
    if (st) return st;
    if (e != hipSuccess) return set_error(PF_ERR_HIP, "blur copy failed: %s", hipGetErrorString(e));
    return PF_OK;
}

}  // extern "C"
