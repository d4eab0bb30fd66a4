// pf_engine.hpp -- internal interface of one filter handle for code that
// chains several handles on the device (the pose tracker of pf_tracker.hip).
//
// Every call enqueues on the handle's own stream, never synchronises the host
// and allocates nothing (except reserve_resample), so a caller may record the
// calls into a HIP graph.  The handle stays opaque here.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pf_rbpf.h"
#include "pf_kernels.hpp"

namespace mkfbt {
namespace eng {

struct Info {
    int N, D, m, J, npose, device;
    bool model, tracks;
    hipStream_t stream;
};

int info(pf_handle *h, Info *out);
// Resampler scratch for weight vectors of up to L entries (allocates).
int reserve_resample(pf_handle *h, int L);
// Device address of the estimate written by estimate(): D doubles, model basis.
const double *estimate_ptr(pf_handle *h);
// Sets both frame counters of the handle to f.
int set_frame(pf_handle *h, int f);
// getEstimator (pf2DRao.cpp:23-31) of the current tracks into estimate_ptr().
int estimate(pf_handle *h);
// ParticleFilter::update (pf2DRao.cpp:125-158) with one measurement column per
// particle: Z (m x N floats, leading dimension N), uniforms uni[2t], uni[2t+1]
// for t = the handle's frame counter `parity`; the frame's log mean weight goes
// to lml[t].  Ancestors are materialised; with_estimate also computes
// getEstimator of the resampled tracks into estimate_ptr() (fused with the
// ancestors), otherwise estimate() may follow.
int step_ref(pf_handle *h, const float *Z, const uint32_t *uni, int parity, double *lml, bool with_estimate);
// The resampler scratch of the handle viewed as a pending resampling of L
// weights (block-local scans cloc, block sums bsum, kWeightsBlock inputs per
// block; a caller's own kernel fills them and resolves the ancestors with
// resolve_prefix / resolve_search), and the frame scratch (U_pending, status).
Pending pending(pf_handle *h, int L);
FrameScratch *scratch(pf_handle *h);
// Whether step_ref_pair can advance h0 and h1 together (same state
// dimension, particle count, device, selection basis with 3 joints, model
// image size and block size).
bool pair_compatible(pf_handle *h0, pf_handle *h1);
// step_ref(h0, ...) and step_ref(h1, ...) with with_estimate, in one launch
// per stage on h0's stream (grid.y = filter); the frame counter `parity` of
// each handle is read.  The caller orders h1's stream around it.  Without
// final_sums the estimates' block sums are left for the caller to add (the
// arguments of estimate_final_args; k_estimate_final's order keeps the
// result identical).
int step_ref_pair(pf_handle *h0, pf_handle *h1, const float *Z0, const float *Z1, const uint32_t *uni0,
                  const uint32_t *uni1, int parity, double *lml0, double *lml1, bool final_sums = true);
// The block sums of the handle's last fused estimate and where it goes.
EstFinalArgs estimate_final_args(pf_handle *h);
// Reads and clears the handle's status word (kStatus* bits; synchronises
// the handle's stream).
int take_status(pf_handle *h, unsigned *word);

}  // namespace eng
}  // namespace mkfbt
