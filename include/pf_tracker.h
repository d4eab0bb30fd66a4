/*
 * pf_tracker.h -- C ABI of the pose tracker's per-frame pipeline on the GPU:
 * the work PFTracker::callback (src/pfPose.cpp:332-349) does for each frame
 * around the two arm filters.
 *
 *   1. GaussianBlur of the hand-likelihood image (pfPose.cpp:213).
 *   2. 10 N hand-position proposals per arm: getSamples around the arm's
 *      projected estimate (pf2DRao.cpp:85-103), or uniform over a box below
 *      the face on the first frame (pfPose.cpp:222-235).
 *   3. getSampleProb of every proposal under both arms (pf2DRao.cpp:105-122,
 *      pfPose.cpp:238-239) and the association weight
 *      L / (Pa p_own + Pa p_other + 1e-4 (1 - 2 Pa)), L = blurred image / 255
 *      at the proposal, 0 outside the image (pfPose.cpp:244-293).
 *   4. resample(weights, N) of the proposals (pfPose.cpp:300-301) and the
 *      6 x N measurement [face centre; proposal; face centre x, y + 1.65 h]
 *      (pfPose.cpp:303-320).
 *   5. ParticleFilter::update of both arms (pfPose.cpp:325-326).
 *   6. The pose output e = pca_proj^T getEstimator() + pca_mean
 *      (pfPose.cpp:347-348), get3Dpose (pfPose.cpp:93-127) and the joint
 *      positions publishTFtree and publish2Dpos send (pfPose.cpp:129-208).
 *
 * The tracker borrows two filter handles (include/pf_rbpf.h) with loaded
 * models and initialised tracks (pf_reset with the bins of pfPose.cpp:68-71);
 * when the two arms have the same shape (state dimension, particle count,
 * model image, selection basis) both advance in one launch per stage on
 * arm 1's stream, otherwise arm 1's work runs on arm 1's stream and arm 2's
 * on arm 2's, joined by events; after a run arm 2's stream is ordered after
 * arm 1's.  The tracker owns its own buffers.  Calls on a tracker and on its two
 * handles are serialised by the caller.  Status codes are those of
 * pf_rbpf.h; PF_WARN_DEGENERATE reports that some resampling of the call
 * fell back to uniform weights (every proposal weight zero, the reference's
 * NaN weights branch).  The proposal weights are resampled as fixed-point
 * values floor(2^32 w z) with z = z_floor (1 - 2 p_assoc), the reciprocal of
 * their bound (pf_tracker_create requires 0 <= p_assoc < 1/2, z_floor > 0).
 *
 * Randomness: the proposal positions and the per-frame uniforms are words
 * of the seeded Philox4x32-10 stream of pf_philox_u32.  For global frame g
 * (frames since pf_tracker_create / pf_tracker_reset), proposal c of arm a
 * uses words 2 (g n_c + c) and +1 of stream 16 + a (n_c = proposals per arm);
 * a Gaussian proposal is y + spread * width * (z0, z1), (z0, z1) the
 * Box-Muller pair of those words, a uniform one is box_lo + (box_hi - box_lo)
 * * word 2^-32.  Frame g's uniforms are words 8 g .. 8 g + 5 of stream 24:
 * u_comp and u_res of arm 1, of arm 2, then the proposal-resampling
 * uniforms of arm 1 and arm 2.
 */
#ifndef PF_TRACKER_H
#define PF_TRACKER_H

#include <stdint.h>

#include "pf_rbpf.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pf_tracker pf_tracker;

typedef struct pf_tracker_config {
    int32_t rows, cols;               /* likelihood image (default 480 x 640) */
    int32_t proposals_per_particle;   /* n_c = this x N (default 10, pfPose.cpp:216) */
    double p_assoc;                   /* Pa (default 0.05, pfPose.cpp:249) */
    double z_floor;                   /* constant of Z: z_floor (1 - 2 Pa) (default 1e-4, pfPose.cpp:262) */
    double spread;                    /* proposal stddev and density variance = spread x ROI width
                                         (default 0.8, pf2DRao.cpp:90,114) */
    int32_t blur_ksize;               /* odd, <= 31 (default 15, pfPose.cpp:213) */
    double blur_sigma;                /* default 3 */
    double fx, fy, cx, cy;            /* camera of get3Dpose (default 525, 525, 319.5, 239.5, pfPose.cpp:101) */
    uint64_t seed;                    /* Philox key (default 1) */
} pf_tracker_config;

void pf_tracker_config_default(pf_tracker_config *cfg);

/* Tracker over two arm handles (same N, same device, 6 measurement rows in
 * the reference order head, hand, neck).  pca_proj_a is d x npose and
 * pca_mean_a npose (the yml matrices of pfPose.cpp:42-49, as double);
 * npose >= 21 (the 22-dim pose of get3Dpose).  The next frame uses the
 * uniform first-frame proposals. */
int pf_tracker_create(pf_tracker **out, const pf_tracker_config *cfg, pf_handle *arm1, pf_handle *arm2,
                      int32_t npose, const double *pca_proj1, const double *pca_mean1,
                      const double *pca_proj2, const double *pca_mean2);
int pf_tracker_destroy(pf_tracker *tr);
/* Back to frame 0: the next frame uses the uniform proposals (the arms'
 * tracks are the caller's, e.g. pf_reset). */
int pf_tracker_reset(pf_tracker *tr);

/* Per-frame outputs; every pointer may be NULL.  For T frames:
 *   lml     2 x T      log mean weight of each arm's update (arm-major)
 *   pose    T x 2 x npose   e of each arm (pfPose.cpp:347-348)
 *   pos3d   T x 2 x 15      get3Dpose of each arm, 3 x 5 row-major: rows
 *                           x, y, z; columns hand, elbow, shoulder, head, neck
 *   tf      T x 24     the joints of publishTFtree (pfPose.cpp:129-171)
 *                      relative to its Head frame, in the order and axes
 *                      extractTFdata records them (extractTFdata.cpp:87-103):
 *                      Head, Left_Hand, Right_Hand, Left_Elbow, Right_Elbow,
 *                      Right_Shoulder, Left_Shoulder, Neck, each (x, y, z);
 *                      arm 1 is published as Right_*, arm 2 as Left_*
 *   pose2d  T x 16     the 8 image points of publish2Dpos (pfPose.cpp:
 *                      173-208): hand 1, hand 2, head, neck, elbow 1,
 *                      elbow 2, shoulder 1, shoulder 2, each (x, y) */
typedef struct pf_tracker_outputs {
    double *lml;
    double *pose;
    double *pos3d;
    double *tf;
    double *pose2d;
} pf_tracker_outputs;

/* One frame with host buffers: likelihood rows x cols uint8, roi = face ROI
 * {x_offset, y_offset, width, height} (facetracking ROIs[0]); outputs for
 * T = 1 (host pointers).  roi == NULL is a frame without a face (the
 * reference's callback else-branch, pfPose.cpp:370-377): no update, pose2d
 * zero, and the next frame with a face starts from uniform proposals. */
int pf_tracker_frame(pf_tracker *tr, const uint8_t *likelihood, const int32_t *roi,
                     const pf_tracker_outputs *out);

/* T frames back to back, all pointers on the device: likelihood T x rows x
 * cols, roi T x 4 (every frame has a face), outputs as above.  flags bit 0:
 * capture 16 frames into a HIP graph and replay them.  Returns after
 * enqueueing; call pf_tracker_synchronize. */
int pf_tracker_run_dev(pf_tracker *tr, int32_t T, const uint8_t *likelihood, const int32_t *roi,
                       const pf_tracker_outputs *out, int32_t flags);
int pf_tracker_synchronize(pf_tracker *tr);
/* Wall time of the last pf_tracker_run_dev between its first and last
 * kernel (ms), valid after pf_tracker_synchronize. */
int pf_tracker_last_run_ms(pf_tracker *tr, double *ms);

/* Intermediates of the last frame, host copies (each may be NULL): the
 * blurred image (rows x cols), the proposals of both arms (2 x 2 x n_c:
 * x row then y row per arm), their association weights (2 x n_c), the
 * resampled proposal indices (2 x N) and the measurements (2 x 6 x N). */
int pf_tracker_get_last(pf_tracker *tr, uint8_t *blurred, double *proposals, double *weights,
                        int32_t *bins, float *measurements);

/* The pose output of the arms' current tracks without a filter update
 * (outputs pose, pos3d, tf and pose2d for T = 1, host pointers; lml is
 * ignored). */
int pf_tracker_pose(pf_tracker *tr, const pf_tracker_outputs *out);

/* GaussianBlur(ksize, sigma, BORDER_REFLECT_101) of the tracker's kernel on
 * a host image of the tracker's size (the pipeline's step 1 alone). */
int pf_tracker_blur(pf_tracker *tr, const uint8_t *src, uint8_t *dst);

/* Number of proposals per arm (n_c). */
int pf_tracker_num_proposals(pf_tracker *tr, int32_t *n_c);

#ifdef __cplusplus
}
#endif

#endif /* PF_TRACKER_H */
