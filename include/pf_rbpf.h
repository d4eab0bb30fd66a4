/*
 * pf_rbpf.h -- C ABI of the MI355X (gfx950) particle-filter hot path.
 *
 * One handle = one ParticleFilter of the reference (src/pf2DRao.h:13-31) with
 * its my_gmm (src/my_gmm.h:20-33) and the per-component KF_model tables
 * (src/KF_model.h:8-16), resident in HBM.  The handle owns its device memory
 * and one HIP stream; calls on one handle are serialised by the caller,
 * different handles may be used concurrently.
 *
 * Conventions: plain C types only, no exceptions cross this boundary.  Every
 * entry point returns an int status: 0 = OK; positive = OK with warnings, a
 * bit mask of PF_WARN_DEGENERATE (1: the resampler fell back to uniform
 * weights because every weight was zero or non-finite, the reference's
 * maxWeight()==0 branch) and PF_WARN_ILLCOND (2: some particle's 2 x 2
 * innovation covariance had cond > ~1e6 and was factored in FP64); negative
 * = error (see PF_ERR_*), with a message available from pf_last_error()
 * (thread-local).  PF_ERR_NOT_SPD from a step means the frame completed but
 * some particle's innovation covariance was not positive definite even in
 * FP64: that particle got zero weight and kept its predicted state (the
 * reference's Cholesky failure, src/pf2DRao.cpp:36-52).  Host matrices are
 * row-major double, i.e. the memory of a continuous CV_64F cv::Mat.  Device
 * state is FP32 with packed symmetric covariances, in rows or per-particle
 * records (pf_set_state_layout); log-likelihood accumulation is FP64.
 *
 * Randomness is never drawn inside the library: every entry point that the
 * reference implements with cv::RNG / randn (src/pf2DRao.cpp:96,179-197)
 * takes host-supplied 32-bit uniforms or normals (use pf_philox_u32 for a
 * seeded Philox4x32-10 stream).
 */
#ifndef PF_RBPF_H
#define PF_RBPF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PF_OK 0
#define PF_WARN_DEGENERATE 1
#define PF_WARN_ILLCOND 2
#define PF_ERR_ARG (-1)
#define PF_ERR_HIP (-2)
#define PF_ERR_NOT_SPD (-3)
#define PF_ERR_STATE (-4)
#define PF_ERR_UNSUPPORTED (-5)
#define PF_ERR_IO (-6)

typedef struct pf_handle pf_handle;

typedef struct pf_config {
    int32_t n_particles;     /* N: ParticleFilter(int nParticles), pf2DRao.cpp:13-16 */
    int32_t device;          /* HIP device ordinal */
    int32_t cand_capacity;   /* detections per joint per frame in candidate mode (0 = 16) */
    int32_t block_size;      /* threads per block of the step kernel (0 = autotuned default) */
    double p_detect;         /* PDAF detection probability P_D (default 0.95) */
    double gate_chi2;        /* PDAF gate on the 2-dof Mahalanobis distance (default 9.2103, 99%) */
    double clutter_density;  /* PDAF clutter density lambda, detections per px^2 (default 3/(640*480)) */
} pf_config;

/* Fills the defaults listed above (n_particles = 500 as pfPose.cpp:57). */
void pf_config_default(pf_config *cfg);
/* Message of the last failing call on this thread ("" if none). */
const char *pf_last_error(void);

/* new ParticleFilter(N) (pfPose.cpp:58-59 -> pf2DRao.cpp:13-16). */
int pf_create(pf_handle **out, const pf_config *cfg);
int pf_destroy(pf_handle *h);

/* The K calls of my_gmm::loadGaussian (my_gmm.cpp:45-75, driven by
 * pfPose.cpp:61-65) in one go.  Arrays are in the layouts of the GMM yml:
 *   means K x d, covs (K*d) x d (component k = rows [k*d,(k+1)*d)),
 *   weights K, gamma K, pca_proj d x npose, pca_mean npose,
 *   sel m indices into the npose pose vector (the rows of H1, my_gmm.cpp:61-67;
 *   NULL = the reference's {9,10,0,1,12,13}, m must then be 6),
 *   R m x m (NULL = 100 I, my_gmm.cpp:54; must be block diagonal in 2x2 joints).
 * Builds F = g I, B = (1-g) mu, Q = (1-g^2) Sigma, H = H1 pca_proj^T,
 * BH = H1 pca_mean^T and uploads them.  Limits: d in {8, 12, 16},
 * m even and <= 16, K <= 256. */
int pf_load_model(pf_handle *h, int32_t K, int32_t d, int32_t npose, int32_t m, const int32_t *sel,
                  const double *means, const double *covs, const double *weights,
                  const double *gamma, const double *pca_proj, const double *pca_mean,
                  const double *R);

/* Per-joint measurement kind for candidate mode: clutter[b] != 0 makes joint b
 * (rows 2b, 2b+1) a PDAF joint; 0 = a certain joint (default: all certain). */
int pf_set_joint_clutter(pf_handle *h, const int32_t *clutter);

/* my_gmm::resetTracker(bins) (my_gmm.cpp:30-42): x_j = mu_{bins_j},
 * P_j = Sigma_{bins_j}. */
int pf_reset(pf_handle *h, const int32_t *bins);
/* Tracks in and out (gmm.tracks, my_gmm.h:31): x N x d, P N x d x d. */
int pf_set_tracks(pf_handle *h, const double *x, const double *P);
int pf_get_tracks(pf_handle *h, double *x, double *P);

/* ParticleFilter::update(measurement) (pf2DRao.cpp:125-158), reference mode:
 * Z is the m x N measurement matrix (column j = particle j).  u_comp is the
 * uniform of the component draw resample(gmm.weight, N) (:128), u_res the
 * uniform of the track resampling (:152), both as 32-bit integers
 * (u = U / 2^32).  Optional outputs (NULL to skip): lw[N] log-weights,
 * anc[N] ancestors, comp[N] component indicators, lml the frame's
 * log((1/N) sum exp(lw)). */
int pf_step(pf_handle *h, const double *Z, uint32_t u_comp, uint32_t u_res,
            float *lw, int32_t *anc, int32_t *comp, double *lml);

/* Same step in candidate (PDAF) mode: count[J] detections per joint and
 * cand[J][cap][2] pixel coordinates shared by all particles (cap =
 * cand_capacity).  Certain joints use their first detection (count 0 =
 * joint not observed this frame). */
int pf_step_candidates(pf_handle *h, const int32_t *count, const double *cand, uint32_t u_comp,
                       uint32_t u_res, float *lw, int32_t *anc, int32_t *comp, double *lml);

/* ParticleFilter::getEstimator (pf2DRao.cpp:23-31): xbar = sum_j x_j / N. */
int pf_estimate(pf_handle *h, double *xbar);

/* ParticleFilter::resample(weights, N) (pf2DRao.cpp:175-210) on the GPU:
 * systematic resampling of L linear weights into N non-decreasing indices,
 * fixed-point prefix sums so the result does not depend on the reduction
 * order.  Returns 1 if every weight was zero (uniform fallback). */
int pf_resample(pf_handle *h, const double *w, int32_t L, int32_t N, uint32_t u, int32_t *out);
/* Same on log-weights (the form the step kernel uses internally). */
int pf_resample_log(pf_handle *h, const double *lw, int32_t L, int32_t N, uint32_t u, int32_t *out);

/* Largest L accepted by pf_resample_log_noalloc. */
#define PF_NOALLOC_MAX_WEIGHTS 524288
/* No counterpart in the reference: pf_resample_log for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops that
 * forbid hipMalloc).  It needs no handle and calls no hipMalloc / hipFree:
 * its scratch is a static device array of 4.1 MiB per device, kept until the
 * process ends.  lw (L FP64 log-weights) and anc (N int32, written) are
 * DEVICE pointers on the calling thread's current device; stream is a
 * hipStream_t (NULL = the null stream).  The work is enqueued and the call
 * returns without waiting.  status (device int32, may be NULL) receives 0, or
 * 1 if every weight was zero or non-finite (uniform fallback).  Same indices
 * as pf_resample_log.  L > PF_NOALLOC_MAX_WEIGHTS is rejected (PF_ERR_ARG)
 * before any launch.  Every call on a device shares the one scratch array:
 * calls that overlap on one device (other streams, other threads) race. */
int pf_resample_log_noalloc(const double *lw, int32_t L, int32_t N, uint32_t u, int32_t *anc,
                            int32_t *status, void *stream);

/* ParticleFilter::getSamples(H, M, n, scale) (pf2DRao.cpp:85-103): with
 * y = Hp xbar + M (Hp npose x d = pca_proj^T, M npose), out (2 x n) =
 * y[0:2] + 0.8*scale*normals (normals 2 x n, host-supplied). */
int pf_get_samples(pf_handle *h, const double *Hp, const double *M, int32_t npose, int32_t n,
                   double scale, const double *normals, double *out);
/* ParticleFilter::getSampleProb (pf2DRao.cpp:105-122): w1[c], w2[c] = density
 * of columns of in1, in2 (2 x n) under N(y[0:2], 0.8*scale*I). */
int pf_get_sample_prob(pf_handle *h, const double *Hp, const double *M, int32_t npose,
                       const double *in1, const double *in2, int32_t n, double scale,
                       double *w1, double *w2);

/* --------------------------- device-resident sequences (throughput path) */
/* T frames back to back without host round trips.  All pointers are DEVICE
 * pointers on the handle's device: uniforms T x 2 (u_comp, u_res per frame),
 * lml T (FP64, written), and either Z (T x m x N floats, reference mode) or
 * count (T x J) + cand (T x J x cap x 2 floats, candidate mode).
 * use_graph bit 0 (value 1): every frame is replayed from hipGraphs of 64,
 * 32, ..., 2, 1 frames (for either parity of their first frame) captured on
 * the handle's first such run (the frames read the run's pointers from
 * device memory, so the graphs serve every later run of any length; they
 * are captured again after pf_load_model or pf_set_joint_clutter); bits 1
 * and 2 time the run (see
 * pf_last_run_kernel_time; bit 1 turns the graphs off).  Returns after
 * enqueueing; pf_synchronize waits. */
int pf_run_ref_dev(pf_handle *h, int32_t T, const float *Z, const uint32_t *uniforms,
                   double *lml, int32_t use_graph);
int pf_run_candidates_dev(pf_handle *h, int32_t T, const int32_t *count, const float *cand,
                          const uint32_t *uniforms, double *lml, int32_t use_graph);
/* Waits for the handle's work and returns its status since the last step or
 * synchronize (warnings and PF_ERR_NOT_SPD as for pf_step). */
int pf_synchronize(pf_handle *h);
/* The handle's HIP stream (hipStream_t) for callers that order their own work. */
int pf_get_stream(pf_handle *h, void **stream);
/* Device time of the last pf_run_*_dev call, in ms, and the number of step
 * launches it covered (HIP events on the handle's stream; used for the
 * roofline figure).  use_graph bit 1 (value 2): an event pair around every
 * step launch, no graph; step_ms is their sum, frame_ms the whole run.
 * Bit 2 (value 4): one event pair around all frames of the run (graph or
 * not); step_ms = frame_ms = that time, which is the step launches back to
 * back when a frame is one launch.  Without either bit nothing is timed. */
int pf_last_run_kernel_time(pf_handle *h, double *step_ms, double *frame_ms, int32_t *launches);
/* Frames of the last pf_run_*_dev call that were replayed from graphs (0 or T). */
int pf_last_run_graph_frames(pf_handle *h, int32_t *frames);

/* Device layout of the filter's state; no counterpart in the reference (a
 * cv::Mat per track).  PF_LAYOUT_ROWS: each state component is a row over
 * the particles; PF_LAYOUT_RECORDS: one record per particle (x, then the
 * packed covariance; d <= 12); PF_LAYOUT_AUTO (the default): records when the
 * step's 256-particle blocks fit at two waves per SIMD at once (N <= 512 x
 * compute units), where a frame is bound by latency and by the gather of
 * thinly spread ancestors, rows above (and for the arms of a pose tracker,
 * whose paired launches have twice the blocks).  Every layout gives the same
 * results bit for bit; changing it converts the current tracks.
 * pf_get_state_layout returns the layout in use (ROWS or RECORDS). */
#define PF_LAYOUT_AUTO 0
#define PF_LAYOUT_ROWS 1
#define PF_LAYOUT_RECORDS 2
int pf_set_state_layout(pf_handle *h, int32_t layout);
int pf_get_state_layout(pf_handle *h, int32_t *layout);

/* ---------------------------------------------------------- KF_model */
/* KF_model::predict (KF_model.cpp:11-15) for n independent states sharing
 * one model: x = F x + B, P = F P F^T + Q.  x n x d, P n x d x d, in place. */
int pf_kf_predict(pf_handle *h, int32_t n, int32_t d, const double *F, const double *B,
                  const double *Q, double *x, double *P);
/* KF_model::update (KF_model.cpp:17-25) for n states: z n x m. */
int pf_kf_update(pf_handle *h, int32_t n, int32_t d, int32_t m, const double *H, const double *BH,
                 const double *R, const double *z, double *x, double *P);

/* ------------------------------------------------------------ inputs */
/* cv::FileStorage reads of an OpenCV-YAML "!!opencv-matrix" node
 * (pfPose.cpp:34-53): rows, cols and up to cap values (converted to double,
 * dt d or f).  out may be NULL to query the shape. */
int pf_yml_read_matrix(const char *path, const char *key, int32_t *rows, int32_t *cols,
                       double *out, int64_t cap);
/* ROS camera_info YAML (cal.yml): width, height, K (3x3), D (5), P (3x4). */
int pf_read_camera_info(const char *path, int32_t *width, int32_t *height, double *K,
                        double *D, double *P);
/* Loads a GMM/PCA yml (means, covs, weights, pca_proj, pca_mean, gamma) into
 * the handle.  gamma_path (NULL = path) is the file gamma is read from: the
 * reference reads both arms' gamma from the second arm's file (pfPose.cpp:52-53). */
int pf_load_model_yml(pf_handle *h, const char *path, const char *gamma_path, int32_t m,
                      const int32_t *sel, const double *R);

/* Seeded Philox4x32-10 stream of 32-bit words (host): word i of stream s is
 * philox(ctr = {(i/4) lo, (i/4) hi, s, 0}, key = {seed lo, seed hi})[i % 4]. */
void pf_philox_u32(uint64_t seed, uint32_t stream, uint64_t first, int64_t count, uint32_t *out);

#ifdef __cplusplus
}
#endif
#endif
