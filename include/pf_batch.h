/*
 * pf_batch.h -- many independent Rao-Blackwellised particle filters on one
 * gfx950 GPU, advanced together.
 *
 * No counterpart in the reference: it runs one ParticleFilter per arm,
 * sequentially (src/pfPose.cpp:325-326).  Every sequence here runs the same
 * algorithm as pf_step_candidates() of pf_rbpf.h (pf2DRao.cpp:125-158 with
 * the PDAF joint likelihood), with its own GMM, particle count and length.
 * One persistent step kernel and one weights kernel per frame serve all
 * sequences still active at that frame; the frame loop is captured into a
 * hipGraph.  Tile size and persistence are picked by pf_batch_autotune().
 */
#ifndef PF_BATCH_H
#define PF_BATCH_H

#include <stdint.h>

#include "pf_rbpf.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pf_batch pf_batch;

typedef struct pf_batch_schedule {
    int32_t block;          /* threads per tile: 64, 128 or 256 */
    int32_t ppl;            /* particles per lane per tile: 1 or 2 (tile = block * ppl) */
    int32_t blocks_per_cu;  /* persistent grid = blocks_per_cu * CUs */
} pf_batch_schedule;

/* S sequences; sequence s has n_particles[s] particles and n_frames[s]
 * frames.  All share K components, state dimension d (12) and m measurement
 * rows.  cfg supplies the PDAF constants and cand_capacity (block_size and
 * n_particles are ignored). */
int pf_batch_create(pf_batch **out, int32_t S, const int32_t *n_particles, const int32_t *n_frames,
                    int32_t K, int32_t d, int32_t m, const pf_config *cfg);
int pf_batch_destroy(pf_batch *b);

/* The layout pf_batch_create would build for these sequences and a schedule
 * of block threads x ppl particles per tile, validated without a GPU: the
 * particle, block and frame offsets and the tile and weight-block tables
 * with their per-frame active counts are checked against the invariants the
 * kernels' indexing relies on (disjoint aligned ranges inside the buffers,
 * every tile and block inside its sequence, only sequences still running
 * active at a frame).  summary (optional, 6 values): particle slots (ld),
 * resampler blocks, tiles, weight blocks, largest per-sequence block count,
 * longest sequence.  pf_batch_create and pf_batch_set_schedule run the same
 * check.  Host only. */
int pf_batch_plan(int32_t S, const int32_t *n_particles, const int32_t *n_frames, int32_t block, int32_t ppl,
                  int64_t *summary);

/* my_gmm::loadGaussian x K for sequence s (arguments as pf_load_model). */
int pf_batch_load_model(pf_batch *b, int32_t s, int32_t npose, const int32_t *sel, const double *means,
                        const double *covs, const double *weights, const double *gamma,
                        const double *pca_proj, const double *pca_mean, const double *R);
/* PDAF joint kinds shared by all sequences (before pf_batch_reset). */
int pf_batch_set_joint_clutter(pf_batch *b, const int32_t *clutter);
/* my_gmm::resetTracker(bins) for sequence s. */
int pf_batch_reset(pf_batch *b, int32_t s, const int32_t *bins);

/* Runs every sequence to its last frame.  Device pointers, sequences
 * concatenated in index order with frame offsets F_s = sum_{r<s} n_frames[r]:
 * uniforms (F x 2), count (F x J), cand (F x J x cap x 2 floats), lml (F,
 * written: log mean weight of every sequence-frame).  flags bit 0: capture
 * the frame loop into a hipGraph.  Waits for the run; returns its status
 * (warning bits or PF_ERR_NOT_SPD as pf_step, over all sequences). */
int pf_batch_run_dev(pf_batch *b, const int32_t *count, const float *cand, const uint32_t *uniforms,
                     double *lml, int32_t flags);
/* Tracks in for sequence s (as pf_set_tracks: x N_s x d, P N_s x d x d,
 * model basis); the next run starts from them. */
int pf_batch_set_tracks(pf_batch *b, int32_t s, const double *x, const double *P);
/* The last frame of sequence s after a run: its log-weights (lw, N_s) and
 * the ancestors its resampling drew (anc, N_s); NULL to skip. */
int pf_batch_get_last_frame(pf_batch *b, int32_t s, float *lw, int32_t *anc);
/* Status of every sequence in the last run (S entries): 0, or the bits
 * PF_WARN_DEGENERATE (a uniform-fallback resampling), PF_WARN_ILLCOND (an
 * innovation factored in FP64) and 4 (an innovation covariance not positive
 * definite: pf_batch_run_dev then returned PF_ERR_NOT_SPD).  The run itself
 * returns the bits of all sequences together. */
int pf_batch_get_status(pf_batch *b, int32_t *status);
/* Tracks and estimate of sequence s after a run (as pf_get_tracks / pf_estimate). */
int pf_batch_get_tracks(pf_batch *b, int32_t s, double *x, double *P);
int pf_batch_estimate(pf_batch *b, int32_t s, double *xbar);

/* Schedule in use / set explicitly. */
int pf_batch_get_schedule(pf_batch *b, pf_batch_schedule *sch);
int pf_batch_set_schedule(pf_batch *b, const pf_batch_schedule *sch);
/* Times `frames` frames (from frame 0 of the inputs) for every candidate
 * schedule, keeps the fastest and reports its time and that of the default
 * schedule (256 threads, 1 particle per lane, 2 blocks per CU).  The tracks
 * are consumed: reset the sequences before the next run. */
int pf_batch_autotune(pf_batch *b, const int32_t *count, const float *cand, const uint32_t *uniforms,
                      double *lml, int32_t frames, pf_batch_schedule *best, double *best_ms,
                      double *default_ms);
/* Step kernel of the 24-dimensional state (d = 24 only; d = 12 always runs
 * one particle per lane):
 *   PF_BATCH_VALU       one particle per lane, the 324 state values in its
 *                       registers, FP32 vector ALU;
 *   PF_BATCH_MFMA       four lanes per particle (16 particles per wave), the
 *                       covariance as 21 upper-triangular 4 x 4 tiles whose
 *                       rank-1 updates run on the matrix cores
 *                       (v_mfma_f32_4x4x1f32: FP32 inputs, FP32 accumulation),
 *                       PDAF candidates split over the four lanes (default);
 *   PF_BATCH_QUAD_VALU  the MFMA variant's layout with the tile updates on
 *                       the vector ALU (for comparison).
 * All three run the algorithm of pf_step_candidates(); results agree to FP32
 * rounding (the order of the sums differs). */
#define PF_BATCH_VALU 0
#define PF_BATCH_MFMA 1
#define PF_BATCH_QUAD_VALU 2
int pf_batch_set_variant(pf_batch *b, int32_t variant);
int pf_batch_get_variant(pf_batch *b, int32_t *variant);
/* Device time of the last pf_batch_run_dev in ms. */
int pf_batch_last_run_ms(pf_batch *b, double *ms);

/* Log-marginal-likelihood and its gradient, per sequence.  No counterpart in
 * the reference (it has no learning); the objective is the sum over frames
 * of the log mean weight that ParticleFilter::update's weights define
 * (pf2DRao.cpp:138-148), with resampling (pf2DRao.cpp:152) a stop-gradient:
 * the run's ancestors and component draws are held fixed, and each
 * log-weight carries log pi_k - stopgrad(log pi_k), whose gradient is the
 * score of the component draws.  Runs the sequences forward as
 * pf_batch_run_dev (same inputs; lml written the same way) recording the
 * choices and checkpointing the state every checkpoint_every frames, then the
 * adjoint backward.  Gradients (host, FP64, model basis, NULL to skip), for
 * sequence s at offset s * (size of one sequence's block):
 *   grad_means  (S x K x d)      d F / d means (my_gmm.cpp:50 u)
 *   grad_logits (S x K)          d F / d logits, weights = softmax(logits)
 *   grad_rates  (S x K)          d F / d r, gamma = 1 - r (my_gmm.cpp:56)
 *   grad_chol   (S x K x d x d)  d F / d L, covs = L L^T, lower triangle (upper 0)
 * The tracks are consumed (reset before the next run). */
int pf_batch_lml_grad(pf_batch *b, const int32_t *count, const float *cand, const uint32_t *uniforms,
                      double *lml, int32_t checkpoint_every, double *grad_means, double *grad_logits,
                      double *grad_rates, double *grad_chol);
/* Component draws and ancestors of sequence s recorded by the last
 * pf_batch_lml_grad (host, n_frames[s] x n_particles[s] each, NULL to skip):
 * comp[t][j] is particle j's component at frame t (pf2DRao.cpp:128), anc[t][j]
 * its ancestor in frame t's resampling (pf2DRao.cpp:152). */
int pf_batch_get_choices(pf_batch *b, int32_t s, int32_t *comp, int32_t *anc);
/* Device times of the last pf_batch_lml_grad: recorded forward, backward (ms). */
int pf_batch_last_adjoint_ms(pf_batch *b, double *forward_ms, double *backward_ms);

#ifdef __cplusplus
}
#endif
#endif
