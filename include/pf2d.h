/*
 * pf2d.h -- C ABI of the plain particle filter pf2D (src/pf2D.h:12-51,
 * src/pf2D.cpp) on gfx950.
 *
 * One handle = one pf2D ParticleFilter(numParticles, numDims, side): N
 * particles of an 8-dim pixel state (FP32 structure-of-arrays in HBM), a
 * static GMM prior (my_gmm of src/pf2D.h:12-23) and the previous update's
 * normalised weights.  Randomness is host-supplied: uniforms for the
 * constructor and the degenerate re-initialisation (cv::randu,
 * pf2D.cpp:64-68,232-250), normals for the random-walk predict (cv::randn,
 * pf2D.cpp:96-98) and one 32-bit uniform per systematic resampling
 * (rand(), pf2D.cpp:255).  Status codes and pf_last_error() are those of
 * pf_rbpf.h.
 */
#ifndef PF2D_H
#define PF2D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pf2d_handle pf2d_handle;

/* ParticleFilter(int numParticles, int numDims, bool side1) (pf2D.cpp:44-71)
 * without the random initialisation (see pf2d_init).  numDims must be 8: the
 * reference's predict draws noise for exactly eight columns (pf2D.cpp:96-98). */
int pf2d_create(pf2d_handle **out, int32_t n_particles, int32_t n_dims, int32_t side, int32_t device);
int pf2d_destroy(pf2d_handle *h);

/* The K calls of my_gmm::loadGaussian(mean, sigma, weight) (pf2D.cpp:28-37):
 * means K x d, covs K x d x d, weights K.  Stores mu_k, Sigma_k^-1 (Cholesky
 * inverse) and 1 / ((2 pi)^(d/2) sqrt(det Sigma_k)). */
int pf2d_load_gmm(pf2d_handle *h, int32_t K, const double *means, const double *covs, const double *weights);

/* The same GMM in the members ParticleFilter::update reads (pf2D.h:18-21,
 * pf2D.cpp:159-163): means K x d, sigma_i K x d x d (inverse covariances;
 * symmetrised), det_s K (1 / ((2 pi)^(d/2) sqrt(det Sigma))), weights K.
 * Component k contributes weights[k] det_s[k] exp(-(x - mu_k) sigma_i_k
 * (x - mu_k)^T / 2) to the prior; weights[k] det_s[k] must be finite and
 * positive.  pf2d_load_gmm computes sigma_i and det_s from covs and calls this. */
int pf2d_load_gmm_precision(pf2d_handle *h, int32_t K, const double *means, const double *sigma_i,
                            const double *det_s, const double *weights);

/* Constructor randomisation (pf2D.cpp:58-70): column c of particle i is
 * lo_c + (hi_c - lo_c) * unif[i * d + c], with column 6 restricted to the
 * side's half of the 640-pixel image; weights 1/N. */
int pf2d_init(pf2d_handle *h, const double *unif);

/* Particles (N x d) and weights (N) in and out. */
int pf2d_set_particles(pf2d_handle *h, const double *particles, const double *weights);
int pf2d_get_particles(pf2d_handle *h, double *particles, double *weights);

/* ParticleFilter::update(measurement) (pf2D.cpp:148-210): meas is the 2 x 2
 * measurement matrix (row 0 vs columns 6:8, row 1 vs columns 0:2).  Log
 * weights = log GMM prior (logsumexp over K) + two isotropic variance-15
 * Gaussians; weights normalised and kept (the estimator uses them, as in the
 * reference); systematic resampling with uniform u; all-zero weights
 * re-draw the particles from reinit_unif (N x d) and set weights 1/N; then
 * the random walk x += 5 n with normals (N x 8) consumed per particle for
 * columns 2..5, 6, 7, 0, 1.  Optional outputs: lw (N), anc (N).  Returns 1
 * on the degenerate re-initialisation. */
int pf2d_update(pf2d_handle *h, const double *meas, uint32_t u, const double *normals,
                const double *reinit_unif, float *lw, int32_t *anc);

/* ParticleFilter::predict() (pf2D.cpp:90-102) on its own: the random walk
 * x += 5 n on the current particles, normals (N x 8) consumed per particle
 * for columns 2..5, 6, 7, 0, 1 as in pf2d_update; the weights are kept. */
int pf2d_predict(pf2d_handle *h, const double *normals);

/* ParticleFilter::getEstimator (pf2D.cpp:79-87): sum_i weights_i particles_i. */
int pf2d_estimate(pf2d_handle *h, double *out);

#ifdef __cplusplus
}
#endif
#endif
