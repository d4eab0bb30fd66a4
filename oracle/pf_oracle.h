/*
 * pf_oracle.h -- FP64 CPU restatement of the reference particle-filter hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * implementation in mkfbodytracker_pdaf_amd/.  Only tests/, the smoke check in
 * __graft_entry__.py and the cpu_baseline leg of bench.py may load it.  Nothing
 * in the product links against it.
 *
 * Every function cites the reference file:line whose behaviour it restates
 * (paths relative to the reference repository root).  Where the reference's
 * literal arithmetic is version-dependent or unintended, the oracle implements
 * the specified algorithm and names the deviation as a quirk ID (Q1..Q15,
 * each stated where it is used).
 *
 * Layouts: all matrices row-major double.  Particle states are AoS:
 * x[N*d], P[N*d*d].  A measurement matrix Z (m x N) is row-major, i.e. the
 * memory layout of a continuous CV_64F cv::Mat: Z[r*N + j] = row r, particle j.
 */
#ifndef PF_ORACLE_H
#define PF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ RNG */
/* Philox4x32-10 block function (Salmon, Moraes, Dror, Shaw, SC'11). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Stream of 32-bit words: word i = philox(ctr={(i/4) lo, (i/4) hi, stream, 0},
 * key={seed lo, seed hi})[i%4], for i in [first, first+count). */
void orc_philox_u32(uint64_t seed, uint32_t stream, uint64_t first, int64_t count, uint32_t *out);

/* -------------------------------------------- deterministic exp, fixed point */
/* exp(x) for x <= 0 built only from IEEE-exact operations (rint, fma, ldexp),
 * so CPU and GPU builds agree bit for bit.  Returns 0 for x < -745 or NaN. */
double orc_det_exp(double x);

/* ---------------------------------------------------- systematic resampling */
/* Spec form of ParticleFilter::resample (pf2DRao.cpp:175-210, pf2D.cpp:225-268):
 *   q_i = floor(2^32 * w_i / max w)           (linear weights)
 *   q_i = floor(2^32 * det_exp(lw_i - max lw)) (log weights)
 *   C = inclusive prefix sum of q (uint64, exact), T = C[L-1]
 *   a_i = min{ idx : C[idx] * N * 2^32 > (i*2^32 + U) * T }     i = 0..N-1
 * i.e. a_i = first index whose cumulative weight exceeds (i + U/2^32)/N of the
 * total (the wheel's stopping rule, pf2DRao.cpp:200-205).  Non-finite or
 * non-positive weights count as 0 and are never selected.  If no weight is
 * positive every q_i = 1 (uniform fallback, Q3) and the return value is 1.
 * Returns 0 on success, 1 on the degenerate fallback, -1 on bad arguments. */
int orc_resample_linear(const double *w, int L, int N, uint32_t U, int *out);
int orc_block_pow(double m);
int orc_resample_scaled(const double *w, int L, int N, uint32_t U, double scale, int *out);
int orc_resample_log(const double *lw, int L, int N, uint32_t U, int *out);
/* Literal FP64 wheel of pf2DRao.cpp:193-208 with beta0 = u/N (quirk mode Q4):
 * the running subtraction, the wrap idx=(idx+1)%L past the FP total, and the
 * maxWeight()==0 branch (pf2DRao.cpp:184-192) replaced by idx_i = fallback[i].
 * Characterisation only. */
int orc_resample_wheel(const double *w, int L, int N, double u, const int *fallback, int *out);

/* ------------------------------------------------------------- KF_model */
/* KF_model::predict (KF_model.cpp:11-15): x = F x + B ; P = F P F^T + Q.  */
void orc_kf_predict(int d, const double *F, const double *B, const double *Q, double *x, double *P);
/* KF_model::update (KF_model.cpp:17-25), literal form:
 *   y = z - (H x + BH); S = H P H^T + R; K = P H^T S^-1 (LU inverse);
 *   x = x + K y; P = (I - K H) P.                                        */
int orc_kf_update(int d, int m, const double *H, const double *BH, const double *R,
                  const double *z, double *x, double *P);
/* Textbook log N(x; u, S) through a Cholesky factor (spec for Q1).
 * Returns -inf if S is not positive definite. */
double orc_mvn_logpdf(int m, const double *x, const double *u, const double *S);

/* ------------------------------------------------------- my_gmm (RBPF) */
/* my_gmm::loadGaussian (my_gmm.cpp:45-75) for one component:
 *   Q = (1-g^2) s ; F = g I ; B = (1-g) u^T ; R = r ;
 *   H = H1 * pca_proj^T ; BH = H1 * pca_mean^T, with H1 the m x npose
 *   selection matrix whose row r has a 1 in column sel[r] (my_gmm.cpp:61-67).
 * pca_proj is d x npose, pca_mean is npose.  Outputs: F,Q d x d; B d;
 * H m x d; BH m; Rout m x m (copy of r). */
void orc_load_gaussian(int d, int npose, int m, const int *sel, const double *u, const double *s,
                       const double *pca_proj, const double *pca_mean, double g, const double *r,
                       double *F, double *B, double *Q, double *H, double *BH, double *Rout);

typedef struct orc_model {
    int K, d, m;
    const double *pi;     /* K      component weights (gmm.weight) */
    const double *gamma;  /* K      mean-reversion coefficient g   */
    const double *mu;     /* K*d    component means                 */
    const double *Sigma;  /* K*d*d  component covariances           */
    const double *H;      /* m*d    shared measurement matrix       */
    const double *BH;     /* m      shared measurement offset       */
    const double *R;      /* m*m    measurement noise               */
} orc_model;

/* my_gmm::resetTracker (my_gmm.cpp:30-42): x_j = mu_{b_j}, P_j = Sigma_{b_j}. */
void orc_reset_tracker(const orc_model *mdl, int N, const int *bins, double *x, double *P);

/* ParticleFilter::update (pf2DRao.cpp:125-158), reference measurement mode:
 * column j of Z (m x N) is particle j's measurement.  Per frame:
 *   comp = resample(pi, N, U1)                                  (:128)
 *   per particle: KF predict with F=g I (:137), lw_j = log N(z_j; Hx+BH, HPH^T+R)
 *     (:138, spec Q1/Q3: log domain, true Cholesky), KF update (:140)
 *   anc = resample_log(lw, N, U2)                               (:152)
 *   tracks[j] = updated[anc_j] as independent copies (:155, spec Q2)
 * On return x,P hold the resampled tracks; xu,Pu (optional, may be NULL) hold
 * the updated states before resampling; lw, anc, comp are outputs;
 * *lml receives log((1/N) sum_j exp(lw_j)) in FP64.
 * Returns the resampler status (0, or 1 for the degenerate fallback). */
int orc_rbpf_update_ref(const orc_model *mdl, int N, double *x, double *P, const double *Z,
                        uint32_t U1, uint32_t U2, double *lw, int *anc, int *comp,
                        double *lml, double *xu, double *Pu);

/* Candidate (PDAF) measurement mode.  Measurement rows are J = m/2 joints of
 * two rows (u, v).  Joint b is either certain (clutter[b]==0: count[b] is 0 or
 * 1, the single detection is used as is) or cluttered (clutter[b]!=0: up to
 * cap detections, PDAF).  Detections cand[(b*cap + c)*2 + {0,1}] are shared by
 * every particle.  Per particle, after the predict, joints are processed as
 * sequential 2-row Kalman blocks: certain joints first, then cluttered
 * joints, each in index order (R must be block diagonal in 2x2 joint blocks).
 * Certain block: lw += log N(z; y, S); standard update.
 * Cluttered block (PDAF, Bar-Shalom & Fortmann 1988):
 *   e_c = (pd/lambda) N(z_c; y, S) if d2_c <= gate else 0 ;
 *   l = (1 - pd*pg) + sum e_c ; lw += log l ;
 *   beta_c = e_c/l ; beta_0 = (1 - pd*pg)/l ;  nu = sum beta_c nu_c ;
 *   x += K nu ; P = P - (1-beta_0) K S K^T + K (sum beta_c nu_c nu_c^T - nu nu^T) K^T
 *   with pg = 1 - exp(-gate/2) (chi-square, 2 dof).
 * Component sampling and resampling as in orc_rbpf_update_ref. */
typedef struct orc_frame {
    int J, cap;
    const int *clutter;    /* J */
    const int *count;      /* J */
    const double *cand;    /* J*cap*2 */
    double pd, gate, lambda;
} orc_frame;
int orc_rbpf_update_pdaf(const orc_model *mdl, int N, double *x, double *P, const orc_frame *fr,
                         uint32_t U1, uint32_t U2, double *lw, int *anc, int *comp,
                         double *lml, double *xu, double *Pu);

/* ParticleFilter::getEstimator (pf2DRao.cpp:23-31): xbar = sum_j x_j / N,
 * accumulated sequentially in j order (Q8). */
void orc_rbpf_estimate(int N, int d, const double *x, double *xbar);

/* ParticleFilter::getSamples (pf2DRao.cpp:85-103): y = Hp xbar + M (Hp is
 * npose x d = pca_proj^T, M npose); sample c = y[0:2] + 0.8*scale*n_c with
 * host normals n (2 x n, row-major like the returned 2 x n matrix; Q14: the
 * 0.8*scale matrix is used as a standard deviation).  out: 2 x n row-major. */
void orc_get_samples(int d, int npose, const double *Hp, const double *M, const double *xbar,
                     int n, double scale, const double *normals, double *out);
/* ParticleFilter::getSampleProb (pf2DRao.cpp:105-122): densities of the
 * columns of in (2 x n) under N(y[0:2], 0.8*scale*I) (variance, Q14). */
void orc_get_sample_prob(int d, int npose, const double *Hp, const double *M, const double *xbar,
                         int n, const double *in, double scale, double *w);

/* ------------------------------------------------------------- pf2D */
/* pf2D my_gmm::loadGaussian (pf2D.cpp:28-37): Sinv = S^-1 (Cholesky inverse),
 * cnorm = 1/((2 pi)^(d/2) sqrt(det S)).  Returns -1 if S is not SPD. */
int orc_pf2d_load_gaussian(int d, const double *S, double *Sinv, double *cnorm);
/* ParticleFilter(N, d, side) (pf2D.cpp:44-71): column i of the N x d particle
 * matrix ~ U[lo_i, hi_i) from host uniforms unif (N*d, row-major):
 * column 6: [320*side + 1, 320 + 320*side); even columns [1, 640); odd [1, 480). */
void orc_pf2d_init(int N, int d, int side, const double *unif, double *particles);
typedef struct orc_pf2d_gmm {
    int K, d;
    const double *pi, *mu, *Sinv, *cnorm;   /* K, K*d, K*d*d, K */
} orc_pf2d_gmm;
/* ParticleFilter::update (pf2D.cpp:148-210), spec form:
 *   lw_i = logsumexp_k [log(pi_k cnorm_k) - q_ik/2]  (Q10 spec; flag bit 0 = as-run expf)
 *        + log N2(x_i[6:8]; z_0, 15 I) + log N2(x_i[0:2]; z_1, 15 I)   (:168, :124-128)
 *   weights = normalised exp(lw) (:174-177)
 *   resample (:184 -> :225-268) with the spec resampler on lw and uniform U;
 *   all weights zero -> particles re-drawn from reinit_unif (N*d) as the
 *   constructor does, weights = 1/N (:232-250)
 *   predict (:205 -> :90-102): x_i += 5 n, normals consumed per particle for
 *   columns 2,3,4,5, then 6,7, then 0,1 (normals N*8 in draw order).
 * weights (N) keep the pre-resample normalised weights (Q9).
 * meas is the 2 x 2 measurement matrix (row 0 vs columns 6:8, row 1 vs 0:2).
 * Returns the resampler status. */
int orc_pf2d_update(const orc_pf2d_gmm *g, int N, double *particles, double *weights, double *lw,
                    int *anc, const double *meas, uint32_t U, const double *normals,
                    const double *reinit_unif, int side, int flags);
/* ParticleFilter::getEstimator (pf2D.cpp:79-87): sum_i weights_i particles_i. */
void orc_pf2d_estimate(int N, int d, const double *particles, const double *weights, double *out);

/* ------------------------------------------------- adjoint reference */
/* The log-marginal-likelihood of a recorded candidate-mode run as a function
 * of theta = [means K*d | logits K | rates K | chol K*d*d (lower used)], with
 * the run's component draws comp[T][N] and ancestors anc[T][N] frozen
 * (resampling as a stop-gradient; anc[t] is the resampling at the end of
 * frame t), initial components bins[N] (my_gmm.cpp:30-42) and the frames'
 * detections.  Each particle's log-weight carries log pi_k(theta) - log pi0_k
 * (value 0 at pi0), so d/d logits is the score of the component draws.
 * Model basis, FP64 (pf_adjoint_ref.cpp). */
typedef struct orc_adj_problem {
    int K, d, m, N, T, cap;
    const double *H, *BH, *R;      /* m*d, m, m*m */
    const int *clutter;            /* J = m/2 */
    const int *count;              /* T*J */
    const double *cand;            /* T*J*cap*2 */
    double pd, gate, lambda;
    const int *bins;               /* N */
    const int *comp, *anc;         /* T*N each */
} orc_adj_problem;
int orc_adj_n_theta(const orc_adj_problem *p);
/* F(theta) with the score term taken against logpi0 (K); lml (T, optional)
 * receives the per-frame log mean weights. */
int orc_adj_value(const orc_adj_problem *p, const double *theta, const double *logpi0, double *F, double *lml);
/* F and dF/dtheta at theta (pi0 = softmax(logits of theta)), by forward-mode
 * dual numbers, 8 directions per pass. */
int orc_adj_grad(const orc_adj_problem *p, const double *theta, double *F, double *grad);
/* Directional derivatives of F at theta along nd <= 8 directions (rows of V,
 * nd x n_theta), in one forward-mode pass: dd[c] = grad F . V[c]. */
int orc_adj_dirderiv(const orc_adj_problem *p, const double *theta, int nd, const double *V, double *F, double *dd);

/* ------------------------------------------------- pose tracker stage */
/* (pf_pose_ref.c) IEEE-exact-operation log, sin/cos of 2 pi w 2^-32, and the
 * Box-Muller pair (u1 = (w0 + 1) 2^-32, angle from w1). */
double orc_det_log(double x);
void orc_det_sincos_turn(uint32_t w, double *sn, double *cs);
void orc_normal_pair(uint32_t w0, uint32_t w1, double *z0, double *z1);
/* GaussianBlur taps and the separable reflect-101 blur (pfPose.cpp:213). */
void orc_gauss_taps(int ksize, double sigma, float *taps);
int orc_blur_u8(const uint8_t *src, int rows, int cols, int ksize, const float *taps, float *tmp, uint8_t *dst);
/* Proposals of one arm for global frame g (pfPose.cpp:218-235). */
void orc_tracker_proposals(uint64_t seed, int arm, int64_t g, int n_c, int uniform, const int roi[4], int rows,
                           int cols, const double y[2], double spread, double *out);
/* Association weights (pfPose.cpp:244-293), unnormalised. */
void orc_tracker_weights(int n_c, const double *props, const double *p_own, const double *p_other, const uint8_t *img,
                         int rows, int cols, double pa, double z_floor, double *w);
/* 6 x N measurement from resampled proposals (pfPose.cpp:303-320). */
void orc_tracker_measurement(int N, const int *bins, int n_c, const double *props, const int roi[4], double *Z);
/* rpy and get3Dpose (pfPose.cpp:84-127). */
void orc_rpy(double roll, double pitch, double yaw, double R[9]);
int orc_get3Dpose(const double *e, double fx, double fy, double cx, double cy, double out[15]);
/* publishTFtree joints relative to Head (extractTFdata layout) and the
 * publish2Dpos points (pfPose.cpp:129-208). */
void orc_tf_rows(const double p1[15], const double p2[15], double out[24]);
void orc_pose2d(const double *e1, const double *e2, double out[16]);

/* ------------------------------------------------------------- misc */
/* Number of OpenMP threads used by the particle loops (1 = sequential). */
void orc_set_threads(int n);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
