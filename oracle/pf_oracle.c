/*
 * pf_oracle.c -- FP64 CPU restatement of the reference particle-filter hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker).  Loaded by tests/, the smoke
 * check and bench.py's cpu_baseline leg; never by the product.
 *
 * Reference files restated (paths relative to the reference repository root):
 *   src/KF_model.cpp:11-25, src/my_gmm.cpp:11-75, src/pf2DRao.cpp:23-210,
 *   src/pf2D.cpp:13-268, src/pfPose.cpp:61-71 (model load + init).
 */
#include "pf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_2PI 6.283185307179586476925286766559

static int g_threads = 1;
void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int orc_get_threads(void) { return g_threads; }

/* ================================================================ Philox */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void orc_philox_u32(uint64_t seed, uint32_t stream, uint64_t first, int64_t count, uint32_t *out)
{
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int64_t i = 0; i < count; ++i) {
        uint64_t w = first + (uint64_t)i;
        uint64_t blk = w >> 2;
        uint32_t ctr[4] = {(uint32_t)blk, (uint32_t)(blk >> 32), stream, 0u};
        uint32_t o[4];
        orc_philox4x32_10(ctr, key, o);
        out[i] = o[w & 3];
    }
}

/* ======================================================= deterministic exp */
/* Cody-Waite reduction x = n ln2 + r with the fdlibm split of ln2, then the
 * degree-12 Taylor polynomial of exp(r) on |r| <= ln2/2 (truncation error
 * < 2e-16 relative).  Only rint, fma and ldexp are used, each exactly
 * specified by IEEE-754, so every conforming build returns the same bits. */
double orc_det_exp(double x)
{
    if (!(x >= -745.0)) return 0.0;
    if (x > 709.0) x = 709.0;
    const double n = rint(x * 1.4426950408889634074);
    double r = fma(-n, 6.93147180369123816490e-01, x);
    r = fma(-n, 1.90821492927058770002e-10, r);
    double p = 1.0 / 479001600.0;
    p = fma(p, r, 1.0 / 39916800.0);
    p = fma(p, r, 1.0 / 3628800.0);
    p = fma(p, r, 1.0 / 362880.0);
    p = fma(p, r, 1.0 / 40320.0);
    p = fma(p, r, 1.0 / 5040.0);
    p = fma(p, r, 1.0 / 720.0);
    p = fma(p, r, 1.0 / 120.0);
    p = fma(p, r, 1.0 / 24.0);
    p = fma(p, r, 1.0 / 6.0);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(p, (int)n);
}

/* v in [0,1] -> floor(v * 2^32) as an integer, exactly. */
static uint64_t fixed32(double v)
{
    if (!(v > 0.0)) return 0;
    double t = floor(ldexp(v, 32));
    if (t >= 4294967296.0) return (uint64_t)1 << 32;
    return (uint64_t)(uint32_t)t;
}

/* ================================================== systematic resampling */
typedef unsigned __int128 u128;

/* a_i = min{idx : C[idx]*N*2^32 > (i*2^32+U)*T}; merge walk, O(N+L). */
static int resample_fixed(uint64_t *q, int L, int N, uint32_t U, int *out)
{
    int status = 0;
    uint64_t T = 0;
    for (int i = 0; i < L; ++i) T += q[i];
    if (T == 0) {                               /* Q3: degenerate -> uniform */
        for (int i = 0; i < L; ++i) q[i] = 1;
        T = (uint64_t)L;
        status = 1;
    }
    /* q becomes the inclusive prefix sum C. */
    uint64_t acc = 0;
    for (int i = 0; i < L; ++i) { acc += q[i]; q[i] = acc; }
    const u128 scale = (u128)(uint64_t)N << 32;
    int idx = 0;
    for (int i = 0; i < N; ++i) {
        const u128 t = (u128)((((uint64_t)i) << 32) | U) * T;
        while ((u128)q[idx] * scale <= t) ++idx;
        out[i] = idx;
    }
    return status;
}

int orc_resample_linear(const double *w, int L, int N, uint32_t U, int *out)
{
    if (L <= 0 || N <= 0 || L > (1 << 28) || N > (1 << 28)) return -1;
    uint64_t *q = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)L);
    double wmax = 0.0;
    for (int i = 0; i < L; ++i)
        if (isfinite(w[i]) && w[i] > wmax) wmax = w[i];
    for (int i = 0; i < L; ++i)
        q[i] = (wmax > 0.0 && isfinite(w[i]) && w[i] > 0.0) ? fixed32(w[i] / wmax) : 0;
    int st = resample_fixed(q, L, N, U, out);
    free(q);
    return st;
}

/* Linear weights with a known bound 1/scale (w scale <= 1): q_i =
 * floor(2^32 w_i scale), no maximum needed.  The pose tracker's proposal
 * weights L/Z (pfPose.cpp:300-301) are bounded by 1/(z_floor (1 - 2 Pa)). */
int orc_resample_scaled(const double *w, int L, int N, uint32_t U, double scale, int *out)
{
    if (L <= 0 || N <= 0 || L > (1 << 28) || N > (1 << 28) || !(scale > 0.0)) return -1;
    uint64_t *q = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)L);
    for (int i = 0; i < L; ++i)
        q[i] = (isfinite(w[i]) && w[i] > 0.0) ? fixed32(w[i] * scale) : 0;
    int st = resample_fixed(q, L, N, U, out);
    free(q);
    return st;
}

/* Block-scaled log-weights (blocks of ORC_WBLOCK consecutive inputs): the
 * largest finite lw of block b gives e_b = ceil(m_b log2 e) (clamped to
 * +-2^30), q_i = fixed32(det_exp(lw_i - e_b ln 2)); with E the largest e_b,
 * the cumulative weight at input k of block b is
 *   C_k = sum_{b' < b} (Q_b' >> s_b') + (cloc_k >> s_b),  s_b = E - e_b,
 * Q the block sums, cloc the block-local inclusive prefix of q, a shift of 64
 * or more giving 0 (empty blocks: no finite lw).  No positive total: the
 * uniform resampling of q_i = 1. */
#define ORC_WBLOCK 256
#define ORC_POW_EMPTY (-2147483647 - 1)
#define ORC_POW_LIMIT (1 << 30)

int orc_block_pow(double m)
{
    const double c = ceil(m * 1.4426950408889634);
    return c > (double)ORC_POW_LIMIT ? ORC_POW_LIMIT : (c < -(double)ORC_POW_LIMIT ? -ORC_POW_LIMIT : (int)c);
}

static double pow_ref(int e) { return (double)e * 0.6931471805599453; }

static uint64_t shr64(uint64_t x, int s) { return s >= 64 ? 0 : x >> s; }

int orc_resample_log(const double *lw, int L, int N, uint32_t U, int *out)
{
    if (L <= 0 || N <= 0 || L > (1 << 28) || N > (1 << 28)) return -1;
    const int nb = (L + ORC_WBLOCK - 1) / ORC_WBLOCK;
    uint64_t *C = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)L);
    int *e = (int *)malloc(sizeof(int) * (size_t)nb);
    int E = ORC_POW_EMPTY;
    for (int b = 0; b < nb; ++b) {
        const int i0 = b * ORC_WBLOCK, i1 = i0 + ORC_WBLOCK < L ? i0 + ORC_WBLOCK : L;
        double m = -INFINITY;
        for (int i = i0; i < i1; ++i)
            if (isfinite(lw[i]) && lw[i] > m) m = lw[i];
        e[b] = isfinite(m) ? orc_block_pow(m) : ORC_POW_EMPTY;
        if (e[b] != ORC_POW_EMPTY && (E == ORC_POW_EMPTY || e[b] > E)) E = e[b];
    }
    uint64_t pre = 0;
    for (int b = 0; b < nb; ++b) {
        const int i0 = b * ORC_WBLOCK, i1 = i0 + ORC_WBLOCK < L ? i0 + ORC_WBLOCK : L;
        const int s = (e[b] == ORC_POW_EMPTY) ? 64 : (((long long)E - e[b]) >= 64 ? 64 : E - e[b]);
        uint64_t loc = 0;
        for (int i = i0; i < i1; ++i) {
            if (e[b] != ORC_POW_EMPTY && isfinite(lw[i])) loc += fixed32(orc_det_exp(lw[i] - pow_ref(e[b])));
            C[i] = pre + shr64(loc, s);
        }
        pre += shr64(loc, s);
    }
    const uint64_t T = pre;
    int status = 0;
    if (T == 0) {                               /* Q3: degenerate -> uniform */
        for (int i = 0; i < L; ++i) C[i] = (uint64_t)i + 1;
        status = 1;
    }
    const uint64_t Tt = T == 0 ? (uint64_t)L : T;
    const u128 scale = (u128)(uint64_t)N << 32;
    int idx = 0;
    for (int i = 0; i < N; ++i) {
        const u128 t = (u128)((((uint64_t)i) << 32) | U) * Tt;
        while ((u128)C[idx] * scale <= t) ++idx;
        out[i] = idx;
    }
    free(C);
    free(e);
    return status;
}

/* pf2DRao.cpp:175-210 as written, with the RNG draws replaced by u / fallback. */
int orc_resample_wheel(const double *w, int L, int N, double u, const int *fallback, int *out)
{
    double mw = 0.0;                                   /* maxWeight, :161-172 */
    for (int i = 0; i < L; ++i)
        if (w[i] > mw) mw = w[i];
    if (mw == 0.0) {                                   /* :184-192 */
        for (int i = 0; i < N; ++i) out[i] = fallback ? fallback[i] : 0;
        return 1;
    }
    int idx = 0;                                       /* :195-208 */
    const double step = 1.0 / (double)N;
    double beta = u * step;
    for (int i = 0; i < N; ++i) {
        while (beta > w[idx]) {
            beta -= w[idx];
            idx = (idx + 1) % L;
        }
        beta += step;
        out[i] = idx;
    }
    return 0;
}

/* ================================================================ linalg */
/* In-place lower Cholesky of an m x m SPD matrix (row-major, lower part
 * written, upper part zeroed).  Returns 0 on success, -1 if not SPD. */
static int chol_lower(int m, double *A)
{
    for (int j = 0; j < m; ++j) {
        double s = A[j * m + j];
        for (int k = 0; k < j; ++k) s -= A[j * m + k] * A[j * m + k];
        if (!(s > 0.0)) return -1;
        const double ljj = sqrt(s);
        A[j * m + j] = ljj;
        for (int i = j + 1; i < m; ++i) {
            double t = A[i * m + j];
            for (int k = 0; k < j; ++k) t -= A[i * m + k] * A[j * m + k];
            A[i * m + j] = t / ljj;
        }
        for (int i = 0; i < j; ++i) A[i * m + j] = 0.0;
    }
    return 0;
}

/* Gauss-Jordan inverse with partial pivoting (the LU family that cv::Mat::inv
 * uses by default).  Returns -1 if singular. */
static int inv_gj(int m, const double *A, double *Ainv)
{
    double *a = (double *)malloc(sizeof(double) * (size_t)(m * m));
    memcpy(a, A, sizeof(double) * (size_t)(m * m));
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) Ainv[i * m + j] = (i == j) ? 1.0 : 0.0;
    for (int c = 0; c < m; ++c) {
        int p = c;
        for (int r = c + 1; r < m; ++r)
            if (fabs(a[r * m + c]) > fabs(a[p * m + c])) p = r;
        if (a[p * m + c] == 0.0) { free(a); return -1; }
        if (p != c)
            for (int j = 0; j < m; ++j) {
                double t = a[c * m + j]; a[c * m + j] = a[p * m + j]; a[p * m + j] = t;
                t = Ainv[c * m + j]; Ainv[c * m + j] = Ainv[p * m + j]; Ainv[p * m + j] = t;
            }
        const double piv = a[c * m + c];
        for (int j = 0; j < m; ++j) { a[c * m + j] /= piv; Ainv[c * m + j] /= piv; }
        for (int r = 0; r < m; ++r) {
            if (r == c) continue;
            const double f = a[r * m + c];
            if (f == 0.0) continue;
            for (int j = 0; j < m; ++j) {
                a[r * m + j] -= f * a[c * m + j];
                Ainv[r * m + j] -= f * Ainv[c * m + j];
            }
        }
    }
    free(a);
    return 0;
}

double orc_mvn_logpdf(int m, const double *x, const double *u, const double *S)
{
    double L[64 * 64];
    double v[64];
    if (m > 64) return NAN;
    memcpy(L, S, sizeof(double) * (size_t)(m * m));
    if (chol_lower(m, L) != 0) return -INFINITY;
    double q = 0.0, logdet = 0.0;
    for (int i = 0; i < m; ++i) {
        double t = x[i] - u[i];
        for (int k = 0; k < i; ++k) t -= L[i * m + k] * v[k];
        v[i] = t / L[i * m + i];
        q += v[i] * v[i];
        logdet += log(L[i * m + i]);
    }
    return -0.5 * q - logdet - 0.5 * (double)m * log(ORC_2PI);
}

/* ============================================================== KF_model */
void orc_kf_predict(int d, const double *F, const double *B, const double *Q, double *x, double *P)
{
    double xn[64], T[64 * 64];
    for (int i = 0; i < d; ++i) {                      /* state = F*state + B */
        double s = 0.0;
        for (int k = 0; k < d; ++k) s += F[i * d + k] * x[k];
        xn[i] = s + B[i];
    }
    for (int i = 0; i < d; ++i)                        /* T = F*cov */
        for (int j = 0; j < d; ++j) {
            double s = 0.0;
            for (int k = 0; k < d; ++k) s += F[i * d + k] * P[k * d + j];
            T[i * d + j] = s;
        }
    for (int i = 0; i < d; ++i)                        /* cov = T*F^T + Q */
        for (int j = 0; j < d; ++j) {
            double s = 0.0;
            for (int k = 0; k < d; ++k) s += T[i * d + k] * F[j * d + k];
            P[i * d + j] = s + Q[i * d + j];
        }
    memcpy(x, xn, sizeof(double) * (size_t)d);
}

int orc_kf_update(int d, int m, const double *H, const double *BH, const double *R,
                  const double *z, double *x, double *P)
{
    double y[64], HP[64 * 64], S[64 * 64], Si[64 * 64], PHt[64 * 64], K[64 * 64], T[64 * 64];
    for (int r = 0; r < m; ++r) {                      /* y = z - (H x + BH) */
        double s = 0.0;
        for (int k = 0; k < d; ++k) s += H[r * d + k] * x[k];
        y[r] = z[r] - (s + BH[r]);
    }
    for (int r = 0; r < m; ++r)                        /* HP = H*cov */
        for (int j = 0; j < d; ++j) {
            double s = 0.0;
            for (int k = 0; k < d; ++k) s += H[r * d + k] * P[k * d + j];
            HP[r * d + j] = s;
        }
    for (int r = 0; r < m; ++r)                        /* S = HP*H^T + R */
        for (int c = 0; c < m; ++c) {
            double s = 0.0;
            for (int k = 0; k < d; ++k) s += HP[r * d + k] * H[c * d + k];
            S[r * m + c] = s + R[r * m + c];
        }
    if (inv_gj(m, S, Si) != 0) return -1;              /* S.inv() */
    for (int i = 0; i < d; ++i)                        /* PH^T = cov*H^T */
        for (int c = 0; c < m; ++c) {
            double s = 0.0;
            for (int k = 0; k < d; ++k) s += P[i * d + k] * H[c * d + k];
            PHt[i * m + c] = s;
        }
    for (int i = 0; i < d; ++i)                        /* K = PH^T * S^-1 */
        for (int c = 0; c < m; ++c) {
            double s = 0.0;
            for (int k = 0; k < m; ++k) s += PHt[i * m + k] * Si[k * m + c];
            K[i * m + c] = s;
        }
    for (int i = 0; i < d; ++i) {                      /* state = state + K*y */
        double s = 0.0;
        for (int c = 0; c < m; ++c) s += K[i * m + c] * y[c];
        x[i] += s;
    }
    for (int i = 0; i < d; ++i)                        /* T = I - K*H */
        for (int j = 0; j < d; ++j) {
            double s = 0.0;
            for (int c = 0; c < m; ++c) s += K[i * m + c] * H[c * d + j];
            T[i * d + j] = (i == j ? 1.0 : 0.0) - s;
        }
    double Pn[64 * 64];
    for (int i = 0; i < d; ++i)                        /* cov = T*cov */
        for (int j = 0; j < d; ++j) {
            double s = 0.0;
            for (int k = 0; k < d; ++k) s += T[i * d + k] * P[k * d + j];
            Pn[i * d + j] = s;
        }
    memcpy(P, Pn, sizeof(double) * (size_t)(d * d));
    return 0;
}

/* ================================================================ my_gmm */
void orc_load_gaussian(int d, int npose, int m, const int *sel, const double *u, const double *s,
                       const double *pca_proj, const double *pca_mean, double g, const double *r,
                       double *F, double *B, double *Q, double *H, double *BH, double *Rout)
{
    for (int i = 0; i < d; ++i)
        for (int j = 0; j < d; ++j) {
            Q[i * d + j] = (1.0 - g * g) * s[i * d + j];    /* :53 */
            F[i * d + j] = (i == j) ? g : 0.0;              /* :56 */
        }
    for (int i = 0; i < d; ++i) B[i] = (1.0 - g) * u[i];    /* :58 */
    for (int a = 0; a < m; ++a) {                           /* :61-72 */
        const int c = sel[a];
        for (int j = 0; j < d; ++j) H[a * d + j] = pca_proj[j * npose + c];
        BH[a] = pca_mean[c];
    }
    memcpy(Rout, r, sizeof(double) * (size_t)(m * m));      /* :54 */
}

void orc_reset_tracker(const orc_model *mdl, int N, const int *bins, double *x, double *P)
{
    const int d = mdl->d;
    for (int j = 0; j < N; ++j) {
        const int b = bins[j];
        memcpy(x + (size_t)j * d, mdl->mu + (size_t)b * d, sizeof(double) * (size_t)d);
        memcpy(P + (size_t)j * d * d, mdl->Sigma + (size_t)b * d * d, sizeof(double) * (size_t)(d * d));
    }
}

/* Per-component F, B, Q (my_gmm.cpp:51-58). */
static void build_components(const orc_model *mdl, double *F, double *B, double *Q)
{
    const int d = mdl->d;
    for (int k = 0; k < mdl->K; ++k) {
        const double g = mdl->gamma[k];
        for (int i = 0; i < d; ++i) {
            B[k * d + i] = (1.0 - g) * mdl->mu[k * d + i];
            for (int j = 0; j < d; ++j) {
                F[(size_t)k * d * d + i * d + j] = (i == j) ? g : 0.0;
                Q[(size_t)k * d * d + i * d + j] = (1.0 - g * g) * mdl->Sigma[(size_t)k * d * d + i * d + j];
            }
        }
    }
}

static double lml_of(const double *lw, int N)
{
    double lmax = -INFINITY;
    for (int j = 0; j < N; ++j)
        if (isfinite(lw[j]) && lw[j] > lmax) lmax = lw[j];
    if (!isfinite(lmax)) return -INFINITY;
    double s = 0.0;
    for (int j = 0; j < N; ++j)
        if (isfinite(lw[j])) s += exp(lw[j] - lmax);
    return lmax + log(s) - log((double)N);
}

/* Shared tail of both update modes: resample on lw, gather independent copies. */
static int finish_step(int N, int d, double *x, double *P, const double *xu, const double *Pu,
                       const double *lw, uint32_t U2, int *anc, double *lml)
{
    const int st = orc_resample_log(lw, N, N, U2, anc);        /* pf2DRao.cpp:152 */
    for (int j = 0; j < N; ++j) {                              /* :153-156, spec Q2 */
        memcpy(x + (size_t)j * d, xu + (size_t)anc[j] * d, sizeof(double) * (size_t)d);
        memcpy(P + (size_t)j * d * d, Pu + (size_t)anc[j] * d * d, sizeof(double) * (size_t)(d * d));
    }
    if (lml) *lml = lml_of(lw, N);
    return st;
}

int orc_rbpf_update_ref(const orc_model *mdl, int N, double *x, double *P, const double *Z,
                        uint32_t U1, uint32_t U2, double *lw, int *anc, int *comp,
                        double *lml, double *xu, double *Pu)
{
    const int K = mdl->K, d = mdl->d, m = mdl->m;
    if (d > 64 || m > 64 || N <= 0) return -1;
    double *F = (double *)malloc(sizeof(double) * (size_t)(K * d * d));
    double *B = (double *)malloc(sizeof(double) * (size_t)(K * d));
    double *Q = (double *)malloc(sizeof(double) * (size_t)(K * d * d));
    build_components(mdl, F, B, Q);
    double *xw = xu ? xu : (double *)malloc(sizeof(double) * (size_t)N * d);
    double *Pw = Pu ? Pu : (double *)malloc(sizeof(double) * (size_t)N * d * d);
    memcpy(xw, x, sizeof(double) * (size_t)N * d);
    memcpy(Pw, P, sizeof(double) * (size_t)N * d * d);

    orc_resample_linear(mdl->pi, K, N, U1, comp);              /* pf2DRao.cpp:128 */

#pragma omp parallel for num_threads(g_threads) if (g_threads > 1) schedule(static)
    for (int j = 0; j < N; ++j) {                              /* :134-142 */
        const int k = comp[j];
        double *xj = xw + (size_t)j * d, *Pj = Pw + (size_t)j * d * d;
        orc_kf_predict(d, F + (size_t)k * d * d, B + (size_t)k * d, Q + (size_t)k * d * d, xj, Pj);
        double z[64], yh[64], S[64 * 64];
        for (int r = 0; r < m; ++r) z[r] = Z[(size_t)r * N + j];
        for (int r = 0; r < m; ++r) {                          /* H*x + BH */
            double s = 0.0;
            for (int c = 0; c < d; ++c) s += mdl->H[r * d + c] * xj[c];
            yh[r] = s + mdl->BH[r];
        }
        for (int r = 0; r < m; ++r)                            /* H*cov*H^T + R */
            for (int c = 0; c < m; ++c) {
                double s = 0.0;
                for (int a = 0; a < d; ++a) {
                    double t = 0.0;
                    for (int b = 0; b < d; ++b) t += mdl->H[r * d + b] * Pj[b * d + a];
                    s += t * mdl->H[c * d + a];
                }
                S[r * m + c] = s + mdl->R[r * m + c];
            }
        lw[j] = orc_mvn_logpdf(m, z, yh, S);                   /* :138 (spec Q1, Q3) */
        if (orc_kf_update(d, m, mdl->H, mdl->BH, mdl->R, z, xj, Pj) != 0) lw[j] = -INFINITY;
    }
    const int st = finish_step(N, d, x, P, xw, Pw, lw, U2, anc, lml);
    if (!xu) free(xw);
    if (!Pu) free(Pw);
    free(F); free(B); free(Q);
    return st;
}

/* One PDAF block on joint rows r0, r0+1 (textbook PDAF form). */
static double pdaf_block(int d, const double *H, const double *BH, const double *Rb,
                         const double *cand, int count, double pd, double gate, double lambda,
                         double *x, double *P)
{
    double yh[2], HP[2 * 64], S[4], Si[4], Kg[64 * 2];
    for (int r = 0; r < 2; ++r) {
        double s = 0.0;
        for (int c = 0; c < d; ++c) s += H[r * d + c] * x[c];
        yh[r] = s + BH[r];
        for (int j = 0; j < d; ++j) {
            double t = 0.0;
            for (int k = 0; k < d; ++k) t += H[r * d + k] * P[k * d + j];
            HP[r * d + j] = t;
        }
    }
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c) {
            double s = 0.0;
            for (int k = 0; k < d; ++k) s += HP[r * d + k] * H[c * d + k];
            S[r * 2 + c] = s + Rb[r * 2 + c];
        }
    const double det = S[0] * S[3] - S[1] * S[2];
    if (!(det > 0.0) || !(S[0] > 0.0)) return -INFINITY;
    Si[0] = S[3] / det; Si[3] = S[0] / det; Si[1] = -S[1] / det; Si[2] = -S[2] / det;
    for (int i = 0; i < d; ++i)                     /* K = P H^T S^-1 = HP^T S^-1 */
        for (int c = 0; c < 2; ++c)
            Kg[i * 2 + c] = HP[0 * d + i] * Si[0 * 2 + c] + HP[1 * d + i] * Si[1 * 2 + c];
    const double pg = 1.0 - exp(-0.5 * gate);
    const double b0 = 1.0 - pd * pg;
    double e[256], nu[256][2];
    double l = b0;
    for (int c = 0; c < count && c < 256; ++c) {
        nu[c][0] = cand[2 * c] - yh[0];
        nu[c][1] = cand[2 * c + 1] - yh[1];
        const double d2 = nu[c][0] * (Si[0] * nu[c][0] + Si[1] * nu[c][1]) +
                          nu[c][1] * (Si[2] * nu[c][0] + Si[3] * nu[c][1]);
        e[c] = (d2 <= gate) ? (pd / lambda) * exp(orc_mvn_logpdf(2, cand + 2 * c, yh, S)) : 0.0;
        l += e[c];
    }
    double v[2] = {0.0, 0.0}, sp[4] = {0.0, 0.0, 0.0, 0.0};
    for (int c = 0; c < count && c < 256; ++c) {
        const double bc = e[c] / l;
        v[0] += bc * nu[c][0];
        v[1] += bc * nu[c][1];
        sp[0] += bc * nu[c][0] * nu[c][0];
        sp[1] += bc * nu[c][0] * nu[c][1];
        sp[3] += bc * nu[c][1] * nu[c][1];
    }
    sp[0] -= v[0] * v[0];
    sp[1] -= v[0] * v[1];
    sp[3] -= v[1] * v[1];
    sp[2] = sp[1];
    const double beta0 = b0 / l;
    for (int i = 0; i < d; ++i) x[i] += Kg[i * 2] * v[0] + Kg[i * 2 + 1] * v[1];
    /* P+ = beta0 P + (1-beta0)(P - K S K^T) + K spread K^T */
    double KSKt, KspK;
    double Pn[64 * 64];
    for (int i = 0; i < d; ++i)
        for (int j = 0; j < d; ++j) {
            KSKt = 0.0; KspK = 0.0;
            for (int a = 0; a < 2; ++a)
                for (int b = 0; b < 2; ++b) {
                    KSKt += Kg[i * 2 + a] * S[a * 2 + b] * Kg[j * 2 + b];
                    KspK += Kg[i * 2 + a] * sp[a * 2 + b] * Kg[j * 2 + b];
                }
            Pn[i * d + j] = beta0 * P[i * d + j] + (1.0 - beta0) * (P[i * d + j] - KSKt) + KspK;
        }
    memcpy(P, Pn, sizeof(double) * (size_t)(d * d));
    return log(l);
}

int orc_rbpf_update_pdaf(const orc_model *mdl, int N, double *x, double *P, const orc_frame *fr,
                         uint32_t U1, uint32_t U2, double *lw, int *anc, int *comp,
                         double *lml, double *xu, double *Pu)
{
    const int K = mdl->K, d = mdl->d, m = mdl->m, J = fr->J;
    if (d > 64 || m != 2 * J || N <= 0 || fr->cap > 256) return -1;
    double *F = (double *)malloc(sizeof(double) * (size_t)(K * d * d));
    double *B = (double *)malloc(sizeof(double) * (size_t)(K * d));
    double *Q = (double *)malloc(sizeof(double) * (size_t)(K * d * d));
    build_components(mdl, F, B, Q);
    double *xw = xu ? xu : (double *)malloc(sizeof(double) * (size_t)N * d);
    double *Pw = Pu ? Pu : (double *)malloc(sizeof(double) * (size_t)N * d * d);
    memcpy(xw, x, sizeof(double) * (size_t)N * d);
    memcpy(Pw, P, sizeof(double) * (size_t)N * d * d);
    int order[32], no = 0;
    for (int b = 0; b < J; ++b) if (!fr->clutter[b]) order[no++] = b;
    for (int b = 0; b < J; ++b) if (fr->clutter[b]) order[no++] = b;

    orc_resample_linear(mdl->pi, K, N, U1, comp);

#pragma omp parallel for num_threads(g_threads) if (g_threads > 1) schedule(static)
    for (int j = 0; j < N; ++j) {
        const int k = comp[j];
        double *xj = xw + (size_t)j * d, *Pj = Pw + (size_t)j * d * d;
        orc_kf_predict(d, F + (size_t)k * d * d, B + (size_t)k * d, Q + (size_t)k * d * d, xj, Pj);
        double acc = 0.0;
        for (int oi = 0; oi < J; ++oi) {
            const int b = order[oi];
            const double *Hb = mdl->H + (size_t)(2 * b) * d;
            const double *BHb = mdl->BH + 2 * b;
            double Rb[4] = {mdl->R[(2 * b) * m + 2 * b], mdl->R[(2 * b) * m + 2 * b + 1],
                            mdl->R[(2 * b + 1) * m + 2 * b], mdl->R[(2 * b + 1) * m + 2 * b + 1]};
            const double *cb = fr->cand + (size_t)b * fr->cap * 2;
            if (!fr->clutter[b]) {
                if (fr->count[b] <= 0) continue;
                double yh[2], S[4];
                for (int r = 0; r < 2; ++r) {
                    double s = 0.0;
                    for (int c = 0; c < d; ++c) s += Hb[r * d + c] * xj[c];
                    yh[r] = s + BHb[r];
                }
                for (int r = 0; r < 2; ++r)
                    for (int c = 0; c < 2; ++c) {
                        double s = 0.0;
                        for (int a = 0; a < d; ++a) {
                            double t = 0.0;
                            for (int bb = 0; bb < d; ++bb) t += Hb[r * d + bb] * Pj[bb * d + a];
                            s += t * Hb[c * d + a];
                        }
                        S[r * 2 + c] = s + Rb[r * 2 + c];
                    }
                acc += orc_mvn_logpdf(2, cb, yh, S);
                if (orc_kf_update(d, 2, Hb, BHb, Rb, cb, xj, Pj) != 0) acc = -INFINITY;
            } else {
                acc += pdaf_block(d, Hb, BHb, Rb, cb, fr->count[b], fr->pd, fr->gate, fr->lambda, xj, Pj);
            }
        }
        lw[j] = acc;
    }
    const int st = finish_step(N, d, x, P, xw, Pw, lw, U2, anc, lml);
    if (!xu) free(xw);
    if (!Pu) free(Pw);
    free(F); free(B); free(Q);
    return st;
}

void orc_rbpf_estimate(int N, int d, const double *x, double *xbar)
{
    for (int i = 0; i < d; ++i) xbar[i] = 0.0;
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < d; ++i) xbar[i] = xbar[i] + 1.0 / (double)N * x[(size_t)j * d + i];
}

void orc_get_samples(int d, int npose, const double *Hp, const double *M, const double *xbar,
                     int n, double scale, const double *normals, double *out)
{
    double y[2];
    for (int r = 0; r < 2; ++r) {                      /* state = H*full_state + M */
        double s = 0.0;
        for (int c = 0; c < d; ++c) s += Hp[r * d + c] * xbar[c];
        y[r] = s + M[r];
    }
    (void)npose;
    const double sd = 0.8 * scale;                     /* C = 0.8*scale*I, a stddev (Q14) */
    for (int c = 0; c < n; ++c) {
        out[c] = y[0] + sd * normals[c];
        out[n + c] = y[1] + sd * normals[n + c];
    }
}

void orc_get_sample_prob(int d, int npose, const double *Hp, const double *M, const double *xbar,
                         int n, const double *in, double scale, double *w)
{
    double y[2];
    for (int r = 0; r < 2; ++r) {
        double s = 0.0;
        for (int c = 0; c < d; ++c) s += Hp[r * d + c] * xbar[c];
        y[r] = s + M[r];
    }
    (void)npose;
    const double var = 0.8 * scale;                    /* used as a variance here (Q14) */
    double S[4] = {var, 0.0, 0.0, var};
    for (int c = 0; c < n; ++c) {
        double xc[2] = {in[c], in[n + c]};
        w[c] = exp(orc_mvn_logpdf(2, xc, y, S));
    }
}

/* ================================================================== pf2D */
int orc_pf2d_load_gaussian(int d, const double *S, double *Sinv, double *cnorm)
{
    double L[64 * 64];
    if (d > 64) return -1;
    memcpy(L, S, sizeof(double) * (size_t)(d * d));
    if (chol_lower(d, L) != 0) return -1;
    double logdet = 0.0;
    for (int i = 0; i < d; ++i) logdet += 2.0 * log(L[i * d + i]);
    /* Sinv = L^-T L^-1, column by column */
    for (int c = 0; c < d; ++c) {
        double e[64], y[64];
        for (int i = 0; i < d; ++i) e[i] = (i == c) ? 1.0 : 0.0;
        for (int i = 0; i < d; ++i) {
            double t = e[i];
            for (int k = 0; k < i; ++k) t -= L[i * d + k] * y[k];
            y[i] = t / L[i * d + i];
        }
        for (int i = d - 1; i >= 0; --i) {
            double t = y[i];
            for (int k = i + 1; k < d; ++k) t -= L[k * d + i] * e[k];
            e[i] = t / L[i * d + i];
        }
        for (int i = 0; i < d; ++i) Sinv[i * d + c] = e[i];
    }
    *cnorm = 1.0 / (pow(ORC_2PI, d / 2.0) * sqrt(exp(logdet)));
    return 0;
}

static void pf2d_bounds(int col, int side, double *lo, double *hi)
{
    if (col == 6) {
        *lo = 640 / 2.0 * side + 1;
        *hi = 640 / 2.0 + 640 / 2.0 * side;
    } else {
        *lo = 1;
        *hi = ((col % 2) == 0) * 640 + (((col + 1) % 2) == 0) * 480;
    }
}

void orc_pf2d_init(int N, int d, int side, const double *unif, double *particles)
{
    for (int i = 0; i < N; ++i)
        for (int c = 0; c < d; ++c) {
            double lo, hi;
            pf2d_bounds(c, side, &lo, &hi);
            particles[(size_t)i * d + c] = lo + (hi - lo) * unif[(size_t)i * d + c];
        }
}

int orc_pf2d_update(const orc_pf2d_gmm *g, int N, double *particles, double *weights, double *lw,
                    int *anc, const double *meas, uint32_t U, const double *normals,
                    const double *reinit_unif, int side, int flags)
{
    const int d = g->d, K = g->K;
    if (d != 8 || N <= 0) return -1;
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1) schedule(static)
    for (int i = 0; i < N; ++i) {                                  /* pf2D.cpp:155-172 */
        const double *xi = particles + (size_t)i * d;
        double terms[256];
        double tmax = -INFINITY, prior_lin = 0.0;
        for (int k = 0; k < K; ++k) {
            double q = 0.0;
            for (int a = 0; a < d; ++a) {
                double t = 0.0;
                for (int b = 0; b < d; ++b) t += (xi[b] - g->mu[k * d + b]) * g->Sinv[(size_t)k * d * d + b * d + a];
                q += t * (xi[a] - g->mu[k * d + a]);
            }
            if (flags & 1) {                                       /* as-run Q10 */
                prior_lin += g->pi[k] * g->cnorm[k] * (double)expf((float)(-0.5 * q));
            } else {
                terms[k] = log(g->pi[k] * g->cnorm[k]) - 0.5 * q;
                if (terms[k] > tmax) tmax = terms[k];
            }
        }
        double lprior;
        if (flags & 1) {
            lprior = log(prior_lin);
        } else {
            double s = 0.0;
            for (int k = 0; k < K; ++k) s += exp(terms[k] - tmax);
            lprior = tmax + log(s);
        }
        /* eyemvnpdf(x[6:8] - z0, 15) * eyemvnpdf(x[0:2] - z1, 15)  (:124-128, :168) */
        const double u0 = xi[6] - meas[0], u1 = xi[7] - meas[1];
        const double v0 = xi[0] - meas[2], v1 = xi[1] - meas[3];
        const double lnorm = -log(ORC_2PI * 15.0);
        lw[i] = lprior + (lnorm - 0.5 * (u0 * u0 + u1 * u1) / 15.0) + (lnorm - 0.5 * (v0 * v0 + v1 * v1) / 15.0);
    }
    double lmax = -INFINITY;                                       /* :174-177 */
    for (int i = 0; i < N; ++i)
        if (isfinite(lw[i]) && lw[i] > lmax) lmax = lw[i];
    double s = 0.0;
    if (isfinite(lmax))
        for (int i = 0; i < N; ++i) s += isfinite(lw[i]) ? exp(lw[i] - lmax) : 0.0;
    for (int i = 0; i < N; ++i)
        weights[i] = (isfinite(lmax) && isfinite(lw[i])) ? exp(lw[i] - lmax) / s : 0.0;

    int st = orc_resample_log(lw, N, N, U, anc);                   /* :184 -> :225-268 */
    double *old = (double *)malloc(sizeof(double) * (size_t)N * d);
    memcpy(old, particles, sizeof(double) * (size_t)N * d);
    if (st == 1) {                                                 /* :232-250 */
        orc_pf2d_init(N, d, side, reinit_unif, particles);
        for (int i = 0; i < N; ++i) { weights[i] = 1.0 / (double)N; anc[i] = i; }
    } else {
        for (int i = 0; i < N; ++i)
            memcpy(particles + (size_t)i * d, old + (size_t)anc[i] * d, sizeof(double) * (size_t)d);
    }
    free(old);
    static const int col_of_draw[8] = {2, 3, 4, 5, 6, 7, 0, 1};     /* :96-98 */
    for (int i = 0; i < N; ++i)                                    /* predict :90-102 */
        for (int t = 0; t < 8; ++t)
            particles[(size_t)i * d + col_of_draw[t]] += 5.0 * normals[(size_t)i * 8 + t];
    return st;
}

void orc_pf2d_estimate(int N, int d, const double *particles, const double *weights, double *out)
{
    for (int c = 0; c < d; ++c) out[c] = 0.0;
    for (int i = 0; i < N; ++i)
        for (int c = 0; c < d; ++c) out[c] = out[c] + weights[i] * particles[(size_t)i * d + c];
}
