// pf_adjoint_ref.cpp -- FP64 reference of the RBPF log-marginal-likelihood as
// a function of the GMM parameters with the discrete choices frozen, and its
// gradient by forward-mode dual numbers.
//
// TEST INFRASTRUCTURE ONLY (part of liboracle.so): the checker of the GPU
// adjoint.  Nothing in the product links against it.
//
// The function (the objective of the adjoint):
//   theta = [means (K x d) | logits (K) | rates (K) | chol (K x d x d, lower)]
//   pi = softmax(logits), gamma = 1 - rates, Sigma_k = L_k L_k^T.
//   tracks_0[j] = (mu_{bins_j}, Sigma_{bins_j})                   my_gmm.cpp:30-42
//   frame t: for particle j with component k = comp[t][j]:
//     predict x = g x + (1-g) mu_k, P = g^2 P + (1-g^2) Sigma_k   KF_model.cpp:11-15, my_gmm.cpp:51-58
//     joints: certain joints first, then cluttered ones (PDAF), exactly as
//     orc_rbpf_update_pdaf;  lw_tj = sum of the joint terms
//            + log pi_k(theta) - log pi0_k                          (score term, value 0)
//     lml_t = log((1/N) sum_j exp(lw_tj))
//     tracks_{t+1}[j] = updated_t[anc[t][j]]                         pf2DRao.cpp:152-156
//   F(theta) = sum_t lml_t.
// comp and anc are the choices a forward run made (they are not functions of
// theta here: resampling is a stop-gradient), and pi0 = softmax(logits) at the
// point where the gradient is taken, so F equals the run's log-marginal-
// likelihood there and d F / d logits is the score of the component draws.
// Gating (d^2 <= gate) is decided on values.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "pf_oracle.h"

namespace {

template <int NT>
struct Dual {
    double v = 0.0;
    double t[NT] = {};
    Dual() = default;
    Dual(double x) : v(x) {}
};

template <int NT> Dual<NT> operator+(const Dual<NT> &a, const Dual<NT> &b)
{ Dual<NT> r(a.v + b.v); for (int i = 0; i < NT; ++i) r.t[i] = a.t[i] + b.t[i]; return r; }
template <int NT> Dual<NT> operator-(const Dual<NT> &a, const Dual<NT> &b)
{ Dual<NT> r(a.v - b.v); for (int i = 0; i < NT; ++i) r.t[i] = a.t[i] - b.t[i]; return r; }
template <int NT> Dual<NT> operator-(const Dual<NT> &a)
{ Dual<NT> r(-a.v); for (int i = 0; i < NT; ++i) r.t[i] = -a.t[i]; return r; }
template <int NT> Dual<NT> operator*(const Dual<NT> &a, const Dual<NT> &b)
{ Dual<NT> r(a.v * b.v); for (int i = 0; i < NT; ++i) r.t[i] = a.t[i] * b.v + a.v * b.t[i]; return r; }
template <int NT> Dual<NT> operator/(const Dual<NT> &a, const Dual<NT> &b)
{ Dual<NT> r(a.v / b.v); for (int i = 0; i < NT; ++i) r.t[i] = (a.t[i] - r.v * b.t[i]) / b.v; return r; }
template <int NT> Dual<NT> &operator+=(Dual<NT> &a, const Dual<NT> &b) { a = a + b; return a; }
template <int NT> Dual<NT> &operator-=(Dual<NT> &a, const Dual<NT> &b) { a = a - b; return a; }
template <int NT> Dual<NT> exp(const Dual<NT> &a)
{ Dual<NT> r(std::exp(a.v)); for (int i = 0; i < NT; ++i) r.t[i] = r.v * a.t[i]; return r; }
template <int NT> Dual<NT> log(const Dual<NT> &a)
{ Dual<NT> r(std::log(a.v)); for (int i = 0; i < NT; ++i) r.t[i] = a.t[i] / a.v; return r; }
template <int NT> double val(const Dual<NT> &a) { return a.v; }

inline double val(double a) { return a; }
using std::exp;
using std::log;

template <class T>
struct Params {
    int K, d;
    std::vector<T> mu, logpi, gamma, Sigma;     // K x d, K, K, K x d x d
};

template <class T>
Params<T> derive(const orc_adj_problem *p, const T *theta)
{
    const int K = p->K, d = p->d;
    Params<T> q;
    q.K = K;
    q.d = d;
    q.mu.assign(theta, theta + (size_t)K * d);
    const T *lg = theta + (size_t)K * d;
    const T *rt = lg + K;
    const T *L = rt + K;
    double mx = -INFINITY;
    for (int k = 0; k < K; ++k) mx = std::fmax(mx, val(lg[k]));
    T s(0.0);
    for (int k = 0; k < K; ++k) s += exp(lg[k] - T(mx));
    const T lse = T(mx) + log(s);
    q.logpi.resize(K);
    q.gamma.resize(K);
    for (int k = 0; k < K; ++k) {
        q.logpi[k] = lg[k] - lse;
        q.gamma[k] = T(1.0) - rt[k];
    }
    q.Sigma.assign((size_t)K * d * d, T(0.0));
    for (int k = 0; k < K; ++k) {
        const T *Lk = L + (size_t)k * d * d;
        for (int i = 0; i < d; ++i)
            for (int j = 0; j < d; ++j) {
                T acc(0.0);
                for (int c = 0; c <= std::min(i, j); ++c) acc += Lk[i * d + c] * Lk[j * d + c];
                q.Sigma[((size_t)k * d + i) * d + j] = acc;
            }
    }
    return q;
}

// One 2-row joint on (x, P): returns the log-likelihood term.
template <class T>
T joint(const orc_adj_problem *p, int b, const double *cz, int cnt, bool clutter, std::vector<T> &x,
        std::vector<T> &P)
{
    const int d = p->d, m = p->m;
    const double *H = p->H + (size_t)(2 * b) * d;
    T y[2], HP[2][64], S[2][2];
    for (int r = 0; r < 2; ++r) {
        T s(0.0);
        for (int c = 0; c < d; ++c) s += T(H[r * d + c]) * x[c];
        y[r] = s + T(p->BH[2 * b + r]);
        for (int c = 0; c < d; ++c) {
            T a(0.0);
            for (int k = 0; k < d; ++k) a += T(H[r * d + k]) * P[(size_t)k * d + c];
            HP[r][c] = a;
        }
    }
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c) {
            T a(0.0);
            for (int k = 0; k < d; ++k) a += HP[r][k] * T(H[c * d + k]);
            S[r][c] = a + T(p->R[(size_t)(2 * b + r) * m + 2 * b + c]);
        }
    const T det = S[0][0] * S[1][1] - S[0][1] * S[1][0];
    T Si[2][2];
    Si[0][0] = S[1][1] / det;
    Si[1][1] = S[0][0] / det;
    Si[0][1] = -S[0][1] / det;
    Si[1][0] = -S[1][0] / det;
    T Kg[64][2];                                   // K = (H P)^T S^-1
    for (int i = 0; i < d; ++i)
        for (int c = 0; c < 2; ++c) Kg[i][c] = HP[0][i] * Si[0][c] + HP[1][i] * Si[1][c];
    const T lnorm = -T(std::log(2.0 * M_PI)) - T(0.5) * log(det);
    if (!clutter) {
        const T v0 = T(cz[0]) - y[0], v1 = T(cz[1]) - y[1];
        const T q = v0 * (Si[0][0] * v0 + Si[0][1] * v1) + v1 * (Si[1][0] * v0 + Si[1][1] * v1);
        for (int i = 0; i < d; ++i) x[i] += Kg[i][0] * v0 + Kg[i][1] * v1;
        std::vector<T> Pn(P.size());
        for (int i = 0; i < d; ++i)
            for (int j = 0; j < d; ++j)
                Pn[(size_t)i * d + j] = P[(size_t)i * d + j] - (Kg[i][0] * HP[0][j] + Kg[i][1] * HP[1][j]);
        P.swap(Pn);
        return lnorm - T(0.5) * q;
    }
    const double pg = 1.0 - std::exp(-0.5 * p->gate);
    const double b0 = 1.0 - p->pd * pg;
    std::vector<T> e(cnt), n0(cnt), n1(cnt);
    T l(b0);
    for (int c = 0; c < cnt; ++c) {
        n0[c] = T(cz[2 * c]) - y[0];
        n1[c] = T(cz[2 * c + 1]) - y[1];
        const T q = n0[c] * (Si[0][0] * n0[c] + Si[0][1] * n1[c]) + n1[c] * (Si[1][0] * n0[c] + Si[1][1] * n1[c]);
        e[c] = (val(q) <= p->gate) ? T(p->pd / p->lambda) * exp(lnorm - T(0.5) * q) : T(0.0);
        l += e[c];
    }
    T v[2] = {T(0.0), T(0.0)}, sp[2][2] = {{T(0.0), T(0.0)}, {T(0.0), T(0.0)}};
    for (int c = 0; c < cnt; ++c) {
        const T bc = e[c] / l;
        v[0] += bc * n0[c];
        v[1] += bc * n1[c];
        sp[0][0] += bc * n0[c] * n0[c];
        sp[0][1] += bc * n0[c] * n1[c];
        sp[1][1] += bc * n1[c] * n1[c];
    }
    sp[0][0] -= v[0] * v[0];
    sp[0][1] -= v[0] * v[1];
    sp[1][1] -= v[1] * v[1];
    sp[1][0] = sp[0][1];
    const T beta0 = T(b0) / l;
    for (int i = 0; i < d; ++i) x[i] += Kg[i][0] * v[0] + Kg[i][1] * v[1];
    // P+ = beta0 P + (1 - beta0)(P - K S K^T) + K spread K^T
    std::vector<T> Pn(P.size());
    for (int i = 0; i < d; ++i)
        for (int j = 0; j < d; ++j) {
            T ksk(0.0), ksp(0.0);
            for (int a = 0; a < 2; ++a)
                for (int c = 0; c < 2; ++c) {
                    ksk += Kg[i][a] * S[a][c] * Kg[j][c];
                    ksp += Kg[i][a] * sp[a][c] * Kg[j][c];
                }
            Pn[(size_t)i * d + j] = beta0 * P[(size_t)i * d + j] + (T(1.0) - beta0) * (P[(size_t)i * d + j] - ksk) + ksp;
        }
    P.swap(Pn);
    return log(l);
}

template <class T>
T replay(const orc_adj_problem *p, const Params<T> &q, const double *logpi0, double *lml_out)
{
    const int N = p->N, d = p->d, J = p->m / 2;
    std::vector<int> order;
    for (int b = 0; b < J; ++b) if (!p->clutter[b]) order.push_back(b);
    for (int b = 0; b < J; ++b) if (p->clutter[b]) order.push_back(b);
    std::vector<std::vector<T>> x(N), P(N), xu(N), Pu(N);
    for (int j = 0; j < N; ++j) {
        const int k = p->bins[j];
        x[j].assign(q.mu.begin() + (size_t)k * d, q.mu.begin() + (size_t)(k + 1) * d);
        P[j].assign(q.Sigma.begin() + (size_t)k * d * d, q.Sigma.begin() + (size_t)(k + 1) * d * d);
    }
    T F(0.0);
    std::vector<T> lw(N);
    for (int t = 0; t < p->T; ++t) {
        // particles are independent within a frame (the sums below are sequential)
#pragma omp parallel for schedule(static)
        for (int j = 0; j < N; ++j) {
            const int k = p->comp[(size_t)t * N + j];
            std::vector<T> xj = x[j], Pj = P[j];
            const T g = q.gamma[k], g2 = g * g;
            for (int i = 0; i < d; ++i) xj[i] = g * xj[i] + (T(1.0) - g) * q.mu[(size_t)k * d + i];
            for (int i = 0; i < d * d; ++i) Pj[i] = g2 * Pj[i] + (T(1.0) - g2) * q.Sigma[(size_t)k * d * d + i];
            T acc = q.logpi[k] - T(logpi0[k]);
            for (int b : order) {
                const int cnt = p->count[(size_t)t * J + b];
                const double *cz = p->cand + ((size_t)t * J + b) * p->cap * 2;
                const bool clut = p->clutter[b] != 0;
                if (!clut && cnt <= 0) continue;
                acc += joint<T>(p, b, cz, clut ? std::min(cnt, p->cap) : 1, clut, xj, Pj);
            }
            lw[j] = acc;
            xu[j].swap(xj);
            Pu[j].swap(Pj);
        }
        double mx = -INFINITY;
        for (int j = 0; j < N; ++j) mx = std::fmax(mx, val(lw[j]));
        T s(0.0);
        for (int j = 0; j < N; ++j) s += exp(lw[j] - T(mx));
        const T lml = T(mx) + log(s) - T(std::log((double)N));
        if (lml_out) lml_out[t] = val(lml);
        F += lml;
        for (int j = 0; j < N; ++j) {
            const int a = p->anc[(size_t)t * N + j];
            x[j] = xu[a];
            P[j] = Pu[a];
        }
    }
    return F;
}

int n_theta(const orc_adj_problem *p) { return p->K * p->d + 2 * p->K + p->K * p->d * p->d; }

int check(const orc_adj_problem *p)
{
    if (!p || p->K <= 0 || p->d <= 0 || p->d > 64 || p->m <= 0 || (p->m & 1) || p->N <= 0 || p->T <= 0 ||
        p->cap <= 0)
        return -1;
    return 0;
}

}  // namespace

extern "C" {

int orc_adj_n_theta(const orc_adj_problem *p) { return check(p) ? -1 : n_theta(p); }

int orc_adj_value(const orc_adj_problem *p, const double *theta, const double *logpi0, double *F, double *lml)
{
    if (check(p)) return -1;
    Params<double> q = derive<double>(p, theta);
    *F = replay<double>(p, q, logpi0, lml);
    return 0;
}

int orc_adj_dirderiv(const orc_adj_problem *p, const double *theta, int nd, const double *V, double *F, double *dd)
{
    constexpr int NT = 8;
    if (check(p) || nd <= 0 || nd > NT) return -1;
    const int n = n_theta(p);
    Params<double> q0 = derive<double>(p, theta);
    std::vector<double> logpi0(q0.logpi);
    std::vector<Dual<NT>> th(n);
    for (int i = 0; i < n; ++i) {
        th[i] = Dual<NT>(theta[i]);
        for (int c = 0; c < nd; ++c) th[i].t[c] = V[(size_t)c * n + i];
    }
    Params<Dual<NT>> q = derive<Dual<NT>>(p, th.data());
    const Dual<NT> f = replay<Dual<NT>>(p, q, logpi0.data(), nullptr);
    for (int c = 0; c < nd; ++c) dd[c] = f.t[c];
    *F = f.v;
    return 0;
}

int orc_adj_grad(const orc_adj_problem *p, const double *theta, double *F, double *grad)
{
    if (check(p)) return -1;
    constexpr int NT = 8;
    const int n = n_theta(p);
    Params<double> q0 = derive<double>(p, theta);
    std::vector<double> logpi0(q0.logpi);
    std::vector<Dual<NT>> th(n);
    for (int c0 = 0; c0 < n; c0 += NT) {
        for (int i = 0; i < n; ++i) {
            th[i] = Dual<NT>(theta[i]);
            if (i >= c0 && i < c0 + NT) th[i].t[i - c0] = 1.0;
        }
        Params<Dual<NT>> q = derive<Dual<NT>>(p, th.data());
        const Dual<NT> f = replay<Dual<NT>>(p, q, logpi0.data(), nullptr);
        for (int i = c0; i < std::min(n, c0 + NT); ++i) grad[i] = f.t[i - c0];
        if (c0 == 0) *F = f.v;
    }
    return 0;
}

}  // extern "C"
