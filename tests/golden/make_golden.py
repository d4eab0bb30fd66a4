"""Writes tests/golden/golden_vectors.npz: known answers for the oracle's
primitives computed independently of it.

* mvn log-densities: scipy.stats.multivariate_normal.logpdf.
* GMM log-densities (pf2D prior): scipy.special.logsumexp over
  scipy.stats.multivariate_normal.logpdf of each component.
* Kalman updates: Gaussian conditioning of the joint [[P, P H^T], [H P, S]]
  in extended precision (numpy.longdouble), i.e. the textbook result without
  any Kalman-gain algebra.
* Systematic resampling: exact rational arithmetic (fractions.Fraction) of
  a_i = first index whose cumulative weight exceeds (i + u)/N of the total.

Run from the repository root: python tests/golden/make_golden.py
"""
from fractions import Fraction
import math
import os

import numpy as np
import scipy.special
import scipy.stats

HERE = os.path.dirname(os.path.abspath(__file__))


def spd(rng, n, scale=1.0):
    A = rng.standard_normal((n, n))
    return scale * (A @ A.T + n * np.eye(n))


def conditioning(x, P, H, BH, R, z):
    """Posterior of x | z for x ~ N(x, P), z = H x + BH + N(0, R), in long double."""
    L = np.longdouble
    x, P, H, BH, R, z = (np.asarray(a, L) for a in (x, P, H, BH, R, z))
    S = H @ P @ H.T + R
    C = P @ H.T
    # solve with long-double Gaussian elimination (numpy.linalg has no longdouble)
    m = S.shape[0]
    A = np.concatenate([S, np.concatenate([C.T, (z - H @ x - BH)[:, None]], axis=1)], axis=1)
    for c in range(m):
        p = c + int(np.argmax(np.abs(A[c:, c])))
        A[[c, p]] = A[[p, c]]
        A[c] /= A[c, c]
        for r in range(m):
            if r != c:
                A[r] -= A[r, c] * A[c]
    Sinv_CT = A[:, m:m + C.shape[0]]          # S^-1 C^T
    Sinv_y = A[:, -1]
    xp = x + C @ Sinv_y
    Pp = P - C @ Sinv_CT
    return xp.astype(np.float64), Pp.astype(np.float64)


def exact_systematic(w, N, U):
    tot = sum(Fraction(v) for v in w)
    cum, acc = [], Fraction(0)
    for v in w:
        acc += Fraction(v)
        cum.append(acc)
    u = Fraction(int(U), 2 ** 32)
    out = []
    idx = 0
    for i in range(N):
        t = (i + u) / N * tot
        while not cum[idx] > t:
            idx += 1
        out.append(idx)
    return np.array(out, np.int32)


def main():
    rng = np.random.default_rng(20240601)
    g = {}
    # --- multivariate normal log densities (m = 2, 6)
    for m in (2, 6):
        S = spd(rng, m, 30.0)
        u = rng.normal(300, 50, m)
        X = u + rng.standard_normal((20, m)) * 40
        g[f"mvn{m}_S"], g[f"mvn{m}_u"], g[f"mvn{m}_X"] = S, u, X
        g[f"mvn{m}_logpdf"] = scipy.stats.multivariate_normal(u, S).logpdf(X)
    # --- GMM prior of pf2D (d = 8, K = 5)
    K, d = 5, 8
    mus = rng.uniform(0, 500, (K, d))
    Ss = np.stack([spd(rng, d, 300.0) for _ in range(K)])
    pis = rng.dirichlet(np.ones(K))
    X = mus[rng.integers(0, K, 30)] + rng.standard_normal((30, d)) * 40
    comp = np.stack([scipy.stats.multivariate_normal(mus[k], Ss[k]).logpdf(X) + np.log(pis[k]) for k in range(K)])
    g.update(gmm_mu=mus, gmm_S=Ss, gmm_pi=pis, gmm_X=X, gmm_logpdf=scipy.special.logsumexp(comp, axis=0))
    # --- Kalman updates (d = 12, m = 6 and m = 2)
    for m in (6, 2):
        dd = 12
        P = spd(rng, dd, 2.0)
        x = rng.standard_normal(dd)
        H = rng.standard_normal((m, dd))
        BH = rng.standard_normal(m) * 10
        R = 100.0 * np.eye(m)
        z = H @ x + BH + rng.standard_normal(m) * 12
        xp, Pp = conditioning(x, P, H, BH, R, z)
        g.update({f"kf{m}_x": x, f"kf{m}_P": P, f"kf{m}_H": H, f"kf{m}_BH": BH, f"kf{m}_R": R,
                  f"kf{m}_z": z, f"kf{m}_xpost": xp, f"kf{m}_Ppost": Pp})
    # --- systematic resampling, exact rationals
    for (L, N) in ((7, 20), (40, 40), (100, 33)):
        w = rng.random(L) ** 2
        w[rng.random(L) < 0.3] = 0.0
        # weights with exact binary fractions so Fraction(w) equals the double
        w = np.round(w * 2 ** 20) / 2 ** 20
        w[int(rng.integers(0, L))] = 1.0       # max exactly 1: the fixed-point weights are exact
        U = int(rng.integers(0, 2 ** 32))
        g[f"sys{L}_{N}_w"] = w
        g[f"sys{L}_{N}_U"] = np.array([U], np.uint64)
        g[f"sys{L}_{N}_anc"] = exact_systematic(w, N, U)
    np.savez_compressed(os.path.join(HERE, "golden_vectors.npz"), **g)
    print("wrote", len(g), "arrays")


if __name__ == "__main__":
    main()
