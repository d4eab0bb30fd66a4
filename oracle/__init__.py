"""ctypes bindings for the FP64 CPU oracle (``oracle/liboracle.so``).

TEST INFRASTRUCTURE ONLY.  The oracle is the parity checker for the HIP
implementation; only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this module.

Every wrapper mirrors one C function of ``oracle/pf_oracle.h``; the C file
cites the reference lines each function restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")


def build() -> None:
    """Compile liboracle.so with the Makefile beside this file."""
    subprocess.run(["make", "-s", "-B", "-C", _HERE, "liboracle.so"], check=True)


class _Model(C.Structure):
    _fields_ = [("K", C.c_int), ("d", C.c_int), ("m", C.c_int),
                ("pi", C.c_void_p), ("gamma", C.c_void_p), ("mu", C.c_void_p),
                ("Sigma", C.c_void_p), ("H", C.c_void_p), ("BH", C.c_void_p),
                ("R", C.c_void_p)]


class _Frame(C.Structure):
    _fields_ = [("J", C.c_int), ("cap", C.c_int), ("clutter", C.c_void_p),
                ("count", C.c_void_p), ("cand", C.c_void_p),
                ("pd", C.c_double), ("gate", C.c_double), ("lam", C.c_double)]


class _Pf2dGmm(C.Structure):
    _fields_ = [("K", C.c_int), ("d", C.c_int), ("pi", C.c_void_p), ("mu", C.c_void_p),
                ("Sinv", C.c_void_p), ("cnorm", C.c_void_p)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_philox4x32_10.argtypes = [_up, _up, _up]
        L.orc_adj_n_theta.argtypes = [C.c_void_p]
        L.orc_adj_value.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_adj_grad.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_adj_dirderiv.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_philox_u32.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, C.c_int64, _up]
        L.orc_det_exp.argtypes = [C.c_double]
        L.orc_det_exp.restype = C.c_double
        L.orc_resample_linear.argtypes = [_dp, C.c_int, C.c_int, C.c_uint32, _ip]
        L.orc_resample_log.argtypes = [_dp, C.c_int, C.c_int, C.c_uint32, _ip]
        L.orc_resample_scaled.argtypes = [_dp, C.c_int, C.c_int, C.c_uint32, C.c_double, _ip]
        L.orc_block_pow.argtypes = [C.c_double]
        L.orc_block_pow.restype = C.c_int
        L.orc_resample_wheel.argtypes = [_dp, C.c_int, C.c_int, C.c_double, C.c_void_p, _ip]
        L.orc_kf_predict.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.orc_kf_update.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_mvn_logpdf.argtypes = [C.c_int, _dp, _dp, _dp]
        L.orc_mvn_logpdf.restype = C.c_double
        L.orc_load_gaussian.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _dp, _dp, _dp, _dp,
                                        C.c_double, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_reset_tracker.argtypes = [C.POINTER(_Model), C.c_int, _ip, _dp, _dp]
        L.orc_rbpf_update_ref.argtypes = [C.POINTER(_Model), C.c_int, _dp, _dp, _dp,
                                          C.c_uint32, C.c_uint32, _dp, _ip, _ip,
                                          C.POINTER(C.c_double), C.c_void_p, C.c_void_p]
        L.orc_rbpf_update_pdaf.argtypes = [C.POINTER(_Model), C.c_int, _dp, _dp, C.POINTER(_Frame),
                                           C.c_uint32, C.c_uint32, _dp, _ip, _ip,
                                           C.POINTER(C.c_double), C.c_void_p, C.c_void_p]
        L.orc_rbpf_estimate.argtypes = [C.c_int, C.c_int, _dp, _dp]
        L.orc_get_samples.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, C.c_int, C.c_double, _dp, _dp]
        L.orc_get_sample_prob.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, C.c_int, _dp, C.c_double, _dp]
        L.orc_pf2d_load_gaussian.argtypes = [C.c_int, _dp, _dp, C.POINTER(C.c_double)]
        L.orc_pf2d_init.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.orc_pf2d_update.argtypes = [C.POINTER(_Pf2dGmm), C.c_int, _dp, _dp, _dp, _ip, _dp,
                                      C.c_uint32, _dp, _dp, C.c_int, C.c_int]
        L.orc_pf2d_estimate.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp]
        L.orc_det_log.argtypes = [C.c_double]
        L.orc_det_log.restype = C.c_double
        L.orc_det_sincos_turn.argtypes = [C.c_uint32, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_normal_pair.argtypes = [C.c_uint32, C.c_uint32, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_gauss_taps.argtypes = [C.c_int, C.c_double, _fp]
        L.orc_blur_u8.argtypes = [_u8p, C.c_int, C.c_int, C.c_int, _fp, _fp, _u8p]
        L.orc_tracker_proposals.argtypes = [C.c_uint64, C.c_int, C.c_int64, C.c_int, C.c_int, _ip, C.c_int,
                                            C.c_int, _dp, C.c_double, _dp]
        L.orc_tracker_weights.argtypes = [C.c_int, _dp, _dp, _dp, _u8p, C.c_int, C.c_int, C.c_double,
                                          C.c_double, _dp]
        L.orc_tracker_measurement.argtypes = [C.c_int, _ip, C.c_int, _dp, _ip, _dp]
        L.orc_rpy.argtypes = [C.c_double, C.c_double, C.c_double, _dp]
        L.orc_get3Dpose.argtypes = [_dp, C.c_double, C.c_double, C.c_double, C.c_double, _dp]
        L.orc_tf_rows.argtypes = [_dp, _dp, _dp]
        L.orc_pose2d.argtypes = [_dp, _dp, _dp]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_get_threads.restype = C.c_int
        _lib = L
    return _lib


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ------------------------------------------------------------------ RNG
def philox4x32_10(ctr, key):
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(np.ascontiguousarray(ctr, np.uint32), np.ascontiguousarray(key, np.uint32), out)
    return out


def philox_u32(seed: int, stream: int, first: int, count: int):
    out = np.zeros(count, np.uint32)
    lib().orc_philox_u32(seed, stream, first, count, out)
    return out


def det_exp(x: float) -> float:
    return lib().orc_det_exp(float(x))


# ----------------------------------------------------------- resampling
def resample_linear(w, N: int, U: int):
    w = _f64(w)
    out = np.zeros(N, np.int32)
    st = lib().orc_resample_linear(w, w.size, N, int(U), out)
    return out, st


def resample_scaled(w, N: int, U: int, scale: float):
    """Systematic resampling of linear weights bounded by 1/scale:
    q_i = floor(2^32 w_i scale) instead of the maximum-relative q."""
    w = _f64(w)
    out = np.zeros(N, np.int32)
    st = lib().orc_resample_scaled(w, w.size, N, int(U), float(scale), out)
    return out, st


def block_pow(m: float) -> int:
    """Exponent of a block of log-weights whose largest finite value is m:
    ceil(m log2 e), clamped to +-2^30."""
    return int(lib().orc_block_pow(float(m)))


def resample_log(lw, N: int, U: int):
    lw = _f64(lw)
    out = np.zeros(N, np.int32)
    st = lib().orc_resample_log(lw, lw.size, N, int(U), out)
    return out, st


def resample_wheel(w, N: int, u: float, fallback=None):
    w = _f64(w)
    out = np.zeros(N, np.int32)
    fb = None if fallback is None else _i32(fallback)
    st = lib().orc_resample_wheel(w, w.size, N, float(u), None if fb is None else fb.ctypes.data, out)
    return out, st


# -------------------------------------------------------------- KF_model
def kf_predict(F, B, Q, x, P):
    x = _f64(x).copy(); P = _f64(P).copy()
    lib().orc_kf_predict(x.size, _f64(F), _f64(B), _f64(Q), x, P)
    return x, P


def kf_update(H, BH, R, z, x, P):
    x = _f64(x).copy(); P = _f64(P).copy()
    H = _f64(H)
    st = lib().orc_kf_update(x.size, H.shape[0], H, _f64(BH), _f64(R), _f64(z), x, P)
    if st != 0:
        raise ArithmeticError("S singular")
    return x, P


def mvn_logpdf(x, u, S):
    x = _f64(x)
    return lib().orc_mvn_logpdf(x.size, x, _f64(u), _f64(S))


def load_gaussian(u, s, pca_proj, pca_mean, g, sel, R):
    d = len(u); npose = pca_proj.shape[1]; m = len(sel)
    F = np.zeros((d, d)); B = np.zeros(d); Q = np.zeros((d, d))
    H = np.zeros((m, d)); BH = np.zeros(m); Ro = np.zeros((m, m))
    lib().orc_load_gaussian(d, npose, m, _i32(sel), _f64(u), _f64(s), _f64(pca_proj),
                            _f64(pca_mean), float(g), _f64(R), F, B, Q, H, BH, Ro)
    return dict(F=F, B=B, Q=Q, H=H, BH=BH, R=Ro)


# ------------------------------------------------------------------ RBPF
class Model:
    """FP64 RBPF model (my_gmm + shared KF_model H/BH/R)."""

    def __init__(self, pi, gamma, mu, Sigma, H, BH, R):
        self.pi = _f64(pi); self.gamma = _f64(gamma); self.mu = _f64(mu)
        self.Sigma = _f64(Sigma); self.H = _f64(H); self.BH = _f64(BH); self.R = _f64(R)
        self.K, self.d = self.mu.shape
        self.m = self.H.shape[0]
        self._s = _Model(self.K, self.d, self.m, self.pi.ctypes.data, self.gamma.ctypes.data,
                         self.mu.ctypes.data, self.Sigma.ctypes.data, self.H.ctypes.data,
                         self.BH.ctypes.data, self.R.ctypes.data)

    def reset(self, bins):
        bins = _i32(bins)
        N = bins.size
        x = np.zeros((N, self.d)); P = np.zeros((N, self.d, self.d))
        lib().orc_reset_tracker(C.byref(self._s), N, bins, x, P)
        return x, P

    def update_ref(self, x, P, Z, U1, U2):
        """One ParticleFilter::update in reference mode.  Returns a dict."""
        x = _f64(x).copy(); P = _f64(P).copy()
        N = x.shape[0]
        lw = np.zeros(N); anc = np.zeros(N, np.int32); comp = np.zeros(N, np.int32)
        xu = np.zeros_like(x); Pu = np.zeros_like(P)
        lml = C.c_double(0.0)
        st = lib().orc_rbpf_update_ref(C.byref(self._s), N, x, P, _f64(Z), int(U1), int(U2),
                                       lw, anc, comp, C.byref(lml), xu.ctypes.data, Pu.ctypes.data)
        return dict(x=x, P=P, lw=lw, anc=anc, comp=comp, lml=lml.value, xu=xu, Pu=Pu, status=st)

    def update_pdaf(self, x, P, clutter, count, cand, U1, U2, pd, gate, lam):
        x = _f64(x).copy(); P = _f64(P).copy()
        N = x.shape[0]
        clutter = _i32(clutter); count = _i32(count); cand = _f64(cand)
        J, cap = cand.shape[0], cand.shape[1]
        fr = _Frame(J, cap, clutter.ctypes.data, count.ctypes.data, cand.ctypes.data,
                    float(pd), float(gate), float(lam))
        lw = np.zeros(N); anc = np.zeros(N, np.int32); comp = np.zeros(N, np.int32)
        xu = np.zeros_like(x); Pu = np.zeros_like(P)
        lml = C.c_double(0.0)
        st = lib().orc_rbpf_update_pdaf(C.byref(self._s), N, x, P, C.byref(fr), int(U1), int(U2),
                                        lw, anc, comp, C.byref(lml), xu.ctypes.data, Pu.ctypes.data)
        if st < 0:
            raise ValueError("bad arguments")
        return dict(x=x, P=P, lw=lw, anc=anc, comp=comp, lml=lml.value, xu=xu, Pu=Pu, status=st)


def rbpf_estimate(x):
    x = _f64(x)
    out = np.zeros(x.shape[1])
    lib().orc_rbpf_estimate(x.shape[0], x.shape[1], x, out)
    return out


def get_samples(Hp, M, xbar, n, scale, normals):
    Hp = _f64(Hp); out = np.zeros((2, n))
    lib().orc_get_samples(Hp.shape[1], Hp.shape[0], Hp, _f64(M), _f64(xbar), n, float(scale), _f64(normals), out)
    return out


def get_sample_prob(Hp, M, xbar, inp, scale):
    Hp = _f64(Hp); inp = _f64(inp); n = inp.shape[1]
    w = np.zeros(n)
    lib().orc_get_sample_prob(Hp.shape[1], Hp.shape[0], Hp, _f64(M), _f64(xbar), n, inp, float(scale), w)
    return w


# ------------------------------------------------------------------ pf2D
def pf2d_load_gaussian(S):
    S = _f64(S); d = S.shape[0]
    Sinv = np.zeros((d, d)); c = C.c_double(0.0)
    if lib().orc_pf2d_load_gaussian(d, S, Sinv, C.byref(c)) != 0:
        raise ArithmeticError("not SPD")
    return Sinv, c.value


def pf2d_init(N, d, side, unif):
    out = np.zeros((N, d))
    lib().orc_pf2d_init(N, d, int(side), _f64(unif), out)
    return out


class Pf2dGmm:
    def __init__(self, pi, mu, Sinv, cnorm):
        self.pi = _f64(pi); self.mu = _f64(mu); self.Sinv = _f64(Sinv); self.cnorm = _f64(cnorm)
        self.K, self.d = self.mu.shape
        self._s = _Pf2dGmm(self.K, self.d, self.pi.ctypes.data, self.mu.ctypes.data,
                           self.Sinv.ctypes.data, self.cnorm.ctypes.data)

    def update(self, particles, meas, U, normals, reinit_unif, side, flags=0):
        p = _f64(particles).copy()
        N = p.shape[0]
        w = np.zeros(N); lw = np.zeros(N); anc = np.zeros(N, np.int32)
        st = lib().orc_pf2d_update(C.byref(self._s), N, p, w, lw, anc, _f64(meas), int(U),
                                   _f64(normals), _f64(reinit_unif), int(side), int(flags))
        return dict(particles=p, weights=w, lw=lw, anc=anc, status=st)


def pf2d_estimate(particles, weights):
    p = _f64(particles); out = np.zeros(p.shape[1])
    lib().orc_pf2d_estimate(p.shape[0], p.shape[1], p, _f64(weights), out)
    return out


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


# ------------------------------------------------------------- adjoint reference
class _AdjProblem(C.Structure):
    _fields_ = [("K", C.c_int), ("d", C.c_int), ("m", C.c_int), ("N", C.c_int), ("T", C.c_int),
                ("cap", C.c_int), ("H", C.c_void_p), ("BH", C.c_void_p), ("R", C.c_void_p),
                ("clutter", C.c_void_p), ("count", C.c_void_p), ("cand", C.c_void_p),
                ("pd", C.c_double), ("gate", C.c_double), ("lam", C.c_double),
                ("bins", C.c_void_p), ("comp", C.c_void_p), ("anc", C.c_void_p)]


class AdjProblem:
    """A recorded candidate-mode run (frozen component draws comp[T][N] and
    ancestors anc[T][N], initial bins[N], detections) whose log-marginal-
    likelihood is differentiated w.r.t. theta = [means | logits | rates | chol]
    (``orc_adj_*`` of ``pf_oracle.h``)."""

    def __init__(self, H, BH, R, clutter, count, cand, bins, comp, anc, pd=0.95,
                 gate=9.210340371976184, lam=3.0 / (640 * 480), K=15):
        self._keep = [_f64(H), _f64(BH), _f64(R), _i32(clutter), _i32(count), _f64(cand),
                      _i32(bins), _i32(comp), _i32(anc)]
        H_, BH_, R_, cl_, cnt_, cand_, bins_, comp_, anc_ = self._keep
        T, N = comp_.shape
        m, d = H_.shape
        cap = cand_.shape[2]
        self.K, self.d, self.m, self.N, self.T = K, d, m, N, T
        self.s = _AdjProblem(K, d, m, N, T, cap, H_.ctypes.data, BH_.ctypes.data, R_.ctypes.data,
                             cl_.ctypes.data, cnt_.ctypes.data, cand_.ctypes.data, pd, gate, lam,
                             bins_.ctypes.data, comp_.ctypes.data, anc_.ctypes.data)
        self.n_theta = lib().orc_adj_n_theta(C.byref(self.s))

    @staticmethod
    def theta(means, weights, gamma, covs):
        """theta from the my_gmm parameters: logits = log(weights), rates = 1 - gamma,
        chol = lower Cholesky factors of the covariances."""
        means = _f64(means)
        K, d = means.shape
        Ls = np.linalg.cholesky(_f64(covs).reshape(K, d, d))
        return np.concatenate([means.ravel(), np.log(_f64(weights)), 1.0 - _f64(gamma), Ls.ravel()])

    def split(self, theta):
        K, d = self.K, self.d
        a = K * d
        return (theta[:a].reshape(K, d), theta[a:a + K], theta[a + K:a + 2 * K],
                theta[a + 2 * K:].reshape(K, d, d))

    def value(self, theta, logpi0=None):
        theta = _f64(theta)
        if logpi0 is None:
            lg = theta[self.K * self.d:self.K * self.d + self.K]
            logpi0 = lg - np.logaddexp.reduce(lg)
        F = C.c_double(0.0)
        lml = np.zeros(self.T)
        rc = lib().orc_adj_value(C.byref(self.s), theta.ctypes.data, _f64(logpi0).ctypes.data, C.byref(F),
                                 lml.ctypes.data)
        assert rc == 0
        return F.value, lml

    def grad(self, theta):
        theta = _f64(theta)
        F = C.c_double(0.0)
        g = np.zeros(self.n_theta)
        rc = lib().orc_adj_grad(C.byref(self.s), theta.ctypes.data, C.byref(F), g.ctypes.data)
        assert rc == 0
        return F.value, g

    def dirderiv(self, theta, V):
        """grad F(theta) . V[c] for the rows of V (at most 8), one forward-mode pass."""
        theta = _f64(theta)
        V = np.ascontiguousarray(_f64(V).reshape(-1, self.n_theta))
        F = C.c_double(0.0)
        dd = np.zeros(V.shape[0])
        rc = lib().orc_adj_dirderiv(C.byref(self.s), theta.ctypes.data, V.shape[0], V.ctypes.data, C.byref(F),
                                    dd.ctypes.data)
        assert rc == 0
        return F.value, dd


# ------------------------------------------------- pose tracker stage
def det_log(x: float) -> float:
    return lib().orc_det_log(float(x))


def det_sincos_turn(w: int):
    s = C.c_double(); c = C.c_double()
    lib().orc_det_sincos_turn(int(w), C.byref(s), C.byref(c))
    return s.value, c.value


def normal_pair(w0: int, w1: int):
    a = C.c_double(); b = C.c_double()
    lib().orc_normal_pair(int(w0), int(w1), C.byref(a), C.byref(b))
    return a.value, b.value


def gauss_taps(ksize: int = 15, sigma: float = 3.0):
    t = np.zeros(ksize, np.float32)
    lib().orc_gauss_taps(int(ksize), float(sigma), t)
    return t


def blur_u8(img, ksize: int = 15, sigma: float = 3.0):
    img = np.ascontiguousarray(img, np.uint8)
    rows, cols = img.shape
    tmp = np.zeros((rows, cols), np.float32)
    out = np.zeros_like(img)
    if lib().orc_blur_u8(img, rows, cols, int(ksize), gauss_taps(ksize, sigma), tmp, out) != 0:
        raise ValueError("image smaller than the blur radius")
    return out


def tracker_proposals(seed, arm, g, n_c, uniform, roi, rows, cols, y, spread=0.8):
    out = np.zeros((2, n_c))
    lib().orc_tracker_proposals(int(seed), int(arm), int(g), int(n_c), int(bool(uniform)), _i32(roi), int(rows),
                                int(cols), _f64(y), float(spread), out)
    return out


def tracker_weights(props, p_own, p_other, img, pa=0.05, z_floor=1e-4):
    props = _f64(props); img = np.ascontiguousarray(img, np.uint8)
    n_c = props.shape[1]
    w = np.zeros(n_c)
    lib().orc_tracker_weights(n_c, props, _f64(p_own), _f64(p_other), img, img.shape[0], img.shape[1],
                              float(pa), float(z_floor), w)
    return w


def tracker_measurement(bins, props, roi):
    bins = _i32(bins); props = _f64(props)
    Z = np.zeros((6, bins.size))
    lib().orc_tracker_measurement(bins.size, bins, props.shape[1], props, _i32(roi), Z)
    return Z


def rpy(roll, pitch, yaw):
    R = np.zeros(9)
    lib().orc_rpy(float(roll), float(pitch), float(yaw), R)
    return R.reshape(3, 3)


def get3Dpose(e, fx=525.0, fy=525.0, cx=319.5, cy=239.5):
    out = np.zeros(15)
    if lib().orc_get3Dpose(_f64(e), float(fx), float(fy), float(cx), float(cy), out) != 0:
        raise ValueError("singular projection")
    return out.reshape(3, 5)


def tf_rows(p1, p2):
    """publishTFtree's joints relative to Head, 8 x 3 (extractTFdata order)."""
    out = np.zeros(24)
    lib().orc_tf_rows(_f64(p1).ravel(), _f64(p2).ravel(), out)
    return out.reshape(8, 3)


def pose2d(e1, e2):
    out = np.zeros(16)
    lib().orc_pose2d(_f64(e1), _f64(e2), out)
    return out.reshape(8, 2)


def project(Hp, M, xbar):
    """y = Hp xbar + M with the products of each row summed in index order
    (the order of orc_get_samples and of the tracker's kernels)."""
    Hp = _f64(Hp); M = _f64(M); xbar = _f64(xbar)
    y = np.zeros(Hp.shape[0])
    for r in range(Hp.shape[0]):
        s = 0.0
        for k in range(Hp.shape[1]):
            s = s + float(Hp[r, k]) * float(xbar[k])
        y[r] = s + float(M[r])
    return y


class PoseTrackerOracle:
    """PFTracker's per-frame stage around two arm models (pfPose.cpp:210-348),
    FP64.  Hp[a] = pca_proj[a]^T (npose x d), M[a] = pca_mean[a]."""

    def __init__(self, arms, Hp, M, N, seed=1, rows=480, cols=640, ppp=10, pa=0.05, z_floor=1e-4,
                 spread=0.8, ksize=15, sigma=3.0, camera=(525.0, 525.0, 319.5, 239.5)):
        self.arms, self.Hp, self.M = list(arms), [_f64(h) for h in Hp], [_f64(m) for m in M]
        self.N, self.seed, self.rows, self.cols = N, seed, rows, cols
        self.n_c = ppp * N
        self.pa, self.z_floor, self.spread = pa, z_floor, spread
        self.ksize, self.sigma, self.camera = ksize, sigma, camera

    def frame_words(self, g):
        """u_comp, u_res of arm 1, of arm 2, proposal uniforms of arm 1, arm 2."""
        return philox_u32(self.seed, 24, 8 * g, 6)

    def proposal_stage(self, xbars, img, roi, g, uniform, weights_from=None):
        """Blur, proposals, weights, proposal resampling and measurements.
        xbars: the two arms' estimates.  weights_from: optional pair of weight
        vectors to resample instead of the oracle's own (teacher forcing)."""
        blurred = blur_u8(img, self.ksize, self.sigma)
        words = self.frame_words(g)
        width = float(roi[2])
        ys = [project(self.Hp[a][:2], self.M[a][:2], xbars[a]) for a in range(2)]
        props = [tracker_proposals(self.seed, a, g, self.n_c, uniform, roi, self.rows, self.cols, ys[a],
                                   self.spread) for a in range(2)]
        out = dict(blurred=blurred, y=ys, props=props, w=[], bins=[], Z=[])
        for a in range(2):
            p_own = get_sample_prob(self.Hp[a], self.M[a], xbars[a], props[a], width)
            p_oth = get_sample_prob(self.Hp[1 - a], self.M[1 - a], xbars[1 - a], props[a], width)
            w = tracker_weights(props[a], p_own, p_oth, blurred, self.pa, self.z_floor)
            out["w"].append(w)
            wr = w if weights_from is None else weights_from[a]
            bins, _ = resample_scaled(wr, self.N, int(words[4 + a]), self.z_floor * (1.0 - 2.0 * self.pa))
            out["bins"].append(bins)
            out["Z"].append(tracker_measurement(bins, props[a], roi))
        return out

    def update_arms(self, states, Zs, g):
        """ParticleFilter::update of both arms with their measurements (Z
        rounded to FP32, the filter's input precision)."""
        words = self.frame_words(g)
        res = []
        for a in range(2):
            x, P = states[a]
            Z32 = np.asarray(Zs[a], np.float32).astype(np.float64)
            res.append(self.arms[a].update_ref(x, P, Z32, int(words[2 * a]), int(words[2 * a + 1])))
        return res

    def pose(self, xbars):
        """e, get3Dpose, TF rows and 2-D points of both arms' estimates."""
        fx, fy, cx, cy = self.camera
        es = [project(self.Hp[a], self.M[a], xbars[a]) for a in range(2)]
        p3 = [get3Dpose(e, fx, fy, cx, cy) for e in es]
        return es, p3, tf_rows(p3[0], p3[1]), pose2d(es[0], es[1])
