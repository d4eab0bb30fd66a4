/*
 * pf_pose_ref.c -- FP64 CPU restatement of the pose tracker's per-frame
 * stage around the two arm filters (src/pfPose.cpp:84-127, 210-329).
 *
 * TEST INFRASTRUCTURE ONLY (see pf_oracle.h).
 *
 * The random proposal positions come from Philox words through a Box-Muller
 * transform whose log, sin and cos are evaluated with IEEE-exact operations
 * in a fixed order (orc_det_log, orc_det_sincos_turn), so that the positions
 * are reproducible bit for bit; those two functions are pinned against the
 * C library's log / sin / cos in tests/test_tracker_oracle.py.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "pf_oracle.h"

/* log x, x > 0 normal: x = m 2^e, m in [sqrt(1/2), sqrt(2)),
 * log m = 2 atanh((m - 1)/(m + 1)) by the series through s^25. */
double orc_det_log(double x)
{
    int e;
    double m = frexp(x, &e);
    if (m < 0.70710678118654752440) {
        m = m + m;
        e -= 1;
    }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    static const double inv_odd[12] = {1.0 / 25.0, 1.0 / 23.0, 1.0 / 21.0, 1.0 / 19.0, 1.0 / 17.0, 1.0 / 15.0,
                                       1.0 / 13.0, 1.0 / 11.0, 1.0 / 9.0,  1.0 / 7.0,  1.0 / 5.0,  1.0 / 3.0};
    double p = inv_odd[0];
    for (int i = 1; i < 12; ++i) p = fma(p, z, inv_odd[i]);
    const double s2 = s + s;
    const double lm = fma(s2 * z, p, s2);
    const double de = (double)e;
    return fma(de, 6.93147180369123816490e-01, fma(de, 1.90821492927058770002e-10, lm));
}

/* sin and cos of 2 pi w 2^-32: quadrant from the top two bits, the angle in
 * the quadrant reduced to [0, pi/4] by the complement, Taylor series. */
void orc_det_sincos_turn(uint32_t w, double *sn, double *cs)
{
    const uint32_t q = w >> 30;
    const uint32_t fr = w & 0x3FFFFFFFu;
    const int flip = fr > 0x20000000u;
    const uint32_t g = flip ? 0x40000000u - fr : fr;
    const double x = (double)g * (1.5707963267948966 / 1073741824.0);
    const double z = x * x;
    /* sin: x + x z (-1/3! + z (1/5! - ...)), through x^19 */
    static const double cs_sin[9] = {-1.0 / 121645100408832000.0, 1.0 / 355687428096000.0, -1.0 / 1307674368000.0,
                                     1.0 / 6227020800.0, -1.0 / 39916800.0, 1.0 / 362880.0, -1.0 / 5040.0,
                                     1.0 / 120.0, -1.0 / 6.0};
    /* cos: 1 + z (-1/2! + z (1/4! - ...)), through x^18 */
    static const double cs_cos[9] = {-1.0 / 6402373705728000.0, 1.0 / 20922789888000.0, -1.0 / 87178291200.0,
                                     1.0 / 479001600.0, -1.0 / 3628800.0, 1.0 / 40320.0, -1.0 / 720.0, 1.0 / 24.0,
                                     -0.5};
    double ps = cs_sin[0], pc = cs_cos[0];
    for (int i = 1; i < 9; ++i) {
        ps = fma(ps, z, cs_sin[i]);
        pc = fma(pc, z, cs_cos[i]);
    }
    const double s0 = fma(x * z, ps, x);
    const double c0 = fma(z, pc, 1.0);
    const double a = flip ? c0 : s0, b = flip ? s0 : c0;
    switch (q) {
    case 0: *sn = a; *cs = b; break;
    case 1: *sn = b; *cs = -a; break;
    case 2: *sn = -a; *cs = -b; break;
    default: *sn = -b; *cs = a; break;
    }
}

void orc_normal_pair(uint32_t w0, uint32_t w1, double *z0, double *z1)
{
    const double u1 = ((double)w0 + 1.0) * 2.3283064365386962890625e-10;
    const double r = sqrt(-2.0 * orc_det_log(u1));
    double sn, cs;
    orc_det_sincos_turn(w1, &sn, &cs);
    *z0 = r * cs;
    *z1 = r * sn;
}

/* Gaussian taps of GaussianBlur(Size(ksize, ksize), sigma) (pfPose.cpp:213):
 * exp(-x^2 / (2 sigma^2)), x = -R..R, divided by their sequential sum,
 * rounded to float. */
void orc_gauss_taps(int ksize, double sigma, float *taps)
{
    const int R = ksize / 2;
    double g[64], sum = 0.0;
    const double s2 = -0.5 / (sigma * sigma);
    for (int i = -R; i <= R; ++i) {
        g[i + R] = orc_det_exp((double)(i * i) * s2);
        sum += g[i + R];
    }
    for (int i = 0; i < ksize; ++i) taps[i] = (float)(g[i] / sum);
}

static int reflect101(int i, int n)
{
    if (i < 0) i = -i;
    if (i >= n) i = 2 * n - 2 - i;
    return i;
}

/* Separable blur with BORDER_REFLECT_101: row pass into float, column pass,
 * each fmaf over the taps in index order; round to nearest even, saturate.
 * Requires rows, cols > ksize / 2.  tmp: rows x cols floats. */
int orc_blur_u8(const uint8_t *src, int rows, int cols, int ksize, const float *taps, float *tmp, uint8_t *dst)
{
    const int R = ksize / 2;
    if (rows <= R || cols <= R) return -1;
    for (int y = 0; y < rows; ++y)
        for (int x = 0; x < cols; ++x) {
            float acc = 0.f;
            for (int k = 0; k < ksize; ++k) acc = fmaf(taps[k], (float)src[(size_t)y * cols + reflect101(x + k - R, cols)], acc);
            tmp[(size_t)y * cols + x] = acc;
        }
    for (int y = 0; y < rows; ++y)
        for (int x = 0; x < cols; ++x) {
            float acc = 0.f;
            for (int k = 0; k < ksize; ++k) acc = fmaf(taps[k], tmp[(size_t)reflect101(y + k - R, rows) * cols + x], acc);
            float r = rintf(acc);
            r = r < 0.f ? 0.f : (r > 255.f ? 255.f : r);
            dst[(size_t)y * cols + x] = (uint8_t)r;
        }
    return 0;
}

/* Proposals of one arm for frame g (pfPose.cpp:218-235 with getSamples,
 * pf2DRao.cpp:85-103): candidate c takes words 2 (g n_c + c), +1 of Philox
 * stream 16 + arm.  uniform != 0: the first-frame box below the face,
 * U[xmin, xmax) x U[ymin, ymax) (cv::randu); else y + spread * width * z
 * with (z0, z1) the Box-Muller pair.  out: 2 x n_c (x row, y row). */
void orc_tracker_proposals(uint64_t seed, int arm, int64_t g, int n_c, int uniform, const int roi[4], int rows,
                           int cols, const double y[2], double spread, double *out)
{
    for (int c = 0; c < n_c; ++c) {
        uint32_t w[2];
        orc_philox_u32(seed, 16u + (uint32_t)arm, 2 * ((uint64_t)g * (uint64_t)n_c + (uint64_t)c), 2, w);
        if (uniform) {
            int xmin = roi[0] - 4 * roi[2];
            if (xmin < 0) xmin = 0;
            int xmax = roi[0] + 5 * roi[2];
            if (xmax > cols) xmax = cols;
            int ymin = roi[1] + roi[3];
            if (ymin > rows) ymin = rows;
            int ymax = roi[1] + 7 * roi[3];
            if (ymax > rows) ymax = rows;
            out[c] = fma((double)xmax - (double)xmin, (double)w[0] * 2.3283064365386962890625e-10, (double)xmin);
            out[n_c + c] = fma((double)ymax - (double)ymin, (double)w[1] * 2.3283064365386962890625e-10, (double)ymin);
        } else {
            double z0, z1;
            orc_normal_pair(w[0], w[1], &z0, &z1);
            const double sd = spread * (double)roi[2];
            out[c] = fma(sd, z0, y[0]);
            out[n_c + c] = fma(sd, z1, y[1]);
        }
    }
}

/* Association weights of one arm's proposals (pfPose.cpp:251-266, arm 2
 * :274-289): L = likelihood(floor y, floor x) / 255 inside the open image
 * rectangle, weight L / (p_own Pa + p_other Pa + z_floor (1 - 2 Pa)), 0 if L
 * is 0 or the proposal is outside.  Left unnormalised: the resampler is
 * invariant to the scale the reference divides out (pfPose.cpp:295-299). */
void orc_tracker_weights(int n_c, const double *props, const double *p_own, const double *p_other, const uint8_t *img,
                         int rows, int cols, double pa, double z_floor, double *w)
{
    for (int j = 0; j < n_c; ++j) {
        const double x = props[j], y = props[n_c + j];
        w[j] = 0.0;
        if (y > 0 && y < rows && x > 0 && x < cols) {
            const double L = (double)img[(size_t)(int)y * cols + (int)x] / 255.0;
            if (L != 0.0) {
                const double Z = p_own[j] * pa + p_other[j] * pa + z_floor * (1 - 2 * pa);
                w[j] = L / Z;
            }
        }
    }
}

/* The 6 x N measurement (pfPose.cpp:303-320): rows face centre x, face
 * y + 0.5 h, proposal x, proposal y, face centre x, face y + 1.65 h. */
void orc_tracker_measurement(int N, const int *bins, int n_c, const double *props, const int roi[4], double *Z)
{
    for (int i = 0; i < N; ++i) {
        Z[i] = roi[0] + roi[2] / 2.0;
        Z[N + i] = roi[1] + 0.5 * roi[3];
        Z[2 * N + i] = props[bins[i]];
        Z[3 * N + i] = props[n_c + bins[i]];
        Z[4 * N + i] = roi[0] + roi[2] / 2.0;
        Z[5 * N + i] = roi[1] + 1.65 * roi[3];
    }
}

/* PFTracker::rpy (pfPose.cpp:84-91): R3(yaw) R2(pitch) R1(roll). */
void orc_rpy(double roll, double pitch, double yaw, double R[9])
{
    const double R1[9] = {1, 0, 0, 0, cos(roll), -sin(roll), 0, sin(roll), cos(roll)};
    const double R2[9] = {cos(pitch), 0, sin(pitch), 0, 1, 0, -sin(pitch), 0, cos(pitch)};
    const double R3[9] = {cos(yaw), -sin(yaw), 0, sin(yaw), cos(yaw), 0, 0, 0, 1};
    double T[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += R3[i * 3 + k] * R2[k * 3 + j];
            T[i * 3 + j] = s;
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += T[i * 3 + k] * R1[k * 3 + j];
            R[i * 3 + j] = s;
        }
}

/* PFTracker::get3Dpose (pfPose.cpp:93-127): P = K [R | t], R = rpy(e16, e17,
 * e15), t = e18..20; image points (u d, v d, d) of joints k = 0..4 (hand,
 * elbow, shoulder, head, neck: e[3k], e[3k+1], e[3k+2]); out (3 x 5, row
 * major) = inv(P[:, 0:3]) (points - P[:, 3])^T, the inverse by Gaussian
 * elimination with partial pivoting (invert, DECOMP_LU).  Returns -1 if the
 * 3 x 3 block is singular. */
int orc_get3Dpose(const double *e, double fx, double fy, double cx, double cy, double out[15])
{
    const double K[9] = {fx, 0.0, cx, 0.0, fy, cy, 0.0, 0.0, 1.0};
    double R[9];
    orc_rpy(e[16], e[17], e[15], R);
    const double t[3] = {e[18], e[19], e[20]};
    double P[12];
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += K[i * 3 + k] * R[k * 3 + j];
            P[i * 4 + j] = s;
        }
        double s = 0.0;
        for (int k = 0; k < 3; ++k) s += K[i * 3 + k] * t[k];
        P[i * 4 + 3] = s;
    }
    /* [A | I] -> [I | A^-1] */
    double A[3][6];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 6; ++j) A[i][j] = j < 3 ? P[i * 4 + j] : (j - 3 == i ? 1.0 : 0.0);
    for (int c = 0; c < 3; ++c) {
        int piv = c;
        for (int r = c + 1; r < 3; ++r)
            if (fabs(A[r][c]) > fabs(A[piv][c])) piv = r;
        if (A[piv][c] == 0.0) return -1;
        if (piv != c)
            for (int j = 0; j < 6; ++j) {
                const double tmp = A[c][j];
                A[c][j] = A[piv][j];
                A[piv][j] = tmp;
            }
        for (int r = 0; r < 3; ++r) {
            if (r == c) continue;
            const double f = A[r][c] / A[c][c];
            for (int j = c; j < 6; ++j) A[r][j] -= f * A[c][j];
        }
    }
    double Pi[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Pi[i * 3 + j] = A[i][3 + j] / A[i][i];
    for (int k = 0; k < 5; ++k) {
        const double d = e[3 * k + 2];
        const double im[3] = {e[3 * k] * d - P[3], e[3 * k + 1] * d - P[7], d - P[11]};
        for (int r = 0; r < 3; ++r) {
            double s = 0.0;
            for (int c = 0; c < 3; ++c) s += Pi[r * 3 + c] * im[c];
            out[r * 5 + k] = s;
        }
    }
    return 0;
}

/* publishTFtree (pfPose.cpp:129-171) as extractTFdata records it
 * (extractTFdata.cpp:87-103): the published child-in-parent offsets
 * (identity rotations; a camera-frame difference v is published as
 * (v.x, v.z, -v.y)), composed from Head down each chain as a translation
 * sum.  p1, p2: get3Dpose of arm 1 (published as Right_*) and arm 2
 * (Left_*), 3 x 5 row-major.  out: Head, Left_Hand, Right_Hand, Left_Elbow,
 * Right_Elbow, Right_Shoulder, Left_Shoulder, Neck, each (x, y, z). */
static void tf_offset(const double a[3], const double b[3], double o[3])
{
    o[0] = a[0] - b[0];
    o[1] = a[2] - b[2];
    o[2] = -a[1] + b[1];
}

void orc_tf_rows(const double p1[15], const double p2[15], double out[24])
{
    double J1[5][3], J2[5][3], head[3], neck[3];
    for (int k = 0; k < 5; ++k)
        for (int r = 0; r < 3; ++r) {
            J1[k][r] = p1[r * 5 + k];
            J2[k][r] = p2[r * 5 + k];
        }
    for (int r = 0; r < 3; ++r) {
        neck[r] = (J1[4][r] + J2[4][r]) / 2.0;
        head[r] = (J1[3][r] + J2[3][r]) / 2.0;
    }
    double neck_in_head[3], sh_in_neck[2][3], el_in_sh[2][3], ha_in_el[2][3];
    tf_offset(neck, head, neck_in_head);
    const double(*J[2])[3] = {J1, J2};
    for (int a = 0; a < 2; ++a) {
        tf_offset(J[a][2], neck, sh_in_neck[a]);
        tf_offset(J[a][1], J[a][2], el_in_sh[a]);
        tf_offset(J[a][0], J[a][1], ha_in_el[a]);
    }
    /* rows: 0 Head, 1 LHand, 2 RHand, 3 LElbow, 4 RElbow, 5 RShoulder, 6 LShoulder, 7 Neck */
    const int arm_of[8] = {-1, 1, 0, 1, 0, 0, 1, -1};
    const int depth_of[8] = {0, 3, 3, 2, 2, 1, 1, 0};
    for (int j = 0; j < 8; ++j)
        for (int c = 0; c < 3; ++c) {
            double v = 0.0;
            if (j == 7) v = neck_in_head[c];
            else if (j > 0) {
                const int a = arm_of[j];
                v = neck_in_head[c] + sh_in_neck[a][c];
                if (depth_of[j] >= 2) v += el_in_sh[a][c];
                if (depth_of[j] >= 3) v += ha_in_el[a][c];
            }
            out[j * 3 + c] = v;
        }
}

/* publish2Dpos (pfPose.cpp:173-208): hand 1, hand 2, head, neck (means of
 * the two arms), elbow 1, elbow 2, shoulder 1, shoulder 2, each (x, y). */
void orc_pose2d(const double *e1, const double *e2, double out[16])
{
    const double pts[16] = {e1[0], e1[1], e2[0], e2[1],
                            0.5 * (e1[9] + e2[9]), 0.5 * (e1[10] + e2[10]),
                            0.5 * (e1[12] + e2[12]), 0.5 * (e1[13] + e2[13]),
                            e1[3], e1[4], e2[3], e2[4], e1[6], e1[7], e2[6], e2[7]};
    memcpy(out, pts, sizeof(pts));
}
