// Per-observation bundle-adjustment arithmetic shared by the device kernels.
//
// Each function restates one step of the reference solver with the reference's
// operation order: dense products start from sum = 0 and accumulate k
// ascending (source/matrix.hpp:492-527), scalar factors multiply element-wise,
// and quaternion products are the Hamilton expansion of source/lie.hpp:337-344.
// Compiled without floating-point contraction, the binary64 instantiation gives
// the reference's results bit for bit; the binary32 instantiation is the mixed
// path's arithmetic.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#define ZB_HD __host__ __device__ __forceinline__

namespace zb {

enum : int32_t { LOSS_NONE = 0, LOSS_TRIVIAL = 1, LOSS_HUBER = 2, LOSS_CAUCHY = 3, LOSS_TUKEY = 4 };

template <class T> ZB_HD T zsqrt(T x) { return sqrt(x); }
template <class T> ZB_HD T zlog(T x) { return log(x); }
template <class T> ZB_HD T zsin(T x) { return sin(x); }
template <class T> ZB_HD T zcos(T x) { return cos(x); }
template <> ZB_HD float zsqrt<float>(float x) { return sqrtf(x); }
template <> ZB_HD float zlog<float>(float x) { return logf(x); }
template <> ZB_HD float zsin<float>(float x) { return sinf(x); }
template <> ZB_HD float zcos<float>(float x) { return cosf(x); }

// Dense product, matrix.hpp:499-507.
template <int R, int K, int C, class T>
ZB_HD void mul(const T* a, const T* b, T* o) {
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int c = 0; c < C; ++c) {
            T s = 0;
#pragma unroll
            for (int k = 0; k < K; ++k) s += a[r * K + k] * b[k * C + c];
            o[r * C + c] = s;
        }
}

// Hamilton product (w, x, y, z), lie.hpp:337-344.
template <class T>
ZB_HD void qmul(const T a[4], const T b[4], T o[4]) {
    o[0] = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
    o[1] = a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
    o[2] = a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1];
    o[3] = a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0];
}

// (q * (0, p)) * conj(q), lie.hpp:347-362.
template <class T>
ZB_HD void qrotate(const T q[4], const T p[3], T o[3]) {
    const T pq[4] = {T(0), p[0], p[1], p[2]};
    const T qc[4] = {+q[0], -q[1], -q[2], -q[3]};
    T t[4], r[4];
    qmul(q, pq, t);
    qmul(t, qc, r);
    o[0] = r[1];
    o[1] = r[2];
    o[2] = r[3];
}

// Rotation matrix of a quaternion, lie.hpp:188-204.
template <class T>
ZB_HD void qmatrix(const T q[4], T m[9]) {
    const T tx = T(2.0) * q[1], ty = T(2.0) * q[2], tz = T(2.0) * q[3];
    const T txx = tx * q[1], txy = tx * q[2], txz = tx * q[3], txw = tx * q[0];
    const T tyy = ty * q[2], tyz = ty * q[3], tyw = ty * q[0];
    const T tzz = tz * q[3], tzw = tz * q[0];
    m[0] = T(1.0) - (tyy + tzz); m[1] = txy - tzw;              m[2] = txz + tyw;
    m[3] = txy + tzw;              m[4] = T(1.0) - (txx + tzz); m[5] = tyz - txw;
    m[6] = txz - tyw;              m[7] = tyz + txw;              m[8] = T(1.0) - (txx + tyy);
}

// Pose parameters [tx, ty, tz, qx, qy, qz, qw] -> quaternion (w, x, y, z),
// factor_graph.hpp:993.
template <class T>
ZB_HD void pose_quat(const T p[7], T q[4]) {
    q[0] = p[6]; q[1] = p[3]; q[2] = p[4]; q[3] = p[5];
}

// Camera-frame point T * X, lie.hpp:529-531.
template <class T>
ZB_HD void pose_apply(const T p[7], const T X[3], T Pc[3]) {
    T q[4], r[3];
    pose_quat(p, q);
    qrotate(q, X, r);
    Pc[0] = r[0] + p[0];
    Pc[1] = r[1] + p[1];
    Pc[2] = r[2] + p[2];
}

// pinhole::project fails for a point at a depth below 1e-12 (camera.hpp:98-100);
// the edge then takes its behind-camera residual and Jacobian.
template <class T>
ZB_HD bool behind_camera(const T Pc[3]) {
    return Pc[2] < T(1.0e-12);
}

// edge_reprojection::compute_residual with pinhole::project,
// factor_graph.hpp:999-1008, camera.hpp:97-111.  K = (fx, fy, cx, cy).
template <class T>
ZB_HD void residual(const T Pc[3], const T uv[2], const T K[4], T r[2]) {
    if (behind_camera(Pc)) {
        const T penalty = T(-0.3) * Pc[2];
        r[0] = penalty;
        r[1] = penalty;
        return;
    }
    const T iz = T(1.0) / Pc[2];
    const T x = Pc[0] * iz, y = Pc[1] * iz;
    r[0] = uv[0] - (K[0] * x + K[2]);
    r[1] = uv[1] - (K[1] * y + K[3]);
}

// edge_reprojection::compute_jacobians, factor_graph.hpp:1012-1048, with the
// projection Jacobian of camera.hpp:114-126 (behind the camera: 1022-1025).
template <class T>
ZB_HD void jacobians(const T Pc[3], const T q[4], const T K[4], T Jp[12], T Jl[6]) {
    T Jc[6];
    if (behind_camera(Pc)) {
        Jc[0] = T(0.0); Jc[1] = T(0.0); Jc[2] = T(0.3);
        Jc[3] = T(0.0); Jc[4] = T(0.0); Jc[5] = T(0.3);
    } else {
        const T iz = T(1.0) / Pc[2];
        const T iz2 = iz * iz;
        Jc[0] = K[0] * iz; Jc[1] = T(0.0); Jc[2] = -K[0] * Pc[0] * iz2;
        Jc[3] = T(0.0);    Jc[4] = K[1] * iz; Jc[5] = -K[1] * Pc[1] * iz2;
    }
    const T X = Pc[0], Y = Pc[1], Z = Pc[2];
    const T Jpp[18] = {T(0), Z, -Y, T(1), T(0), T(0), -Z, T(0), X, T(0), T(1), T(0), Y, -X, T(0), T(0), T(0), T(1)};
    T t[12];
    mul<2, 3, 6>(Jc, Jpp, t);
#pragma unroll
    for (int i = 0; i < 12; ++i) Jp[i] = -t[i];
    T R[9], u[6];
    qmatrix(q, R);
    mul<2, 3, 3>(Jc, R, u);
#pragma unroll
    for (int i = 0; i < 6; ++i) Jl[i] = -u[i];
}

// edge_base::chi2: (r^T * Omega) * r, factor_graph.hpp:217-219.
template <class T>
ZB_HD T chi2(const T r[2], const T W[4]) {
    T t[2];
    mul<1, 2, 2>(r, W, t);
    T s = 0;
    s += t[0] * r[0];
    s += t[1] * r[1];
    return s;
}

// Robust losses, factor_graph.hpp:851-934.
template <class T>
ZB_HD void loss(int32_t kind, T delta, T e2, T rho[3]) {
    if (kind == LOSS_HUBER) {
        const T d2 = delta * delta;
        if (e2 <= d2) {
            rho[0] = e2; rho[1] = T(1.0); rho[2] = T(0.0);
        } else {
            const T e = zsqrt(e2);
            rho[0] = T(2.0) * e * delta - d2;
            rho[1] = delta / e;
            rho[2] = T(-0.5) * rho[1] / e2;
        }
    } else if (kind == LOSS_CAUCHY) {
        const T d2 = delta * delta;
        const T d2r = T(1.0) / d2;
        const T aux = d2r * e2 + T(1.0);
        rho[0] = d2 * zlog(aux);
        rho[1] = T(1.0) / aux;
        rho[2] = -d2r * (rho[1] * rho[1]);
    } else if (kind == LOSS_TUKEY) {
        const T d2 = delta * delta;
        if (e2 <= d2) {
            const T aux = e2 / d2;
            const T oma = T(1.0) - aux;
            const T oma2 = oma * oma;
            const T oma3 = oma2 * oma;
            rho[0] = d2 * (T(1.0) - oma3) / T(3.0);
            rho[1] = oma2;
            rho[2] = T(-2.0) * oma / d2;
        } else {
            rho[0] = d2 / T(3.0); rho[1] = T(0); rho[2] = T(0);
        }
    } else {  // trivial (and no loss, where only rho[0] is used)
        rho[0] = e2; rho[1] = T(1); rho[2] = T(0);
    }
}

// edge_base::robust_chi2, factor_graph.hpp:221-228.
template <class T>
ZB_HD T robust_chi2(int32_t kind, T delta, const T r[2], const T W[4]) {
    const T c = chi2(r, W);
    if (kind == LOSS_NONE) return c;
    T rho[3];
    loss(kind, delta, c, rho);
    return rho[0];
}

// Inlier flag: quadratic region of the loss (Huber / Tukey e2 <= delta^2,
// factor_graph.hpp:873, 919); every other kind counts as inlier.
template <class T>
ZB_HD bool inlier(int32_t kind, T delta, const T r[2], const T W[4]) {
    if (kind == LOSS_HUBER || kind == LOSS_TUKEY) return chi2(r, W) <= delta * delta;
    return true;
}

// edge_base::robust_info, factor_graph.hpp:254-269.
template <class T>
ZB_HD void robust_info(int32_t kind, T delta, const T W[4], const T r[2], T& drho, T Wr[4]) {
    if (kind == LOSS_NONE) {
        drho = T(1.0);
#pragma unroll
        for (int i = 0; i < 4; ++i) Wr[i] = W[i];
        return;
    }
    const T e2 = chi2(r, W);
    T rho[3];
    loss(kind, delta, e2, rho);
    drho = rho[1];
#pragma unroll
    for (int i = 0; i < 4; ++i) Wr[i] = W[i] * rho[1];
    if ((rho[1] + T(2.0) * rho[2] * e2) > T(0.0)) {
        T we[2];
        mul<2, 2, 1>(W, r, we);
        const T s = T(2.0) * rho[2];
        const T swe[2] = {we[0] * s, we[1] * s};
        T outer[4];
        mul<2, 1, 2>(swe, we, outer);
#pragma unroll
        for (int i = 0; i < 4; ++i) Wr[i] = Wr[i] + outer[i];
    }
}

// Transposed Jacobian (R x C -> C x R).
template <int R, int C, class T>
ZB_HD void transpose(const T* a, T* o) {
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int c = 0; c < C; ++c) o[c * R + r] = a[r * C + c];
}

// Gradient terms of one edge, factor_graph.hpp:680-689, 725:
//   v = ((drho J^T) Omega) r  (b -= v), for the pose (6) and the point (3).
template <class T>
ZB_HD void pose_grad(const T Jp[12], const T W[4], const T r[2], T drho, T v[6]) {
    T JpT[12], dJ[12], t[12];
    transpose<2, 6>(Jp, JpT);
#pragma unroll
    for (int i = 0; i < 12; ++i) dJ[i] = JpT[i] * drho;
    mul<6, 2, 2>(dJ, W, t);
    mul<6, 2, 1>(t, r, v);
}
template <class T>
ZB_HD void point_grad(const T Jl[6], const T W[4], const T r[2], T drho, T v[3]) {
    T JlT[6], dJ[6], t[6];
    transpose<2, 3>(Jl, JlT);
#pragma unroll
    for (int i = 0; i < 6; ++i) dJ[i] = JlT[i] * drho;
    mul<3, 2, 2>(dJ, W, t);
    mul<3, 2, 1>(t, r, v);
}

// Hessian terms of one edge without the gradient.
template <class T>
ZB_HD void pose_hessian(const T Jp[12], const T Jl[6], const T Wr[4], T Hpp21[21], T Wpl[18]) {
    T JpT[12], jtw[12];
    transpose<2, 6>(Jp, JpT);
    mul<6, 2, 2>(JpT, Wr, jtw);
    int idx = 0;
#pragma unroll
    for (int k = 0; k < 6; ++k)
#pragma unroll
        for (int m = 0; m <= k; ++m) {
            T s = 0;
            s += jtw[k * 2 + 0] * Jp[0 * 6 + m];
            s += jtw[k * 2 + 1] * Jp[1 * 6 + m];
            Hpp21[idx++] = s;
        }
    mul<6, 2, 3>(jtw, Jl, Wpl);
}
template <class T>
ZB_HD void point_hessian(const T Jl[6], const T Wr[4], T Hll[9]) {
    T JlT[6], jtw[6];
    transpose<2, 3>(Jl, JlT);
    mul<3, 2, 2>(JlT, Wr, jtw);
    mul<3, 2, 3>(jtw, Jl, Hll);
}

// Pose-vertex terms of make_hessian for one edge, factor_graph.hpp:680-696, 725:
//   jtw = Jp^T Omega',  H_pp = jtw Jp (lower triangle, 21 entries row-major),
//   W = jtw Jl (6x3),  v = ((drho Jp^T) Omega) r  (b_p -= v).
template <class T>
ZB_HD void pose_terms(const T Jp[12], const T Jl[6], const T Wr[4], const T W[4], const T r[2], T drho, T Hpp21[21],
                      T Wpl[18], T v[6]) {
    pose_hessian(Jp, Jl, Wr, Hpp21, Wpl);
    pose_grad(Jp, W, r, drho, v);
}

// Point-vertex terms, factor_graph.hpp:680-689, 720, 725:
//   H_ll = (Jl^T Omega') Jl (3x3),  v = ((drho Jl^T) Omega) r  (b_l -= v).
template <class T>
ZB_HD void point_terms(const T Jl[6], const T Wr[4], const T W[4], const T r[2], T drho, T Hll[9], T v[3]) {
    point_hessian(Jl, Wr, Hll);
    point_grad(Jl, W, r, drho, v);
}

// Closed-form 3x3 inverse with the singular-to-zero rule, matrix.hpp:749-774.
template <class T>
ZB_HD bool invert3(const T* v, T* out) {
    const T det = +v[0] * (v[4] * v[8] - v[7] * v[5]) - v[1] * (v[3] * v[8] - v[5] * v[6]) +
                  v[2] * (v[3] * v[7] - v[4] * v[6]);
    T scale = 0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const T a = fabs(v[i]);
        scale = scale < a ? a : scale;
    }
    if (!(fabs(det) > scale * scale * scale * T(1e-12))) {
#pragma unroll
        for (int i = 0; i < 9; ++i) out[i] = T(0);
        return false;
    }
    const T di = T(1.0) / det;
    out[0] = (v[4] * v[8] - v[7] * v[5]) * di;
    out[1] = (v[2] * v[7] - v[1] * v[8]) * di;
    out[2] = (v[1] * v[5] - v[2] * v[4]) * di;
    out[3] = (v[5] * v[6] - v[3] * v[8]) * di;
    out[4] = (v[0] * v[8] - v[2] * v[6]) * di;
    out[5] = (v[3] * v[2] - v[0] * v[5]) * di;
    out[6] = (v[3] * v[7] - v[6] * v[4]) * di;
    out[7] = (v[6] * v[1] - v[0] * v[7]) * di;
    out[8] = (v[0] * v[4] - v[3] * v[1]) * di;
    return true;
}

// ---- correctly rounded sin, cos and cube (binary64) ---------------------------
// The reference takes sin/cos/pow from the C library.  These evaluate the
// function in double-double arithmetic (error-free transformations with fma,
// Dekker / Knuth) to ~2^-100 relative and round once, so they return the
// correctly rounded result except when the exact value lies within ~2^-100 of a
// rounding midpoint.  Argument reduction subtracts k*pi/2 with pi/2 split into
// three binary64 parts (Cody-Waite), exact for |k| < 2^20.
struct dd {
    double hi, lo;
};
ZB_HD dd dd_two_sum(double a, double b) {
    const double s = a + b;
    const double bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}
ZB_HD dd dd_quick(double a, double b) {
    const double s = a + b;
    return {s, b - (s - a)};
}
ZB_HD dd dd_two_prod(double a, double b) {
    const double p = a * b;
    return {p, fma(a, b, -p)};
}
ZB_HD dd dd_add(dd a, dd b) {
    dd s = dd_two_sum(a.hi, b.hi);
    const dd t = dd_two_sum(a.lo, b.lo);
    s.lo = s.lo + t.hi;
    s = dd_quick(s.hi, s.lo);
    s.lo = s.lo + t.lo;
    return dd_quick(s.hi, s.lo);
}
ZB_HD dd dd_mul(dd a, dd b) {
    dd p = dd_two_prod(a.hi, b.hi);
    p.lo = p.lo + (a.hi * b.lo + a.lo * b.hi);
    return dd_quick(p.hi, p.lo);
}
ZB_HD dd dd_mul_d(dd a, double b) {
    dd p = dd_two_prod(a.hi, b);
    p.lo = p.lo + a.lo * b;
    return dd_quick(p.hi, p.lo);
}
ZB_HD dd dd_div_d(dd a, double b) {
    const double q1 = a.hi / b;
    const dd p = dd_two_prod(q1, b);
    dd r = dd_two_sum(a.hi, -p.hi);
    r.lo = r.lo - p.lo;
    r.lo = r.lo + a.lo;
    const double q2 = (r.hi + r.lo) / b;
    return dd_quick(q1, q2);
}
// sin(r) / cos(r) by Taylor series in double-double, |r| <= pi/4.
ZB_HD dd dd_sin_kernel(dd r) {
    const dd r2 = dd_mul(r, r);
    dd term = r, sum = r;
    for (int k = 1; k < 16; ++k) {
        term = dd_mul(term, r2);
        term = dd_div_d(term, -static_cast<double>((2 * k) * (2 * k + 1)));
        sum = dd_add(sum, term);
        if (fabs(term.hi) < 1e-36 * fabs(sum.hi)) break;
    }
    return sum;
}
ZB_HD dd dd_cos_kernel(dd r) {
    const dd r2 = dd_mul(r, r);
    dd term = {1.0, 0.0}, sum = {1.0, 0.0};
    for (int k = 1; k < 16; ++k) {
        term = dd_mul(term, r2);
        term = dd_div_d(term, -static_cast<double>((2 * k - 1) * (2 * k)));
        sum = dd_add(sum, term);
        if (fabs(term.hi) < 1e-36) break;
    }
    return sum;
}
// x - k*pi/2 in double-double, returns the quadrant k mod 4.
ZB_HD int dd_reduce(double x, dd& r) {
    const double pio2_1 = 1.5707963267948966;       // nearest binary64 to pi/2
    const double pio2_2 = 6.123233995736766e-17;    // next 53 bits of pi/2
    const double pio2_3 = -1.4973849048591698e-33;  // next 53 bits of pi/2
    if (fabs(x) <= 0.7853981633974483) {
        r = {x, 0.0};
        return 0;
    }
    const double k = rint(x * 0.6366197723675814);
    dd t = dd_two_sum(x, -dd_two_prod(k, pio2_1).hi);
    t.lo = t.lo - dd_two_prod(k, pio2_1).lo;
    t = dd_quick(t.hi, t.lo);
    t = dd_add(t, dd_two_prod(-k, pio2_2));
    t = dd_add(t, dd_two_prod(-k, pio2_3));
    r = t;
    long long q = static_cast<long long>(k);
    return static_cast<int>(q & 3);
}
ZB_HD double cr_sin(double x) {
    if (!(fabs(x) < 1.0e5)) return sin(x);
    if (x == 0.0) return x;
    dd r;
    const int q = dd_reduce(x, r);
    dd v = (q & 1) ? dd_cos_kernel(r) : dd_sin_kernel(r);
    if (q & 2) v = {-v.hi, -v.lo};
    return v.hi + v.lo;
}
ZB_HD double cr_cos(double x) {
    if (!(fabs(x) < 1.0e5)) return cos(x);
    dd r;
    const int q = dd_reduce(x, r);
    dd v = (q & 1) ? dd_sin_kernel(r) : dd_cos_kernel(r);
    if (q == 1 || q == 2) v = {-v.hi, -v.lo};
    return v.hi + v.lo;
}
// x * x * x rounded once (the value of pow(x, 3.0)).
ZB_HD double cr_cube(double x) {
    const dd x2 = dd_two_prod(x, x);
    dd c = dd_two_prod(x2.hi, x);
    c.lo = c.lo + x2.lo * x;
    return c.hi + c.lo;
}

template <class T> ZB_HD T trig_sin(T x) { return zsin(x); }
template <class T> ZB_HD T trig_cos(T x) { return zcos(x); }
template <> ZB_HD double trig_sin<double>(double x) { return cr_sin(x); }
template <> ZB_HD double trig_cos<double>(double x) { return cr_cos(x); }

// so3::exp and so3::left_jacobian, lie.hpp:247-264, 284-294; vertex_pose::plus
// T <- exp(delta) * T, factor_graph.hpp:947-960 via se3::exp / se3::operator*
// (lie.hpp:446-450, 521-526).  delta = (omega, upsilon).
template <class T>
ZB_HD void pose_plus(T p[7], const T d[6]) {
    const T w[3] = {d[0], d[1], d[2]};
    T t2 = 0;
    t2 += w[0] * w[0];
    t2 += w[1] * w[1];
    t2 += w[2] * w[2];
    T real, imag, a, b;
    if (t2 < T(1e-6 * 1e-6)) {
        const T t4 = t2 * t2;
        real = T(1.0) - T(1.0 / 8.0) * t2 + T(1.0 / 384.0) * t4;
        imag = T(0.5) - T(1.0 / 48.0) * t2 + T(1.0 / 3840.0) * t4;
        a = T(0.5);
        b = T(1.0 / 6.0);
    } else {
        const T theta = zsqrt(t2);
        const T half = T(0.5) * theta;
        real = trig_cos(half);
        imag = trig_sin(half) / theta;
        a = (T(1) - trig_cos(theta)) / (t2);
        b = (theta - trig_sin(theta)) / (t2 * theta);
    }
    const T qe[4] = {real, w[0] * imag, w[1] * imag, w[2] * imag};
    const T Wh[9] = {T(0), -w[2], w[1], w[2], T(0), -w[0], -w[1], w[0], T(0)};
    T WW[9], J[9];
    mul<3, 3, 3>(Wh, Wh, WW);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const T I = (i % 4 == 0) ? T(1.0) : T(0.0);
        J[i] = (I + Wh[i] * a) + WW[i] * b;
    }
    T te[3];
    mul<3, 3, 1>(J, d + 3, te);
    T q[4], qn[4], tr[3];
    pose_quat(p, q);
    const T t[3] = {p[0], p[1], p[2]};
    qmul(qe, q, qn);
    qrotate(qe, t, tr);
    p[0] = tr[0] + te[0];
    p[1] = tr[1] + te[1];
    p[2] = tr[2] + te[2];
    p[3] = qn[1];
    p[4] = qn[2];
    p[5] = qn[3];
    p[6] = qn[0];
}

}  // namespace zb
