// Device/host math for the S^2 visual factor, the frozen-IMU factor and the left-SE(3)
// manifold. Plain structs of scalars (no Eigen), templated on the arithmetic type so the FP64
// and the mixed-precision kernels share one implementation.
//
// Reference semantics followed here:
//   S^2 factor        src/backend/VisualSphereBatchMath.hpp:29-174
//   sphere basis      src/backend/WindowVisualFactor.cpp:324-328
//   Huber weighting   src/backend/VisualCudaBatch.cu:46-72,125-136
//   left SE(3)        src/backend/GpuRetainedFactorMath.hpp:14-51, LeftSE3Manifold.cpp
//   IMU residual      src/backend/GpuRetainedFactorMath.hpp:53-140
//
// The pose Jacobians are produced directly in tangent coordinates: for a left perturbation of
// the anchor pose the world point moves by [rho_t - [p_w]x theta], so J_anchor = [D, -D[p_w]x]
// with D = d r / d p_C * R_WCb^T, and the observer Jacobian is exactly -J_anchor. This equals
// the reference's ambient Jacobian times LeftSE3Manifold::PlusJacobian (the quaternion lift
// times its plus-Jacobian is |q|^2 I).
#pragma once

#include <cmath>
#include <cstdint>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define VIB_HD __host__ __device__ __forceinline__
#else
#define VIB_HD inline
#endif

namespace vib {

template <class T>
struct V3 {
    T x, y, z;
};
template <class T>
struct M3 {  // row-major
    T m[9];
};

// 1 / sqrt(x) for x > 0. In device code, inside the normal range, the hardware estimate refined
// by two Newton steps (a few ulp, no division); elsewhere the quotient of the rounded root.
VIB_HD double inv_sqrt(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    if (x > 1e-280 && x < 1e280) {
        double q = __builtin_amdgcn_rsq(x);
        const double h = 0.5 * x;
        q = q * fma(-h * q, q, 1.5);
        q = q * fma(-h * q, q, 1.5);
        return q;
    }
#endif
    return 1.0 / sqrt(x);
}
VIB_HD float inv_sqrt(float x) { return 1.0f / sqrtf(x); }
// sin and cos of one argument (one range reduction).
VIB_HD void sin_cos(double a, double& s, double& c) { sincos(a, &s, &c); }
VIB_HD void sin_cos(float a, float& s, float& c) { sincosf(a, &s, &c); }

template <class T> VIB_HD V3<T> v3(T x, T y, T z) { return V3<T>{x, y, z}; }
// (any pointer: decltype(+p[0]) is the element type without qualifiers or address space)
template <class P> VIB_HD auto load3(P p) -> V3<decltype(+p[0])> { return {p[0], p[1], p[2]}; }
template <class T> VIB_HD V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> VIB_HD V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> VIB_HD V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> VIB_HD T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> VIB_HD V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> VIB_HD T sqnorm(V3<T> a) { return dot(a, a); }

template <class T> VIB_HD V3<T> mul(const M3<T>& R, V3<T> v) {
    return {R.m[0] * v.x + R.m[1] * v.y + R.m[2] * v.z, R.m[3] * v.x + R.m[4] * v.y + R.m[5] * v.z,
            R.m[6] * v.x + R.m[7] * v.y + R.m[8] * v.z};
}
template <class T> VIB_HD V3<T> mulT(const M3<T>& R, V3<T> v) {  // R^T v
    return {R.m[0] * v.x + R.m[3] * v.y + R.m[6] * v.z, R.m[1] * v.x + R.m[4] * v.y + R.m[7] * v.z,
            R.m[2] * v.x + R.m[5] * v.y + R.m[8] * v.z};
}
template <class T> VIB_HD M3<T> mul(const M3<T>& A, const M3<T>& B) {
    M3<T> o;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            o.m[r * 3 + c] = A.m[r * 3] * B.m[c] + A.m[r * 3 + 1] * B.m[3 + c] + A.m[r * 3 + 2] * B.m[6 + c];
    return o;
}

// Eigen convention xyzw quaternion -> rotation matrix of the normalized quaternion
// (QuaternionBase::toRotationMatrix on q.normalized()).
template <class P, class T = decltype(+(*(P)nullptr))>
VIB_HD M3<T> quat_to_matrix(P q) {
    const T n = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    const T x = q[0] / n, y = q[1] / n, z = q[2] / n, w = q[3] / n;
    const T tx = 2 * x, ty = 2 * y, tz = 2 * z;
    const T twx = tx * w, twy = ty * w, twz = tz * w, txx = tx * x, txy = ty * x, txz = tz * x;
    const T tyy = ty * y, tyz = tz * y, tzz = tz * z;
    M3<T> r;
    r.m[0] = 1 - (tyy + tzz); r.m[1] = txy - twz; r.m[2] = txz + twy;
    r.m[3] = txy + twz; r.m[4] = 1 - (txx + tzz); r.m[5] = tyz - twx;
    r.m[6] = txz - twy; r.m[7] = tyz + twx; r.m[8] = 1 - (txx + tyy);
    return r;
}

// ---- quaternion algebra (Hamilton, storage x y z w) -------------------------------------
template <class T> struct Qt { T x, y, z, w; };
template <class P> VIB_HD auto qload(P p) -> Qt<decltype(+p[0])> { return {p[0], p[1], p[2], p[3]}; }
template <class T> VIB_HD Qt<T> qmul(Qt<T> a, Qt<T> b) {
    return {a.w * b.x + b.w * a.x + a.y * b.z - a.z * b.y, a.w * b.y + b.w * a.y + a.z * b.x - a.x * b.z,
            a.w * b.z + b.w * a.z + a.x * b.y - a.y * b.x, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
template <class T> VIB_HD Qt<T> qconj(Qt<T> a) { return {-a.x, -a.y, -a.z, a.w}; }
template <class T> VIB_HD Qt<T> qnormalized(Qt<T> a) {
    const T n = sqrt(a.x * a.x + a.y * a.y + a.z * a.z + a.w * a.w);
    return {a.x / n, a.y / n, a.z / n, a.w / n};
}
template <class T> VIB_HD V3<T> qvec(Qt<T> a) { return {a.x, a.y, a.z}; }
template <class T> VIB_HD V3<T> qrotate(Qt<T> q, V3<T> v) {  // Eigen _transformVector
    V3<T> uv = cross(qvec(q), v);
    uv = uv + uv;
    return v + q.w * uv + cross(qvec(q), uv);
}
template <class T> VIB_HD M3<T> qmatrix(Qt<T> q) {  // toRotationMatrix (no normalization)
    const T tx = 2 * q.x, ty = 2 * q.y, tz = 2 * q.z;
    const T twx = tx * q.w, twy = ty * q.w, twz = tz * q.w, txx = tx * q.x, txy = ty * q.x, txz = tz * q.x;
    const T tyy = ty * q.y, tyz = tz * q.y, tzz = tz * q.z;
    M3<T> r;
    r.m[0] = 1 - (tyy + tzz); r.m[1] = txy - twz; r.m[2] = txz + twy;
    r.m[3] = txy + twz; r.m[4] = 1 - (txx + tzz); r.m[5] = tyz - twx;
    r.m[6] = txz - twy; r.m[7] = tyz + twx; r.m[8] = 1 - (txx + tyy);
    return r;
}

// ---- left SE(3) (GpuRetainedFactorMath.hpp:14-51) ----------------------------------------
template <class T>
VIB_HD Qt<T> rotation_exp(V3<T> a) {
    const T n = sqrt(sqnorm(a));
    const T s = n < T(1e-8) ? T(0.5) - n * n / T(48) : sin(T(0.5) * n) / n;
    return {s * a.x, s * a.y, s * a.z, cos(T(0.5) * n)};
}
// J_l(theta) applied to v
template <class T>
VIB_HD V3<T> rotation_jacobian_mul(V3<T> a, V3<T> v) {
    const T a2 = sqnorm(a), n = sqrt(a2);
    const T c1 = a2 < T(1e-10) ? T(0.5) - a2 / T(24) : (1 - cos(n)) / a2;
    const T c2 = a2 < T(1e-10) ? T(1) / T(6) - a2 / T(120) : (n - sin(n)) / (a2 * n);
    const V3<T> wv = cross(a, v);
    return v + c1 * wv + c2 * cross(a, wv);
}
template <class PX, class PD, class PY, class T = decltype(+(*(PX)nullptr))>
VIB_HD void pose_plus(PX x, PD d, PY y) {
    const V3<T> angle = load3(d + 3);
    const Qt<T> q = rotation_exp(angle);
    const V3<T> p = qrotate(q, load3(x)) + rotation_jacobian_mul(angle, load3(d));
    const Qt<T> yq = qnormalized(qmul(q, qload(x + 3)));
    y[0] = p.x; y[1] = p.y; y[2] = p.z;
    y[3] = yq.x; y[4] = yq.y; y[5] = yq.z; y[6] = yq.w;
}
template <class PX, class PR, class PD, class T = decltype(+(*(PX)nullptr))>
VIB_HD void pose_minus(PX x, PR ref, PD d) {
    Qt<T> q = qnormalized(qmul(qload(x + 3), qconj(qload(ref + 3))));
    if (q.w < 0) q = {-q.x, -q.y, -q.z, -q.w};
    const T s = sqrt(q.x * q.x + q.y * q.y + q.z * q.z);
    const T k = s < T(1e-10) ? T(2) : T(2) * atan2(s, q.w) / s;
    const V3<T> th = {k * q.x, k * q.y, k * q.z};
    const T a2 = sqnorm(th), a = sqrt(a2);
    const T c = a2 < T(1e-10) ? T(1) / T(12) + a2 / T(720) : (1 - T(0.5) * a / tan(T(0.5) * a)) / a2;
    const V3<T> p = load3(x) - qrotate(q, load3(ref));
    const V3<T> wp = cross(th, p);
    const V3<T> r = p - T(0.5) * wp + c * cross(th, wp);
    d[0] = r.x; d[1] = r.y; d[2] = r.z; d[3] = th.x; d[4] = th.y; d[5] = th.z;
}

// ---- S^2 factor ----------------------------------------------------------------------------
// sphereBasis (WindowVisualFactor.cpp:324-328): columns c0, c1.
template <class T>
VIB_HD void sphere_basis(V3<T> b, V3<T>& c0, V3<T>& c1) {
    const V3<T> axis = fabs(b.x) < T(0.8) ? v3<T>(1, 0, 0) : v3<T>(0, 1, 0);
    V3<T> t = axis - dot(b, axis) * b;
    const T in = inv_sqrt(sqnorm(t));
    c0 = {t.x * in, t.y * in, t.z * in};
    c1 = cross(b, c0);
}

// timeBearing (VisualSphereBatchMath.hpp:29-49)
template <class T>
VIB_HD V3<T> time_bearing(V3<T> raw, V3<T> v, T dt, bool differentiate, V3<T>& derivative) {
    if (!differentiate) {
        const V3<T> u = dt * v;
        const T angle = sqrt(sqnorm(u));
        if (angle < T(1e-8)) {
            const V3<T> s = raw + u;
            const T in = inv_sqrt(sqnorm(s));
            return {s.x * in, s.y * in, s.z * in};
        }
        T sn, cs;
        sin_cos(angle, sn, cs);
        return cs * raw + (sn / angle) * u;
    }
    const T speed = sqrt(sqnorm(v)), angle = speed * dt;
    if (fabs(angle) < T(1e-8)) {
        const V3<T> dir = raw + dt * v;
        const T il = inv_sqrt(sqnorm(dir));
        const V3<T> unit = {dir.x * il, dir.y * il, dir.z * il};
        const V3<T> pv = v - dot(unit, v) * unit;
        derivative = {-pv.x * il, -pv.y * il, -pv.z * il};
        return unit;
    }
    T sn, cs;
    sin_cos(angle, sn, cs);
    derivative = (speed * sn) * raw - cs * v;
    return cs * raw + (sn / speed) * v;
}

// Residual and the 2x3 derivative d r / d p_C of the S^2 log residual for a predicted camera
// point pc. P (2x3) is the measurement projection (whitening * basis^T, transported). Also
// returns the pieces needed for the td derivative. Returns false on the reference's
// invalid-factor conditions.
template <class T>
struct SphereEval {
    T r[2];
    T reduce[6];  // 2x3 row-major, d r / d p_C
    V3<T> n, u, logv;
    T c, f, g, range;
};
template <class T>
VIB_HD bool sphere_residual(V3<T> pc, V3<T> bb, const T* P, T denominator, bool want_derivative,
                            SphereEval<T>& e) {
    const T range = sqrt(sqnorm(pc));
    if (!isfinite(pc.x) || !isfinite(pc.y) || !isfinite(pc.z) || !isfinite(range) || range < T(1e-10)) return false;
    const T ir = T(1) / range;
    const V3<T> n = {pc.x * ir, pc.y * ir, pc.z * ir};
    const T c = fmin(T(1), fmax(T(-1), dot(bb, n)));
    const T s = sqrt(sqnorm(cross(bb, n)));
    if (c < T(-1) + T(1e-8) || denominator < T(1e-8)) return false;
    const V3<T> u = n - c * bb;
    const T d = 1 - c;
    const bool series = c > 0 && s < T(1e-3);
    const T theta = series ? T(0) : atan2(s, c);
    const T is = series ? T(0) : T(1) / s;  // only read where s >= 1e-3
    const T f = s < T(1e-8) ? T(1) : (series ? 1 + d / 3 + 2 * d * d / 15 + 2 * d * d * d / 35 : theta * is);
    const V3<T> lg = f * u;
    e.r[0] = P[0] * lg.x + P[1] * lg.y + P[2] * lg.z;
    e.r[1] = P[3] * lg.x + P[4] * lg.y + P[5] * lg.z;
    if (!isfinite(e.r[0]) || !isfinite(e.r[1])) return false;
    e.n = n; e.u = u; e.logv = lg; e.c = c; e.f = f; e.range = range;
    if (!want_derivative) return true;
    const T g = s < T(1e-8) ? T(0) : (series ? (T(-1) / T(3) - 4 * d / T(15) - 6 * d * d / T(35))
                                            : (theta * c - s) * (is * is * is));
    e.g = g;
    for (int r = 0; r < 2; ++r) {
        const T* Pr = P + 3 * r;
        const T Pu = Pr[0] * u.x + Pr[1] * u.y + Pr[2] * u.z;
        const T Pb = Pr[0] * bb.x + Pr[1] * bb.y + Pr[2] * bb.z;
        const T k = g * Pu - f * Pb;
        const T pr0 = f * Pr[0] + k * bb.x, pr1 = f * Pr[1] + k * bb.y, pr2 = f * Pr[2] + k * bb.z;
        const T pn = pr0 * n.x + pr1 * n.y + pr2 * n.z;
        e.reduce[3 * r + 0] = (pr0 - pn * n.x) * ir;
        e.reduce[3 * r + 1] = (pr1 - pn * n.y) * ir;
        e.reduce[3 * r + 2] = (pr2 - pn * n.z) * ir;
    }
    return true;
}

// Huber (VisualCudaBatch.cu:55-57,133)
template <class T> VIB_HD T huber_weight(T delta, T sq) {
    return delta > 0 && sq > delta * delta ? sqrt(delta / sqrt(sq)) : T(1);
}
template <class T> VIB_HD T huber_cost(T delta, T sq) {
    return delta > 0 && sq > delta * delta ? delta * (sqrt(sq) - T(0.5) * delta) : T(0.5) * sq;
}

}  // namespace vib
