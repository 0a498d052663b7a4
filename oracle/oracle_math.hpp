// ORACLE -- test infrastructure only. FP64 CPU restatement of the reference algorithm used
// to check the HIP path. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// may load it. It is never part of the product path.
//
// Small fixed-size linear algebra mirroring the Eigen operations the reference uses
// (column vectors, row-major 3x3, Eigen-convention xyzw quaternions).
#pragma once
#include <cmath>
#include <cstring>
#include <algorithm>

namespace oracle {

struct V3 {
    double x = 0, y = 0, z = 0;
    V3() = default;
    V3(double a, double b, double c) : x(a), y(b), z(c) {}
    explicit V3(const double* p) : x(p[0]), y(p[1]), z(p[2]) {}
    double& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    void store(double* p) const { p[0] = x; p[1] = y; p[2] = z; }
};
inline V3 operator+(const V3& a, const V3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(const V3& a, const V3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(const V3& a) { return {-a.x, -a.y, -a.z}; }
inline V3 operator*(double s, const V3& a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator*(const V3& a, double s) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator/(const V3& a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(const V3& a, const V3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double sqnorm(const V3& a) { return dot(a, a); }
inline double norm(const V3& a) { return std::sqrt(sqnorm(a)); }
inline V3 normalized(const V3& a) { return a / norm(a); }
inline bool finite(const V3& a) { return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z); }

// Dense row-major small matrix.
template <int R, int C>
struct M {
    double a[R * C];
    M() { std::memset(a, 0, sizeof(a)); }
    double& operator()(int r, int c) { return a[r * C + c]; }
    double operator()(int r, int c) const { return a[r * C + c]; }
    static M identity() { M m; for (int i = 0; i < std::min(R, C); ++i) m(i, i) = 1; return m; }
};
using M3 = M<3, 3>;
using M23 = M<2, 3>;

template <int R, int K, int C>
inline M<R, C> operator*(const M<R, K>& a, const M<K, C>& b) {
    M<R, C> o;
    for (int r = 0; r < R; ++r)
        for (int c = 0; c < C; ++c) {
            double s = 0;
            for (int k = 0; k < K; ++k) s += a(r, k) * b(k, c);
            o(r, c) = s;
        }
    return o;
}
template <int R, int C>
inline M<R, C> operator+(const M<R, C>& a, const M<R, C>& b) { M<R, C> o; for (int i = 0; i < R * C; ++i) o.a[i] = a.a[i] + b.a[i]; return o; }
template <int R, int C>
inline M<R, C> operator-(const M<R, C>& a, const M<R, C>& b) { M<R, C> o; for (int i = 0; i < R * C; ++i) o.a[i] = a.a[i] - b.a[i]; return o; }
template <int R, int C>
inline M<R, C> operator-(const M<R, C>& a) { M<R, C> o; for (int i = 0; i < R * C; ++i) o.a[i] = -a.a[i]; return o; }
template <int R, int C>
inline M<R, C> operator*(double s, const M<R, C>& a) { M<R, C> o; for (int i = 0; i < R * C; ++i) o.a[i] = s * a.a[i]; return o; }
template <int R, int C>
inline M<C, R> transpose(const M<R, C>& a) { M<C, R> o; for (int r = 0; r < R; ++r) for (int c = 0; c < C; ++c) o(c, r) = a(r, c); return o; }
template <int R, int C>
inline bool finite(const M<R, C>& a) { for (int i = 0; i < R * C; ++i) if (!std::isfinite(a.a[i])) return false; return true; }

inline V3 operator*(const M3& m, const V3& v) {
    return {m(0, 0) * v.x + m(0, 1) * v.y + m(0, 2) * v.z,
            m(1, 0) * v.x + m(1, 1) * v.y + m(1, 2) * v.z,
            m(2, 0) * v.x + m(2, 1) * v.y + m(2, 2) * v.z};
}
inline void mul23(const M23& m, const V3& v, double* out) {
    out[0] = m(0, 0) * v.x + m(0, 1) * v.y + m(0, 2) * v.z;
    out[1] = m(1, 0) * v.x + m(1, 1) * v.y + m(1, 2) * v.z;
}
inline M3 skew(const V3& v) {
    M3 m;
    m(0, 1) = -v.z; m(0, 2) = v.y; m(1, 0) = v.z; m(1, 2) = -v.x; m(2, 0) = -v.y; m(2, 1) = v.x;
    return m;
}
inline M3 outer(const V3& a, const V3& b) {
    M3 m; for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) m(r, c) = a[r] * b[c]; return m;
}
inline M3 diag3(double s) { M3 m; m(0, 0) = m(1, 1) = m(2, 2) = s; return m; }
inline M3 fromPtr(const double* p) { M3 m; for (int i = 0; i < 9; ++i) m.a[i] = p[i]; return m; }

// Eigen-convention quaternion (Hamilton product, storage x y z w).
struct Q {
    double x = 0, y = 0, z = 0, w = 1;
    Q() = default;
    Q(double w_, double x_, double y_, double z_) : x(x_), y(y_), z(z_), w(w_) {}
    static Q fromXyzw(const double* p) { return Q(p[3], p[0], p[1], p[2]); }
    void storeXyzw(double* p) const { p[0] = x; p[1] = y; p[2] = z; p[3] = w; }
    V3 vec() const { return {x, y, z}; }
    double sqnorm() const { return x * x + y * y + z * z + w * w; }
    Q normalized() const { const double n = std::sqrt(sqnorm()); return Q(w / n, x / n, y / n, z / n); }
    Q conjugate() const { return Q(w, -x, -y, -z); }
    M3 matrix() const {  // Eigen QuaternionBase::toRotationMatrix
        const double tx = 2 * x, ty = 2 * y, tz = 2 * z;
        const double twx = tx * w, twy = ty * w, twz = tz * w;
        const double txx = tx * x, txy = ty * x, txz = tz * x;
        const double tyy = ty * y, tyz = tz * y, tzz = tz * z;
        M3 r;
        r(0, 0) = 1 - (tyy + tzz); r(0, 1) = txy - twz; r(0, 2) = txz + twy;
        r(1, 0) = txy + twz; r(1, 1) = 1 - (txx + tzz); r(1, 2) = tyz - twx;
        r(2, 0) = txz - twy; r(2, 1) = tyz + twx; r(2, 2) = 1 - (txx + tyy);
        return r;
    }
};
inline Q operator*(const Q& a, const Q& b) {
    return Q(a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z,
             a.w * b.x + b.w * a.x + a.y * b.z - a.z * b.y,
             a.w * b.y + b.w * a.y + a.z * b.x - a.x * b.z,
             a.w * b.z + b.w * a.z + a.x * b.y - a.y * b.x);
}
// Eigen QuaternionBase::_transformVector (q * v)
inline V3 rotate(const Q& q, const V3& v) {
    V3 uv = cross(q.vec(), v);
    uv = uv + uv;
    return v + q.w * uv + cross(q.vec(), uv);
}

}  // namespace oracle
