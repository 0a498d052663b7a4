// Host-side measurement preparation: IMU preintegration into the frozen record consumed by the
// device IMU factor, and pinhole observation whitening into S^2 bearings.
//
//   preintegration   ImuError::redoPreintegration (src/factor/ImuError.cpp:254-462), midpoint
//                    rule, 15x15 covariance propagated per noise source (4 sensitivity
//                    matrices), whitening W = Lambda^-1/2 U^T (PseudoInverse::symmSqrtU,
//                    src/factor/PseudoInverse.hpp:116-145)
//   whitening        FrameStore::observation (src/backend/FrameStore.cpp:102-124) with the
//                    undistorted pinhole back-projection (src/cameras/PinholeCamera.hpp:778-807)
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "../../include/vib.h"

namespace {

struct Mat3 {
    double a[9];
    static Mat3 zero() { Mat3 m; std::memset(m.a, 0, sizeof(m.a)); return m; }
    static Mat3 eye() { Mat3 m = zero(); m.a[0] = m.a[4] = m.a[8] = 1; return m; }
    double& operator()(int r, int c) { return a[3 * r + c]; }
    double operator()(int r, int c) const { return a[3 * r + c]; }
};
Mat3 operator*(const Mat3& x, const Mat3& y) {
    Mat3 o;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) o(r, c) = x(r, 0) * y(0, c) + x(r, 1) * y(1, c) + x(r, 2) * y(2, c);
    return o;
}
Mat3 operator+(const Mat3& x, const Mat3& y) { Mat3 o; for (int i = 0; i < 9; ++i) o.a[i] = x.a[i] + y.a[i]; return o; }
Mat3 operator-(const Mat3& x, const Mat3& y) { Mat3 o; for (int i = 0; i < 9; ++i) o.a[i] = x.a[i] - y.a[i]; return o; }
Mat3 scale(double s, const Mat3& x) { Mat3 o; for (int i = 0; i < 9; ++i) o.a[i] = s * x.a[i]; return o; }
void mv(const Mat3& m, const double* v, double* o) {
    for (int r = 0; r < 3; ++r) o[r] = m(r, 0) * v[0] + m(r, 1) * v[1] + m(r, 2) * v[2];
}
Mat3 skew(const double* v) {
    Mat3 m = Mat3::zero();
    m(0, 1) = -v[2]; m(0, 2) = v[1]; m(1, 0) = v[2]; m(1, 2) = -v[0]; m(2, 0) = -v[1]; m(2, 1) = v[0];
    return m;
}

struct Quat { double w, x, y, z; };
Quat qmul(const Quat& a, const Quat& b) {
    return {a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z, a.w * b.x + b.w * a.x + a.y * b.z - a.z * b.y,
            a.w * b.y + b.w * a.y + a.z * b.x - a.x * b.z, a.w * b.z + b.w * a.z + a.x * b.y - a.y * b.x};
}
Mat3 qmat(const Quat& q) {  // Eigen toRotationMatrix convention
    const double tx = 2 * q.x, ty = 2 * q.y, tz = 2 * q.z;
    const double twx = tx * q.w, twy = ty * q.w, twz = tz * q.w, txx = tx * q.x, txy = ty * q.x, txz = tz * q.x;
    const double tyy = ty * q.y, tyz = tz * q.y, tzz = tz * q.z;
    Mat3 r;
    r(0, 0) = 1 - (tyy + tzz); r(0, 1) = txy - twz; r(0, 2) = txz + twy;
    r(1, 0) = txy + twz; r(1, 1) = 1 - (txx + tzz); r(1, 2) = tyz - twx;
    r(2, 0) = txz - twy; r(2, 1) = tyz + twx; r(2, 2) = 1 - (txx + tyy);
    return r;
}

double sinc(double x) {  // ode::sinc, src/factor/ode.hpp:15-27
    if (std::fabs(x) > 1e-6) return std::sin(x) / x;
    const double x2 = x * x;
    return 1.0 - x2 / 6.0 + x2 * x2 / 120.0 - x2 * x2 * x2 / 5040.0;
}

Mat3 right_jacobian(const double* phi) {  // Forster et al. RSS 2015 eq. (8)
    const double n = std::sqrt(phi[0] * phi[0] + phi[1] * phi[1] + phi[2] * phi[2]);
    const Mat3 px = skew(phi), px2 = px * px;
    if (n < 1.0e-4) return Mat3::eye() + scale(-0.5, px) + scale(1.0 / 6.0, px2);
    return Mat3::eye() + scale(-(1.0 - std::cos(n)) / (n * n), px) + scale((n - std::sin(n)) / (n * n * n), px2);
}

// 15x15 covariance propagation helper: dP <- F dP F^T + Q
void propagate(const double* F, double* P, const double* Q) {
    double T[225], O[225];
    for (int i = 0; i < 15; ++i)
        for (int j = 0; j < 15; ++j) {
            double s = 0;
            for (int k = 0; k < 15; ++k) s += F[i * 15 + k] * P[k * 15 + j];
            T[i * 15 + j] = s;
        }
    for (int i = 0; i < 15; ++i)
        for (int j = 0; j < 15; ++j) {
            double s = 0;
            for (int k = 0; k < 15; ++k) s += T[i * 15 + k] * F[j * 15 + k];
            O[i * 15 + j] = s + Q[i * 15 + j];
        }
    std::memcpy(P, O, sizeof(O));
}

}  // namespace

namespace vib {

// Cyclic Jacobi eigen-decomposition of a symmetric n x n matrix (row-major); eigenvalues
// ascending, eigenvectors as columns of V.
void symmetric_eigen(std::vector<double> A, int n, std::vector<double>& ev, std::vector<double>& V) {
    V.assign((size_t)n * n, 0);
    for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1;
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0, tot = 0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                const double v = A[(size_t)i * n + j] * A[(size_t)i * n + j];
                tot += v;
                if (i != j) off += v;
            }
        if (off == 0 || off <= 1e-32 * tot) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = A[(size_t)p * n + q];
                if (apq == 0) continue;
                const double theta = (A[(size_t)q * n + q] - A[(size_t)p * n + p]) / (2 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
                const double c = 1 / std::sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < n; ++k) {
                    const double x = A[(size_t)k * n + p], y = A[(size_t)k * n + q];
                    A[(size_t)k * n + p] = c * x - s * y;
                    A[(size_t)k * n + q] = s * x + c * y;
                }
                for (int k = 0; k < n; ++k) {
                    const double x = A[(size_t)p * n + k], y = A[(size_t)q * n + k];
                    A[(size_t)p * n + k] = c * x - s * y;
                    A[(size_t)q * n + k] = s * x + c * y;
                }
                for (int k = 0; k < n; ++k) {
                    const double x = V[(size_t)k * n + p], y = V[(size_t)k * n + q];
                    V[(size_t)k * n + p] = c * x - s * y;
                    V[(size_t)k * n + q] = s * x + c * y;
                }
            }
    }
    ev.assign(n, 0);
    for (int i = 0; i < n; ++i) ev[i] = A[(size_t)i * n + i];
    // selection sort ascending
    for (int i = 0; i < n; ++i) {
        int m = i;
        for (int j = i + 1; j < n; ++j) if (ev[j] < ev[m]) m = j;
        if (m != i) {
            std::swap(ev[i], ev[m]);
            for (int k = 0; k < n; ++k) std::swap(V[(size_t)k * n + i], V[(size_t)k * n + m]);
        }
    }
}

}  // namespace vib

extern "C" {

int vib_preintegrate(const double* smp, int32_t count, double t0, double t1, const double* sbl,
                     const vib_imu_params* prm, const double* gyro_start, const double* gyro_end, double* out) {
    if (!smp || !sbl || !prm || !gyro_start || !gyro_end || !out || count < 2 || !(t1 > t0)) return VIB_ERR_INVALID_ARGUMENT;
    const int S = VIB_IMU_SAMPLE_STRIDE;
    if (!(smp[0] <= t0) || !(smp[(count - 1) * S] >= t1)) return VIB_ERR_INVALID_ARGUMENT;
    Quat Dq{1, 0, 0, 0};
    Mat3 Cint = Mat3::zero(), Cdint = Mat3::zero(), cr = Mat3::zero(), dadbg = Mat3::zero(), dvdbg = Mat3::zero(),
         dpdbg = Mat3::zero();
    double vint[3] = {0, 0, 0}, pint[3] = {0, 0, 0};
    std::vector<double> dP(4 * 225, 0.0);
    double time = t0;
    bool started = false;
    int steps = 0;
    for (int i = 0; i + 1 < count; ++i) {
        const double* s0 = smp + i * S;
        const double* s1 = smp + (i + 1) * S;
        double w0[3] = {s0[1], s0[2], s0[3]}, a0[3] = {s0[4], s0[5], s0[6]};
        double w1[3] = {s1[1], s1[2], s1[3]}, a1[3] = {s1[4], s1[5], s1[6]};
        double next = s1[0];
        double dt = next - time;
        if (t1 < next) {
            const double span = next - s0[0];
            next = t1;
            dt = next - time;
            const double r = dt / span;
            for (int k = 0; k < 3; ++k) { w1[k] = (1 - r) * w0[k] + r * w1[k]; a1[k] = (1 - r) * a0[k] + r * a1[k]; }
        }
        if (dt <= 0.0) continue;
        if (!started) {
            started = true;
            const double r = dt / (next - s0[0]);
            for (int k = 0; k < 3; ++k) { w0[k] = r * w0[k] + (1 - r) * w1[k]; a0[k] = r * a0[k] + (1 - r) * a1[k]; }
        }
        double gsat = 1, asat = 1;
        for (int k = 0; k < 3; ++k) {
            if (std::fabs(w0[k]) > prm->g_max || std::fabs(w1[k]) > prm->g_max) gsat = 100;
            if (std::fabs(a0[k]) > prm->a_max || std::fabs(a1[k]) > prm->a_max) asat = 100;
        }
        double om[3], ac[3];
        for (int k = 0; k < 3; ++k) {
            om[k] = 0.5 * (w0[k] + w1[k]) - sbl[3 + k];
            ac[k] = 0.5 * (a0[k] + a1[k]) - sbl[6 + k];
        }
        const double half = 0.5 * dt * std::sqrt(om[0] * om[0] + om[1] * om[1] + om[2] * om[2]);
        const double sc = sinc(half) * 0.5 * dt;
        const Quat dq{std::cos(half), sc * om[0], sc * om[1], sc * om[2]};
        const Quat Dq1 = qmul(Dq, dq);
        const Mat3 C0 = qmat(Dq), C1 = qmat(Dq1), CC = C0 + C1;
        double CCa[3];
        mv(CC, ac, CCa);
        const Mat3 Cint1 = Cint + scale(0.5 * dt, CC);
        double vint1[3];
        for (int k = 0; k < 3; ++k) vint1[k] = vint[k] + 0.5 * dt * CCa[k];
        Cdint = Cdint + scale(dt, Cint) + scale(0.25 * dt * dt, CC);
        for (int k = 0; k < 3; ++k) pint[k] += dt * vint[k] + 0.25 * dt * dt * CCa[k];
        double odt[3] = {om[0] * dt, om[1] * dt, om[2] * dt};
        const Mat3 Jr = right_jacobian(odt);
        dadbg = dadbg + scale(dt, C1 * Jr);
        const Quat dqi{dq.w, -dq.x, -dq.y, -dq.z};
        const Mat3 cr1 = qmat(dqi) * cr + scale(dt, Jr);
        const Mat3 ax = skew(ac);
        const Mat3 mix = C0 * ax * cr + C1 * ax * cr1;
        const Mat3 dvdbg1 = dvdbg + scale(0.5 * dt, mix);
        dpdbg = dpdbg + scale(dt, dvdbg) + scale(0.25 * dt * dt, mix);
        double F[225];
        for (int k = 0; k < 225; ++k) F[k] = (k % 16 == 0) ? 1.0 : 0.0;
        auto put = [&](int r0, int c0, const Mat3& m) {
            for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) F[(r0 + r) * 15 + c0 + c] = m(r, c);
        };
        double tmp[3];
        for (int k = 0; k < 3; ++k) tmp[k] = vint[k] * dt + 0.25 * dt * dt * CCa[k];
        put(0, 3, scale(-1, skew(tmp)));
        put(0, 6, scale(dt, Mat3::eye()));
        put(0, 9, scale(dt, dvdbg) + scale(0.25 * dt * dt, mix));
        put(0, 12, scale(-dt, Cint) + scale(0.25 * dt * dt, CC));
        put(3, 9, scale(-dt, C1));
        for (int k = 0; k < 3; ++k) tmp[k] = 0.5 * dt * CCa[k];
        put(6, 3, scale(-1, skew(tmp)));
        put(6, 9, scale(0.5 * dt, mix));
        put(6, 12, scale(-0.5 * dt, CC));
        double Q[4][225];
        std::memset(Q, 0, sizeof(Q));
        for (int k = 0; k < 3; ++k) {
            Q[0][(3 + k) * 16] = gsat * dt;
            Q[1][k * 16] = 0.5 * dt * dt * dt * asat * asat * asat;
            Q[1][(6 + k) * 16] = asat * dt;
            Q[2][(9 + k) * 16] = dt;
            Q[3][(12 + k) * 16] = dt;
        }
        for (int j = 0; j < 4; ++j) propagate(F, dP.data() + 225 * j, Q[j]);
        Dq = Dq1; Cint = Cint1;
        for (int k = 0; k < 3; ++k) vint[k] = vint1[k];
        cr = cr1; dvdbg = dvdbg1;
        time = next;
        ++steps;
        if (next == t1) break;
    }
    if (steps <= 0) return VIB_ERR_INVALID_ARGUMENT;
    const double sig[4] = {prm->sigma_g_c, prm->sigma_a_c, prm->sigma_gw_c, prm->sigma_aw_c};
    std::vector<double> P(225, 0.0);
    for (int j = 0; j < 4; ++j)
        for (int r = 0; r < 15; ++r)
            for (int c = 0; c < 15; ++c)
                P[r * 15 + c] += (0.5 * dP[225 * j + r * 15 + c] + 0.5 * dP[225 * j + c * 15 + r]) * sig[j] * sig[j];
    std::vector<double> ev, U;
    vib::symmetric_eigen(P, 15, ev, U);
    const double eps = std::numeric_limits<double>::epsilon();
    const double tol = std::max(eps, eps * 15 * ev[14]);
    std::memset(out, 0, sizeof(double) * VIB_IMU_STRIDE);
    out[0] = t1 - t0;
    out[1] = prm->g;
    for (int k = 0; k < 3; ++k) out[2 + k] = prm->p_BS[k];
    for (int k = 0; k < 4; ++k) out[5 + k] = prm->q_BS[k];
    for (int k = 0; k < 3; ++k) { out[9 + k] = gyro_start[k]; out[12 + k] = gyro_end[k]; }
    out[15] = Dq.x; out[16] = Dq.y; out[17] = Dq.z; out[18] = Dq.w;
    for (int k = 0; k < 3; ++k) { out[19 + k] = vint[k]; out[22 + k] = pint[k]; }
    for (int k = 0; k < 6; ++k) out[25 + k] = sbl[3 + k];
    for (int k = 0; k < 9; ++k) {
        out[31 + k] = dadbg.a[k]; out[40 + k] = dvdbg.a[k]; out[49 + k] = dpdbg.a[k];
        out[58 + k] = Cint.a[k]; out[67 + k] = Cdint.a[k];
    }
    for (int i = 0; i < 15; ++i) {
        const double s = std::sqrt(ev[i] > tol ? 1.0 / ev[i] : 1.0 / tol);
        for (int k = 0; k < 15; ++k) out[76 + i * 15 + k] = s * U[(size_t)k * 15 + i];
    }
    return VIB_OK;
}

int vib_pinhole_observations(double fx, double fy, double cx, double cy, double sigma, int32_t count,
                             const double* px, double* bearings, double* sqrt_info) {
    if (!px || !bearings || !sqrt_info || count < 0 || !(fx > 0) || !(fy > 0) || !(sigma > 0)) return VIB_ERR_INVALID_ARGUMENT;
    for (int i = 0; i < count; ++i) {
        const double ray[3] = {(px[2 * i] - cx) / fx, (px[2 * i + 1] - cy) / fy, 1.0};
        const double rn = std::sqrt(ray[0] * ray[0] + ray[1] * ray[1] + ray[2] * ray[2]);
        const double b[3] = {ray[0] / rn, ray[1] / rn, ray[2] / rn};
        // sphere basis of b
        const double ax[3] = {std::fabs(b[0]) < 0.8 ? 1.0 : 0.0, std::fabs(b[0]) < 0.8 ? 0.0 : 1.0, 0.0};
        const double bd = b[0] * ax[0] + b[1] * ax[1] + b[2] * ax[2];
        double c0[3] = {ax[0] - b[0] * bd, ax[1] - b[1] * bd, ax[2] - b[2] * bd};
        const double c0n = std::sqrt(c0[0] * c0[0] + c0[1] * c0[1] + c0[2] * c0[2]);
        for (double& v : c0) v /= c0n;
        const double c1[3] = {b[1] * c0[2] - b[2] * c0[1], b[2] * c0[0] - b[0] * c0[2], b[0] * c0[1] - b[1] * c0[0]};
        // D = (I - b b^T) J / |ray| with J = [1/fx 0; 0 1/fy; 0 0]; M = B^T D (2x2)
        double D[3][2];
        const double J[3][2] = {{1 / fx, 0}, {0, 1 / fy}, {0, 0}};
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 2; ++c) {
                double s = 0;
                for (int k = 0; k < 3; ++k) s += ((r == k ? 1.0 : 0.0) - b[r] * b[k]) * J[k][c];
                D[r][c] = s / rn;
            }
        double M[2][2];
        for (int c = 0; c < 2; ++c) {
            M[0][c] = c0[0] * D[0][c] + c0[1] * D[1][c] + c0[2] * D[2][c];
            M[1][c] = c1[0] * D[0][c] + c1[1] * D[1][c] + c1[2] * D[2][c];
        }
        const double s2 = sigma * sigma;
        const double cov00 = s2 * (M[0][0] * M[0][0] + M[0][1] * M[0][1]);
        const double cov10 = s2 * (M[1][0] * M[0][0] + M[1][1] * M[0][1]);
        const double cov11 = s2 * (M[1][0] * M[1][0] + M[1][1] * M[1][1]);
        const double l00 = std::sqrt(cov00), l10 = cov10 / l00, l11 = std::sqrt(cov11 - l10 * l10);
        if (!(l00 > 0) || !(l11 > 0)) return VIB_ERR_INVALID_ARGUMENT;
        for (int k = 0; k < 3; ++k) bearings[3 * i + k] = b[k];
        sqrt_info[3 * i] = 1 / l00;
        sqrt_info[3 * i + 1] = -l10 / (l00 * l11);
        sqrt_info[3 * i + 2] = 1 / l11;
    }
    return VIB_OK;
}

}  // extern "C"
