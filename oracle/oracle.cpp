// ORACLE -- test infrastructure only (see oracle_math.hpp). FP64 CPU restatement of the
// reference `backend_visual_batch: schur` path. Each function cites the reference code it
// follows. Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg only.
//
// Parity pins (checked in tests/test_oracle.py): zero residual at noiseless truth, central
// finite differences of every Jacobian on the left-SE3 tangent, closed-form d r / d rho,
// same-epoch zero pose/rig Jacobians, td Jacobian zero for static bearings, the FD numeric
// factor path of WindowVisualFactor.cpp:578-611, dense brute-force normal equations and
// Schur solves, Huber closed forms, IMU residual ~0 on analytic motion and FD IMU Jacobians,
// linear-quadratic marginalization agreement, eigen-cutoff invariants, Cholesky/Jacobi vs
// numpy/scipy. See the test file for each.
#include "oracle_math.hpp"
#include "../include/vib.h"

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <limits>
#include <map>
#include <set>
#include <string>
#include <thread>
#include <vector>

namespace oracle {
namespace {

const double kInf = std::numeric_limits<double>::infinity();

// ------------------------------------------------------------------------------------
// Window copy (owning) of the C-ABI vib_window.
// ------------------------------------------------------------------------------------
struct Window {
    int K = 0, C = 0, L = 0, F = 0, E = 0;
    uint32_t flags = 0;
    double huber = 1.0, minRho = 1e-8, tdLower = -kInf, tdUpper = kInf;
    std::vector<int64_t> frameId;
    std::vector<double> pose, sb;  // K*7, K*9
    std::vector<uint8_t> poseFree, sbWidth;
    double rig[7]{}, td = 0;
    std::vector<double> camR, camT;
    std::vector<int> lmFrame, lmCam;
    std::vector<double> lmBearing, lmVel, lmTref, rho;
    std::vector<uint8_t> lmFree;
    std::vector<int> obsLm, obsFrame, obsCam;
    std::vector<double> obsBearing, obsVel, obsTref, obsSqrt;
    std::vector<int> imuFrom, imuTo;
    std::vector<double> imu;
    int priorRows = 0, priorBlocks = 0;
    std::vector<int> priorKind;
    std::vector<int64_t> priorId;
    std::vector<double> priorRef, priorA, priorR0;
};

template <class T>
std::vector<T> copyArr(const T* p, size_t n) { return p ? std::vector<T>(p, p + n) : std::vector<T>(); }

Window load(const vib_window& w) {
    Window o;
    o.K = w.num_frames; o.C = w.num_cameras; o.L = w.num_landmarks; o.F = w.num_observations; o.E = w.num_imu;
    o.flags = w.flags; o.huber = w.huber_delta; o.minRho = w.min_inverse_depth;
    o.tdLower = w.td_lower; o.tdUpper = w.td_upper;
    o.frameId = copyArr(w.frame_id, o.K);
    o.pose = copyArr(w.pose, 7 * o.K); o.sb = copyArr(w.speed_bias, 9 * o.K);
    o.poseFree = w.pose_free ? copyArr(w.pose_free, o.K) : std::vector<uint8_t>(o.K, 1);
    o.sbWidth = w.sb_width ? copyArr(w.sb_width, o.K) : std::vector<uint8_t>(o.K, 9);
    std::copy_n(w.rig, 7, o.rig); o.td = *w.td;
    o.camR = copyArr(w.cam_rotation, 9 * o.C); o.camT = copyArr(w.cam_translation, 3 * o.C);
    o.lmFrame = copyArr(w.lm_anchor_frame, o.L); o.lmCam = copyArr(w.lm_anchor_cam, o.L);
    o.lmBearing = copyArr(w.lm_bearing, 3 * o.L); o.lmVel = copyArr(w.lm_velocity, 3 * o.L);
    o.lmTref = copyArr(w.lm_tref, o.L); o.rho = copyArr(w.lm_rho, o.L);
    o.lmFree = w.lm_free ? copyArr(w.lm_free, o.L) : std::vector<uint8_t>(o.L, 1);
    o.obsLm = copyArr(w.obs_landmark, o.F); o.obsFrame = copyArr(w.obs_frame, o.F); o.obsCam = copyArr(w.obs_cam, o.F);
    o.obsBearing = copyArr(w.obs_bearing, 3 * o.F); o.obsVel = copyArr(w.obs_velocity, 3 * o.F);
    o.obsTref = copyArr(w.obs_tref, o.F); o.obsSqrt = copyArr(w.obs_sqrt_info, 3 * o.F);
    o.imuFrom = copyArr(w.imu_from, o.E); o.imuTo = copyArr(w.imu_to, o.E);
    o.imu = copyArr(w.imu_data, (size_t)VIB_IMU_STRIDE * o.E);
    o.priorRows = w.prior_rows; o.priorBlocks = w.prior_rows > 0 ? w.prior_blocks : 0;
    o.priorKind = copyArr(w.prior_key_kind, o.priorBlocks); o.priorId = copyArr(w.prior_key_id, o.priorBlocks);
    int refSize = 0, cols = 0;
    for (int b = 0; b < o.priorBlocks; ++b) {
        const int k = o.priorKind[b];
        refSize += k == VIB_KEY_POSE || k == VIB_KEY_RIG ? 7 : (k == VIB_KEY_SPEED_BIAS ? 9 : 1);
        cols += k == VIB_KEY_POSE || k == VIB_KEY_RIG ? 6 : (k == VIB_KEY_SPEED_BIAS ? 9 : 1);
    }
    o.priorRef = copyArr(w.prior_ref, refSize);
    o.priorA = copyArr(w.prior_A, (size_t)o.priorRows * cols);
    o.priorR0 = copyArr(w.prior_r0, o.priorRows);
    return o;
}

// ------------------------------------------------------------------------------------
// Column layout: [pose_k (6, free), rig (6), td (1), sb_k (width)] then free rho.
// ------------------------------------------------------------------------------------
struct Layout {
    std::vector<int> poseCol, sbCol, lmCol;
    int rigCol = -1, tdCol = -1, nv = 0, nr = 0, ne = 0;
};
Layout layoutOf(const Window& w) {
    Layout l;
    int c = 0;
    l.poseCol.assign(w.K, -1); l.sbCol.assign(w.K, -1); l.lmCol.assign(w.L, -1);
    for (int k = 0; k < w.K; ++k) if (w.poseFree[k]) { l.poseCol[k] = c; c += 6; }
    if (w.flags & VIB_RIG_FREE) { l.rigCol = c; c += 6; }
    if (w.flags & VIB_TD_FREE) { l.tdCol = c; c += 1; }
    l.nv = c;
    for (int k = 0; k < w.K; ++k) if (w.sbWidth[k]) { l.sbCol[k] = c; c += w.sbWidth[k]; }
    l.nr = c;
    for (int i = 0; i < w.L; ++i) if (w.lmFree[i]) { l.lmCol[i] = c; c += 1; }
    l.ne = c - l.nr;
    return l;
}

// ------------------------------------------------------------------------------------
// Manifolds (src/backend/LeftSE3Manifold.cpp, src/backend/GpuRetainedFactorMath.hpp:14-51)
// ------------------------------------------------------------------------------------
Q rotationExp(const V3& angle) {  // GpuRetainedFactorMath.hpp:14-18
    const double a = norm(angle);
    const double s = a < 1e-8 ? 0.5 - a * a / 48.0 : std::sin(0.5 * a) / a;
    return Q(std::cos(0.5 * a), s * angle.x, s * angle.y, s * angle.z);
}
M3 rotationJacobian(const V3& angle) {  // GpuRetainedFactorMath.hpp:20-26
    const double a2 = sqnorm(angle), a = std::sqrt(a2);
    const M3 w = skew(angle);
    return M3::identity() + (a2 < 1e-10 ? 0.5 - a2 / 24 : (1 - std::cos(a)) / a2) * w +
           (a2 < 1e-10 ? 1.0 / 6 - a2 / 120 : (a - std::sin(a)) / (a2 * a)) * (w * w);
}
// y = x [+] d (left SE3, GpuRetainedFactorMath.hpp:28-37)
void posePlus(const double* x, const double* d, double* y, bool rotationOnly) {
    const V3 angle(d + (rotationOnly ? 0 : 3));
    const Q q = rotationExp(angle);
    V3 yp(x);
    if (!rotationOnly) yp = rotate(q, V3(x)) + rotationJacobian(angle) * V3(d);
    const Q yq = (q * Q::fromXyzw(x + 3)).normalized();
    yp.store(y);
    yq.storeXyzw(y + 3);
}
// d = x [-] ref (GpuRetainedFactorMath.hpp:39-51)
void poseMinus(const double* x, const double* ref, double* d) {
    Q q = Q::fromXyzw(x + 3) * Q::fromXyzw(ref + 3).conjugate();
    q = q.normalized();
    if (q.w < 0) q = Q(-q.w, -q.x, -q.y, -q.z);
    const double s = norm(q.vec());
    const V3 theta = (s < 1e-10 ? 2.0 : 2.0 * std::atan2(s, q.w) / s) * q.vec();
    const double a2 = sqnorm(theta), a = std::sqrt(a2);
    const M3 w = skew(theta);
    const double c = a2 < 1e-10 ? 1.0 / 12 + a2 / 720 : (1 - 0.5 * a / std::tan(0.5 * a)) / a2;
    const V3 p = (M3::identity() - 0.5 * w + c * (w * w)) * (V3(x) - rotate(q, V3(ref)));
    p.store(d);
    theta.store(d + 3);
}
// LeftSE3Manifold::PlusJacobian (LeftSE3Manifold.cpp:95-105), 7x6 row-major.
void plusJacobian(const double* x, double* J) {
    for (int i = 0; i < 42; ++i) J[i] = 0;
    const M3 s = skew(V3(x));
    for (int r = 0; r < 3; ++r) {
        J[r * 6 + r] = 1;
        for (int c = 0; c < 3; ++c) J[r * 6 + 3 + c] = -s(r, c);
    }
    const Q q = Q::fromXyzw(x + 3);
    const M3 top = 0.5 * (diag3(q.w) - skew(q.vec()));
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) J[(3 + r) * 6 + 3 + c] = top(r, c);
    J[6 * 6 + 3] = -0.5 * q.x; J[6 * 6 + 4] = -0.5 * q.y; J[6 * 6 + 5] = -0.5 * q.z;
}

// ------------------------------------------------------------------------------------
// S^2 visual factor (VisualSphereBatchMath.hpp:29-174, WindowVisualFactor.cpp:324-410)
// ------------------------------------------------------------------------------------
M<3, 2> sphereBasis(const V3& b) {  // WindowVisualFactor.cpp:324-328
    const V3 axis = std::fabs(b.x) < 0.8 ? V3(1, 0, 0) : V3(0, 1, 0);
    const V3 c0 = normalized(axis - b * dot(b, axis));
    const V3 c1 = cross(b, c0);
    M<3, 2> m;
    for (int r = 0; r < 3; ++r) { m(r, 0) = c0[r]; m(r, 1) = c1[r]; }
    return m;
}
// measurement_ = sqrtInformation * basis^T (WindowVisualFactor.cpp:407)
M23 measurementMatrix(const double* bearing, const double* s) {
    const M<3, 2> B = sphereBasis(V3(bearing));
    M<2, 2> S; S(0, 0) = s[0]; S(1, 0) = s[1]; S(1, 1) = s[2];
    return S * transpose(B);
}
V3 timeBearing(const V3& raw, const V3& v, double dt, bool differentiate, V3& derivative) {
    if (!differentiate) {  // VisualSphereBatchMath.hpp:32-37
        const V3 u = v * dt;
        const double angle = norm(u);
        if (angle < 1e-8) return normalized(raw + u);
        return std::cos(angle) * raw + (std::sin(angle) / angle) * u;
    }
    const double speed = norm(v), angle = speed * dt;  // :38-48
    if (std::fabs(angle) < 1e-8) {
        const V3 direction = raw + v * dt;
        const double length = norm(direction);
        const V3 unit = direction / length;
        derivative = -((M3::identity() - outer(unit, unit)) * v) / length;
        return unit;
    }
    const double sine = std::sin(angle), cosine = std::cos(angle);
    derivative = speed * sine * raw - cosine * v;
    return cosine * raw + (sine / speed) * v;
}
// Quaternion lift of writePose (VisualSphereBatchMath.hpp:51-63). J 2x7 row-major.
bool writePose(const double* pose, double* out, const M23& translation, const M23& rotation) {
    if (!out) return true;
    const Q q = Q::fromXyzw(pose + 3);
    if (!std::isfinite(q.sqnorm()) || std::fabs(q.sqnorm() - 1) > 1e-6) return false;
    M<3, 4> lift;
    const M3 top = 2.0 * (diag3(q.w) + skew(q.vec()));
    for (int r = 0; r < 3; ++r) { for (int c = 0; c < 3; ++c) lift(r, c) = top(r, c); }
    lift(0, 3) = -2 * q.x; lift(1, 3) = -2 * q.y; lift(2, 3) = -2 * q.z;
    const M<2, 4> rot = rotation * lift;
    for (int r = 0; r < 2; ++r) {
        for (int c = 0; c < 3; ++c) out[r * 7 + c] = translation(r, c);
        for (int c = 0; c < 4; ++c) out[r * 7 + 3 + c] = rot(r, c);
    }
    for (int i = 0; i < 14; ++i) if (!std::isfinite(out[i])) return false;
    return true;
}

// Ambient Jacobians of one visual factor: anchor pose (2x7), observer pose (2x7), rig (2x7),
// td (2), rho (2). Null pointers skip blocks (derivativeMask).
struct VisualJ {
    double anchor[14], observer[14], rig[14], td[2], rho[2];
};
// sphere_batch::evaluate (VisualSphereBatchMath.hpp:67-174)
bool visualEvaluate(const Window& w, int f, const double* poseA, const double* poseB, const double* rig,
                    double td, double rho, double* residual, VisualJ* J, bool wantAnchor, bool wantObserver,
                    bool wantRig, bool wantTd, bool wantRho) {
    const int l = w.obsLm[f];
    const int a = w.lmFrame[l], b = w.obsFrame[f];
    const bool sameEpoch = a == b;
    if (!(rho > 0) || !std::isfinite(rho) || !std::isfinite(td)) return false;
    const double* blocks7[3] = {poseA, sameEpoch ? nullptr : poseB, rig};
    for (auto* s : blocks7) {
        if (!s) continue;
        for (int i = 0; i < 7; ++i) if (!std::isfinite(s[i])) return false;
        if (s[3] * s[3] + s[4] * s[4] + s[5] * s[5] + s[6] * s[6] <= 1e-16) return false;
    }
    const bool jAnchor = J && wantAnchor, jObs = J && wantObserver && !sameEpoch, jRig = J && wantRig;
    const bool timeDerivative = J && wantTd;
    const V3 anchorRaw(&w.lmBearing[3 * l]), measured(&w.obsBearing[3 * f]);
    const V3 anchorVel(&w.lmVel[3 * l]), measuredVel(&w.obsVel[3 * f]);
    const bool movingAnchor = sqnorm(anchorVel) > 0, movingMeasured = sqnorm(measuredVel) > 0;
    V3 ba = normalized(anchorRaw), bb = normalized(measured), da, db;
    if (movingAnchor) ba = timeBearing(anchorRaw, anchorVel, w.lmTref[l] - td, timeDerivative, da);
    if (movingMeasured) bb = timeBearing(measured, measuredVel, w.obsTref[f] - td, timeDerivative, db);
    if (!finite(ba) || !finite(bb)) return false;
    const M3 Rca = fromPtr(&w.camR[9 * w.lmCam[l]]), Rcb = fromPtr(&w.camR[9 * w.obsCam[f]]);
    const V3 tca(&w.camT[3 * w.lmCam[l]]), tcb(&w.camT[3 * w.obsCam[f]]);
    const V3 local = Rca * (ba / rho) + tca;
    M3 Ra, Rb, Re, W, cameraRotation;
    V3 body, relative, rotatedBody, observerBody, pc, te;
    if (sameEpoch) {
        pc = transpose(Rcb) * (local - tcb);
        cameraRotation = transpose(Rcb) * Rca;
    } else {
        Ra = Q::fromXyzw(poseA + 3).normalized().matrix();
        Rb = Q::fromXyzw(poseB + 3).normalized().matrix();
        Re = Q::fromXyzw(rig + 3).normalized().matrix();
        te = V3(rig);
        W = transpose(Rb * Re * Rcb);
        cameraRotation = W * (Ra * Re * Rca);
        const V3 separation = V3(poseA) - V3(poseB);
        const V3 offset = W * (Ra * (Re * tca + te) - Rb * (Re * tcb + te) + separation);
        pc = cameraRotation * (ba / rho) + offset;
        if (jAnchor || jObs || jRig) {
            body = Re * local + te;
            rotatedBody = Ra * body;
            relative = rotatedBody + separation;
            observerBody = transpose(Rb) * relative;
        }
    }
    const double range = norm(pc);
    if (!finite(pc) || !std::isfinite(range) || range < 1e-10) return false;
    const V3 n = pc / range;
    const double c = std::fmin(1.0, std::fmax(-1.0, dot(bb, n))), s = norm(cross(bb, n));
    const double denominator = movingMeasured ? 1 + dot(measured, bb) : 2;
    if (c < -1 + 1e-8 || denominator < 1e-8) return false;
    const V3 u = n - c * bb;
    const double d = 1 - c;
    const bool series = c > 0 && s < 1e-3;
    const double theta = series ? 0 : std::atan2(s, c);
    const double fcoef = s < 1e-8 ? 1.0 : (series ? 1 + d / 3 + 2 * d * d / 15 + 2 * d * d * d / 35 : theta / s);
    const V3 logarithm = fcoef * u;
    const M23 measurement = measurementMatrix(&w.obsBearing[3 * f], &w.obsSqrt[3 * f]);
    M23 P = measurement;
    M3 K;
    if (movingMeasured) {
        K = skew(cross(measured, bb));
        const M3 transport = M3::identity() + K + (1.0 / denominator) * (K * K);
        P = measurement * transpose(transport);
    }
    mul23(P, logarithm, residual);
    if (!std::isfinite(residual[0]) || !std::isfinite(residual[1])) return false;
    if (!J) return true;
    const double g = s < 1e-8 ? 0.0 : (series ? (-1.0 / 3.0 - 4 * d / 15.0 - 6 * d * d / 35.0)
                                               : (theta * c - s) / (s * s * s));
    double Pu[2], Pbb[2];
    mul23(P, u, Pu); mul23(P, bb, Pbb);
    M23 projected;
    for (int r = 0; r < 2; ++r)
        for (int k = 0; k < 3; ++k)
            projected(r, k) = fcoef * P(r, k) + (g * Pu[r] - fcoef * Pbb[r]) * bb[k];
    double pn[2]; mul23(projected, n, pn);
    M23 reduce;
    for (int r = 0; r < 2; ++r)
        for (int k = 0; k < 3; ++k) reduce(r, k) = (projected(r, k) - pn[r] * n[k]) / range;
    if (sameEpoch) {
        if (jAnchor) for (int i = 0; i < 14; ++i) J->anchor[i] = 0;
        if (jRig) for (int i = 0; i < 14; ++i) J->rig[i] = 0;
    } else if (jAnchor || jObs || jRig) {
        const M23 D = reduce * W;
        if (jAnchor && !writePose(poseA, J->anchor, D, -(D * skew(rotatedBody)))) return false;
        if (jObs && !writePose(poseB, J->observer, -D, D * skew(relative))) return false;
        if (jRig && !writePose(rig, J->rig, D * (Ra - Rb),
                               D * ((-1.0) * (Ra * skew(body - te)) + Rb * skew(observerBody - te))))
            return false;
    }
    if (J && wantRho) {
        mul23(reduce * cameraRotation, -1.0 * ba / (rho * rho), J->rho);
        if (!std::isfinite(J->rho[0]) || !std::isfinite(J->rho[1])) return false;
    }
    if (timeDerivative) {
        if (!movingAnchor && !movingMeasured) { J->td[0] = J->td[1] = 0; }
        else {
            const M3 Jb = fcoef * ((-1.0) * outer(bb, n) - diag3(c)) + g * outer(u, n);
            const M3 dK = skew(cross(measured, db));
            const M3 dR = dK + (1.0 / denominator) * (dK * K + K * dK) -
                          (dot(measured, db) / (denominator * denominator)) * (K * K);
            double t1[2], t2[2], t3[2];
            mul23(reduce * cameraRotation, da / rho, t1);
            mul23(P * Jb, db, t2);
            mul23(measurement * transpose(dR), logarithm, t3);
            J->td[0] = t1[0] + t2[0] + t3[0];
            J->td[1] = t1[1] + t2[1] + t3[1];
            if (!std::isfinite(J->td[0]) || !std::isfinite(J->td[1])) return false;
        }
    }
    return true;
}

double huberWeight(double delta, double squared) {  // VisualCudaBatch.cu:55-57
    return delta > 0 && squared > delta * delta ? std::sqrt(delta / std::sqrt(squared)) : 1.0;
}
double robustCost(double delta, double squared) {  // VisualCudaBatch.cu:133
    return delta > 0 && squared > delta * delta ? delta * (std::sqrt(squared) - 0.5 * delta) : 0.5 * squared;
}

// ------------------------------------------------------------------------------------
// IMU factor from frozen preintegration (GpuRetainedFactorMath.hpp:53-140)
// ------------------------------------------------------------------------------------
struct Frozen {
    double duration, gravity;
    V3 pBS; Q qBS; V3 gyro[2];
    Q dq; V3 dv, dp; V3 refBg, refBa;
    M3 A, V, P, Va, Pa;
    double W[225];
};
Frozen unpackImu(const double* d) {
    Frozen f;
    f.duration = d[0]; f.gravity = d[1];
    f.pBS = V3(d + 2); f.qBS = Q::fromXyzw(d + 5);
    f.gyro[0] = V3(d + 9); f.gyro[1] = V3(d + 12);
    f.dq = Q::fromXyzw(d + 15); f.dv = V3(d + 19); f.dp = V3(d + 22);
    f.refBg = V3(d + 25); f.refBa = V3(d + 28);
    f.A = fromPtr(d + 31); f.V = fromPtr(d + 40); f.P = fromPtr(d + 49); f.Va = fromPtr(d + 58); f.Pa = fromPtr(d + 67);
    std::copy_n(d + 76, 225, f.W);
    return f;
}
struct SensorState { V3 p, v, arm, lever; Q q; };
SensorState sensorState(const Frozen& imu, const double* pose, const double* sb, int end) {  // :58-67
    const Q q = Q::fromXyzw(pose + 3).normalized();
    SensorState s;
    s.q = (q * imu.qBS).normalized();
    s.arm = rotate(q, imu.pBS);
    s.lever = cross(rotate(s.q, imu.gyro[end] - V3(sb + 3)), s.arm);
    s.p = V3(pose) + s.arm;
    s.v = V3(sb) + s.lever;
    return s;
}
// Writes 15 residuals and (if J) four 15x9 tangent blocks (pose blocks use 6 columns).
bool imuEvaluate(const Frozen& imu, const double* pose0, const double* sb0, const double* pose1,
                 const double* sb1, double* r, double (*J)[15 * 9]) {
    const SensorState s0 = sensorState(imu, pose0, sb0, 0), s1 = sensorState(imu, pose1, sb1, 1);
    const M3 R = s0.q.conjugate().matrix();
    const double dt = imu.duration;
    const V3 gravity(0, 0, imu.gravity);
    const V3 bg = V3(sb0 + 3) - imu.refBg, ba = V3(sb0 + 6) - imu.refBa;
    const V3 dp = s0.p - s1.p + dt * s0.v - (0.5 * dt * dt) * gravity;
    const V3 dv = s0.v - s1.v - dt * gravity;
    const V3 phi = -1.0 * (imu.A * bg);
    const Q correction = rotationExp(phi), dq = correction * imu.dq;
    const Q tail = imu.dq * s1.q.conjugate() * s0.q;
    double e[15];
    const V3 ep = R * dp + imu.dp + imu.P * bg - imu.Pa * ba;
    const V3 eq = 2.0 * (dq * (s1.q.conjugate() * s0.q)).vec();
    const V3 ev = R * dv + imu.dv + imu.V * bg - imu.Va * ba;
    for (int i = 0; i < 3; ++i) { e[i] = ep[i]; e[3 + i] = eq[i]; e[6 + i] = ev[i]; }
    for (int i = 0; i < 6; ++i) e[9 + i] = sb0[3 + i] - sb1[3 + i];
    for (int i = 0; i < 15; ++i) {
        double s = 0;
        for (int k = 0; k < 15; ++k) s += imu.W[i * 15 + k] * e[k];
        r[i] = s;
        if (!std::isfinite(s)) return false;
    }
    if (!J) return true;
    const M3 gyroAngle = -1.0 * (rotationJacobian(-1.0 * phi) * imu.A);
    M3 Jrotation, Jbias;
    for (int j = 0; j < 3; ++j) {
        V3 axis; axis[j] = 1;
        const V3 db(gyroAngle(0, j), gyroAngle(1, j), gyroAngle(2, j));
        const V3 c1 = (dq * s1.q.conjugate() * Q(0, axis.x, axis.y, axis.z) * s0.q).vec();
        const V3 c2 = (correction * Q(0, db.x, db.y, db.z) * tail).vec();
        for (int i = 0; i < 3; ++i) { Jrotation(i, j) = c1[i]; Jbias(i, j) = c2[i]; }
    }
    for (int k = 0; k < 4; ++k) {
        const bool first = k < 2;
        const double sign = first ? 1 : -1;
        const SensorState& s = first ? s0 : s1;
        double Jr[15 * 9] = {0};
        auto put = [&](int r0, int c0, const M3& m) {
            for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) Jr[(r0 + i) * 9 + c0 + j] = m(i, j);
        };
        auto add = [&](int r0, int c0, const M3& m) {
            for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) Jr[(r0 + i) * 9 + c0 + j] += m(i, j);
        };
        if (k % 2 == 0) {
            put(0, 0, sign * R);
            put(0, 3, (-sign) * (R * skew(s.p)));
            put(3, 3, sign * Jrotation);
            put(6, 3, (-sign) * (R * skew(s.lever)));
            if (first) {
                add(0, 3, R * (skew(dp) - dt * skew(s.lever)));
                add(6, 3, R * skew(dv));
            }
        } else {
            const M3 leverBias = skew(s.arm) * s.q.matrix();
            put(6, 0, sign * R);
            put(6, 3, sign * (R * leverBias));
            put(9, 3, diag3(sign));
            put(12, 6, diag3(sign));
            if (first) {
                put(0, 0, dt * R);
                put(0, 3, dt * (R * leverBias) + imu.P);
                put(0, 6, -1.0 * imu.Pa);
                put(3, 3, Jbias);
                add(6, 3, imu.V);
                put(6, 6, -1.0 * imu.Va);
            }
        }
        for (int i = 0; i < 15; ++i)
            for (int j = 0; j < 9; ++j) {
                double s = 0;
                for (int m = 0; m < 15; ++m) s += imu.W[i * 15 + m] * Jr[m * 9 + j];
                J[k][i * 9 + j] = s;
                if (!std::isfinite(s)) return false;
            }
    }
    return true;
}

// ------------------------------------------------------------------------------------
// Linearization into normal equations (VisualCudaBatch.cu:46-96,207-289)
// ------------------------------------------------------------------------------------
struct State {
    std::vector<double> pose, sb, rho;
    double rig[7], td;
};
State stateOf(const Window& w) {
    State s; s.pose = w.pose; s.sb = w.sb; s.rho = w.rho;
    std::copy_n(w.rig, 7, s.rig); s.td = w.td;
    return s;
}

int slotOfFrameId(const Window& w, int64_t id) {
    for (int k = 0; k < w.K; ++k) if (w.frameId[k] == id) return k;
    return -1;
}

struct Linear {
    int nr = 0, ne = 0;
    std::vector<double> H;      // nr*nr (full symmetric, undamped)
    std::vector<double> cross;  // ne*nr
    std::vector<double> diag;   // ne
    std::vector<double> g;      // nr + ne
    double cost = 0;
    bool valid = true;
};

// Prior residual and chart derivative (MarginalizationPrior.cpp:20-66, VisualCudaBatch.cu:216-268).
// Fills r (m) and, if Jcols != nullptr, dense tangent Jacobian J (m x nr).
bool priorEvaluate(const Window& w, const Layout& lay, const State& x, std::vector<double>& r,
                   std::vector<double>* J) {
    const int m = w.priorRows;
    if (!m) { r.clear(); return true; }
    int cols = 0;
    std::vector<int> chartOff(w.priorBlocks), tsize(w.priorBlocks), refOff(w.priorBlocks);
    int ro = 0;
    for (int b = 0; b < w.priorBlocks; ++b) {
        const int k = w.priorKind[b];
        tsize[b] = k == VIB_KEY_POSE || k == VIB_KEY_RIG ? 6 : (k == VIB_KEY_SPEED_BIAS ? 9 : 1);
        chartOff[b] = cols; cols += tsize[b];
        refOff[b] = ro; ro += k == VIB_KEY_POSE || k == VIB_KEY_RIG ? 7 : (k == VIB_KEY_SPEED_BIAS ? 9 : 1);
    }
    std::vector<double> chart(cols);
    std::vector<std::vector<double>> D(w.priorBlocks);
    std::vector<int> colOf(w.priorBlocks, -1), widthOf(w.priorBlocks, 0);
    for (int b = 0; b < w.priorBlocks; ++b) {
        const int k = w.priorKind[b];
        const double* ref = &w.priorRef[refOff[b]];
        const double* xv = nullptr;
        if (k == VIB_KEY_POSE || k == VIB_KEY_SPEED_BIAS) {
            const int slot = slotOfFrameId(w, w.priorId[b]);
            if (slot < 0) return false;
            xv = k == VIB_KEY_POSE ? &x.pose[7 * slot] : &x.sb[9 * slot];
            colOf[b] = k == VIB_KEY_POSE ? lay.poseCol[slot] : lay.sbCol[slot];
            widthOf[b] = k == VIB_KEY_POSE ? (lay.poseCol[slot] >= 0 ? 6 : 0) : (lay.sbCol[slot] >= 0 ? w.sbWidth[slot] : 0);
        } else if (k == VIB_KEY_RIG) {
            xv = x.rig; colOf[b] = lay.rigCol; widthOf[b] = lay.rigCol >= 0 ? 6 : 0;
        } else {
            xv = &x.td; colOf[b] = lay.tdCol; widthOf[b] = lay.tdCol >= 0 ? 1 : 0;
        }
        const bool se3 = k == VIB_KEY_POSE || k == VIB_KEY_RIG;
        if (se3) poseMinus(xv, ref, &chart[chartOff[b]]);
        else for (int c = 0; c < tsize[b]; ++c) chart[chartOff[b] + c] = xv[c] - ref[c];
        if (!J || !widthOf[b]) continue;
        D[b].assign(tsize[b] * widthOf[b], 0);
        if (!se3) {
            for (int rr = 0; rr < tsize[b]; ++rr)
                for (int c = 0; c < widthOf[b]; ++c) D[b][rr * widthOf[b] + c] = rr == c;
        } else {
            const double h = 1e-7;
            double step[6] = {0}, plus[7], minus[7], dp[6], dm[6];
            for (int c = 0; c < 6; ++c) {
                step[c] = h; posePlus(xv, step, plus, false);
                step[c] = -h; posePlus(xv, step, minus, false);
                step[c] = 0;
                poseMinus(plus, ref, dp); poseMinus(minus, ref, dm);
                for (int rr = 0; rr < 6; ++rr) D[b][rr * 6 + c] = (dp[rr] - dm[rr]) / (2 * h);
            }
        }
    }
    r.assign(m, 0);
    if (J) J->assign((size_t)m * lay.nr, 0);
    for (int row = 0; row < m; ++row) {
        const double* A = &w.priorA[(size_t)row * cols];
        double v = w.priorR0[row];
        for (int c = 0; c < cols; ++c) v += A[c] * chart[c];
        r[row] = v;
        if (!std::isfinite(v)) return false;
        if (!J) continue;
        for (int b = 0; b < w.priorBlocks; ++b) {
            if (!widthOf[b]) continue;
            for (int c = 0; c < widthOf[b]; ++c) {
                double s = 0;
                for (int j = 0; j < tsize[b]; ++j) s += A[chartOff[b] + j] * D[b][j * widthOf[b] + c];
                (*J)[(size_t)row * lay.nr + colOf[b] + c] = s;
            }
        }
    }
    return true;
}

// IMU and prior rows added into the dense retained system H (nr x nr, full), g (first nr
// entries) and cost (GpuRetainedFactorMath.hpp:53-140, VisualCudaBatch.cu:216-289).
bool retainedInto(const Window& w, const Layout& lay, const State& x, std::vector<double>& H,
                  std::vector<double>& g, double& cost) {
    bool valid = true;
    // IMU
    for (int e = 0; e < w.E; ++e) {
        const Frozen imu = unpackImu(&w.imu[(size_t)VIB_IMU_STRIDE * e]);
        const int s0 = w.imuFrom[e], s1 = w.imuTo[e];
        double r[15], J[4][15 * 9];
        if (!imuEvaluate(imu, &x.pose[7 * s0], &x.sb[9 * s0], &x.pose[7 * s1], &x.sb[9 * s1], r, J)) {
            valid = false; continue;
        }
        const int col[4] = {lay.poseCol[s0], lay.sbCol[s0], lay.poseCol[s1], lay.sbCol[s1]};
        const int width[4] = {col[0] >= 0 ? 6 : 0, col[1] >= 0 ? w.sbWidth[s0] : 0,
                              col[2] >= 0 ? 6 : 0, col[3] >= 0 ? w.sbWidth[s1] : 0};
        for (int i = 0; i < 15; ++i) cost += 0.5 * r[i] * r[i];
        for (int bi = 0; bi < 4; ++bi)
            for (int ci = 0; ci < width[bi]; ++ci) {
                const int gi = col[bi] + ci;
                double gsum = 0;
                for (int i = 0; i < 15; ++i) gsum += J[bi][i * 9 + ci] * r[i];
                g[gi] += gsum;
                for (int bj = 0; bj < 4; ++bj)
                    for (int cj = 0; cj < width[bj]; ++cj) {
                        double s = 0;
                        for (int i = 0; i < 15; ++i) s += J[bi][i * 9 + ci] * J[bj][i * 9 + cj];
                        H[(size_t)gi * lay.nr + col[bj] + cj] += s;
                    }
            }
    }
    // prior
    if (w.priorRows) {
        std::vector<double> r, J;
        if (!priorEvaluate(w, lay, x, r, &J)) valid = false;
        else {
            for (int i = 0; i < w.priorRows; ++i) cost += 0.5 * r[i] * r[i];
            for (int i = 0; i < w.priorRows; ++i) {
                const double* row = &J[(size_t)i * lay.nr];
                for (int a = 0; a < lay.nr; ++a) {
                    if (row[a] == 0) continue;
                    g[a] += row[a] * r[i];
                    for (int b = 0; b < lay.nr; ++b) H[(size_t)a * lay.nr + b] += row[a] * row[b];
                }
            }
        }
    }
    return valid;
}

// Full linearization at x (visual + IMU + prior). Mirrors launchVisualLinearization +
// evaluateRetained: robustified tangent rows accumulated into pose block, cross, diag, g.
Linear linearize(const Window& w, const Layout& lay, const State& x) {
    Linear L;
    L.nr = lay.nr; L.ne = lay.ne;
    L.H.assign((size_t)lay.nr * lay.nr, 0);
    L.cross.assign((size_t)lay.ne * lay.nr, 0);
    L.diag.assign(lay.ne, 0);
    L.g.assign(lay.nr + lay.ne, 0);
    std::vector<double> plusA(42), plusB(42), plusR(42);
    for (int f = 0; f < w.F; ++f) {
        const int l = w.obsLm[f], a = w.lmFrame[l], b = w.obsFrame[f];
        const bool same = a == b;
        const int ca = lay.poseCol[a], cb = same ? -1 : lay.poseCol[b];
        double r[2];
        VisualJ J;
        const bool ok = visualEvaluate(w, f, &x.pose[7 * a], &x.pose[7 * b], x.rig, x.td, x.rho[l], r, &J,
                                       ca >= 0, cb >= 0, lay.rigCol >= 0, lay.tdCol >= 0, lay.lmCol[l] >= 0);
        if (!ok) { L.valid = false; continue; }
        const double sq = r[0] * r[0] + r[1] * r[1];
        if (!std::isfinite(sq)) { L.valid = false; continue; }
        L.cost += robustCost(w.huber, sq);
        const double wgt = huberWeight(w.huber, sq);
        // tangent rows: list of (column, value row0, value row1)
        std::vector<int> cols;
        std::vector<double> v0, v1;
        auto addBlock = [&](int col, const double* amb, int size, const double* lift, int width) {
            for (int c = 0; c < width; ++c) {
                double s0 = 0, s1 = 0;
                for (int k = 0; k < size; ++k) {
                    s0 += amb[k] * lift[k * width + c];
                    s1 += amb[size + k] * lift[k * width + c];
                }
                cols.push_back(col + c); v0.push_back(wgt * s0); v1.push_back(wgt * s1);
            }
        };
        if (ca >= 0) { plusJacobian(&x.pose[7 * a], plusA.data()); addBlock(ca, J.anchor, 7, plusA.data(), 6); }
        if (cb >= 0) { plusJacobian(&x.pose[7 * b], plusB.data()); addBlock(cb, J.observer, 7, plusB.data(), 6); }
        if (lay.rigCol >= 0) { plusJacobian(x.rig, plusR.data()); addBlock(lay.rigCol, J.rig, 7, plusR.data(), 6); }
        const double one = 1;
        if (lay.tdCol >= 0) addBlock(lay.tdCol, J.td, 1, &one, 1);
        if (lay.lmCol[l] >= 0) addBlock(lay.lmCol[l], J.rho, 1, &one, 1);
        const double rr0 = wgt * r[0], rr1 = wgt * r[1];
        for (size_t i = 0; i < cols.size(); ++i) {
            const int ci = cols[i];
            L.g[ci] += v0[i] * rr0 + v1[i] * rr1;
            for (size_t j = 0; j < cols.size(); ++j) {
                const int cj = cols[j];
                const double val = v0[i] * v0[j] + v1[i] * v1[j];
                if (ci < lay.nr && cj < lay.nr) L.H[(size_t)ci * lay.nr + cj] += val;
                else if (ci >= lay.nr && cj < lay.nr) L.cross[(size_t)(ci - lay.nr) * lay.nr + cj] += val;
                else if (ci >= lay.nr && cj == ci) L.diag[ci - lay.nr] += val;
            }
        }
    }
    if (!retainedInto(w, lay, x, L.H, L.g, L.cost)) L.valid = false;
    if (!std::isfinite(L.cost)) L.valid = false;
    return L;
}

// Cost only (trial): VisualCudaBatch.cu:715-721
bool costAt(const Window& w, const Layout& lay, const State& x, double& cost) {
    cost = 0;
    for (int f = 0; f < w.F; ++f) {
        const int l = w.obsLm[f], a = w.lmFrame[l], b = w.obsFrame[f];
        double r[2];
        if (!visualEvaluate(w, f, &x.pose[7 * a], &x.pose[7 * b], x.rig, x.td, x.rho[l], r, nullptr,
                            false, false, false, false, false)) return false;
        const double sq = r[0] * r[0] + r[1] * r[1];
        if (!std::isfinite(sq)) return false;
        cost += robustCost(w.huber, sq);
    }
    for (int e = 0; e < w.E; ++e) {
        const Frozen imu = unpackImu(&w.imu[(size_t)VIB_IMU_STRIDE * e]);
        const int s0 = w.imuFrom[e], s1 = w.imuTo[e];
        double r[15];
        if (!imuEvaluate(imu, &x.pose[7 * s0], &x.sb[9 * s0], &x.pose[7 * s1], &x.sb[9 * s1], r, nullptr)) return false;
        for (int i = 0; i < 15; ++i) cost += 0.5 * r[i] * r[i];
    }
    if (w.priorRows) {
        std::vector<double> r;
        if (!priorEvaluate(w, lay, x, r, nullptr)) return false;
        for (double v : r) cost += 0.5 * v * v;
    }
    return std::isfinite(cost);
}

// Dense Cholesky (LL^T) of an n x n symmetric matrix (upper triangle read). Returns false if
// a pivot is not positive (the potrf info != 0 failure class).
bool cholesky(std::vector<double>& A, int n) {
    for (int j = 0; j < n; ++j) {
        double d = A[(size_t)j * n + j];
        for (int k = 0; k < j; ++k) d -= A[(size_t)j * n + k] * A[(size_t)j * n + k];
        if (!(d > 0) || !std::isfinite(d)) return false;
        d = std::sqrt(d);
        A[(size_t)j * n + j] = d;
        for (int i = j + 1; i < n; ++i) {
            double s = A[(size_t)std::min(i, j) * n + std::max(i, j)];  // symmetric source (upper)
            for (int k = 0; k < j; ++k) s -= A[(size_t)i * n + k] * A[(size_t)j * n + k];
            A[(size_t)i * n + j] = s / d;
        }
    }
    return true;
}
void choleskySolve(const std::vector<double>& L, int n, std::vector<double>& b) {
    for (int i = 0; i < n; ++i) {
        double s = b[i];
        for (int k = 0; k < i; ++k) s -= L[(size_t)i * n + k] * b[k];
        b[i] = s / L[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = b[i];
        for (int k = i + 1; k < n; ++k) s -= L[(size_t)k * n + i] * b[k];
        b[i] = s / L[(size_t)i * n + i];
    }
}

struct Bounds { std::vector<double> lower, upper; };  // absolute, per tangent column
Bounds boundsOf(const Window& w, const Layout& lay) {
    Bounds b;
    const int n = lay.nr + lay.ne;
    b.lower.assign(n, -kInf); b.upper.assign(n, kInf);
    if (lay.tdCol >= 0) { b.lower[lay.tdCol] = w.tdLower; b.upper[lay.tdCol] = w.tdUpper; }
    for (int l = 0; l < w.L; ++l) if (lay.lmCol[l] >= 0) b.lower[lay.lmCol[l]] = w.minRho;
    return b;
}

double damped(double d, double lambda) { return d + lambda * std::fmax(d, 1e-6); }  // :145-147

struct Step {
    bool ok = false;            // false: Cholesky failure
    std::vector<double> delta;  // nr + ne
    std::vector<double> S, rhs; // damped reduced system (full symmetric) and rhs
    double predicted = 0, trialCost = 0, maxIncrement = 0;
    bool trialValid = false;
    State candidate;
};

// solveOnline restated (VisualCudaBatch.cu:530-604, kernels :138-205,291-340).
Step solveStep(const Window& w, const Layout& lay, const Linear& L, const State& x, const Bounds& abs,
               double lambda) {
    Step st;
    const int nr = lay.nr, ne = lay.ne;
    // relative bounds (prepareResidentVariables :291-300): lower = absLower - x
    // Only td and rho carry finite bounds; the other blocks stay at +-inf after the shift.
    std::vector<double> lower(nr + ne), upper(nr + ne);
    for (int i = 0; i < nr; ++i) {
        const double v = i == lay.tdCol ? x.td : 0.0;
        lower[i] = abs.lower[i] - v;
        upper[i] = abs.upper[i] - v;
    }
    std::vector<int> lmOfCol(ne, -1);
    for (int l = 0; l < w.L; ++l) if (lay.lmCol[l] >= 0) lmOfCol[lay.lmCol[l] - nr] = l;
    for (int k = 0; k < ne; ++k) {
        lower[nr + k] = abs.lower[nr + k] - x.rho[lmOfCol[k]];
        upper[nr + k] = abs.upper[nr + k] - x.rho[lmOfCol[k]];
    }
    // initializeReduced + subtractPointSchur
    std::vector<double> S(L.H), gr(L.g.begin(), L.g.begin() + nr);
    for (int k = 0; k < ne; ++k) {
        const double inv = 1 / damped(L.diag[k], lambda);
        const double* c = &L.cross[(size_t)k * nr];
        for (int a = 0; a < nr; ++a) {
            if (c[a] == 0) continue;
            for (int b = 0; b < nr; ++b) if (c[b] != 0) S[(size_t)a * nr + b] -= inv * c[a] * c[b];
            gr[a] -= inv * c[a] * L.g[nr + k];
        }
    }
    // prepareStateSolve: damping on the undamped full diagonal, rhs = -g
    for (int i = 0; i < nr; ++i) S[(size_t)i * nr + i] += lambda * std::fmax(L.H[(size_t)i * nr + i], 1e-6);
    st.S = S;
    std::vector<double> rhs(nr);
    for (int i = 0; i < nr; ++i) rhs[i] = -gr[i];
    st.rhs = rhs;
    std::vector<double> chol(S);
    if (!cholesky(chol, nr)) return st;
    st.ok = true;
    std::vector<double> ds(rhs);
    choleskySolve(chol, nr, ds);
    // clampStateStep
    for (int i = 0; i < nr; ++i) if (std::isfinite(ds[i])) ds[i] = std::fmax(lower[i], std::fmin(upper[i], ds[i]));
    // statePrediction (undamped H)
    double pred = 0;
    for (int i = 0; i < nr; ++i) {
        double hd = 0;
        for (int j = 0; j < nr; ++j) hd += L.H[(size_t)std::min(i, j) * nr + std::max(i, j)] * ds[j];
        pred += -ds[i] * (L.g[i] + 0.5 * hd);
    }
    st.delta.assign(nr + ne, 0);
    for (int i = 0; i < nr; ++i) st.delta[i] = ds[i];
    // backSubstitute
    for (int k = 0; k < ne; ++k) {
        const double* c = &L.cross[(size_t)k * nr];
        double cd = 0;
        for (int j = 0; j < nr; ++j) if (c[j] != 0) cd += c[j] * ds[j];
        const int i = nr + k;
        const double d = -(L.g[i] + cd) / damped(L.diag[k], lambda);
        st.delta[i] = std::isfinite(d) ? std::fmax(lower[i], std::fmin(upper[i], d)) : d;
        pred += -st.delta[i] * (L.g[i] + cd + 0.5 * L.diag[k] * st.delta[i]);
    }
    st.predicted = pred;
    // residentTrial
    State y = x;
    bool invalid = false;
    double maxInc = 0;
    auto track = [&](const double* d, int width) {
        for (int k = 0; k < width; ++k) {
            if (!std::isfinite(d[k])) invalid = true;
            maxInc = std::fmax(maxInc, std::fabs(d[k]));
        }
    };
    for (int k = 0; k < w.K; ++k) {
        if (lay.poseCol[k] >= 0) {
            const double* d = &st.delta[lay.poseCol[k]];
            track(d, 6);
            posePlus(&x.pose[7 * k], d, &y.pose[7 * k], false);
        }
        if (lay.sbCol[k] >= 0) {
            const double* d = &st.delta[lay.sbCol[k]];
            track(d, w.sbWidth[k]);
            for (int j = 0; j < w.sbWidth[k]; ++j) y.sb[9 * k + j] = x.sb[9 * k + j] + d[j];
        }
    }
    if (lay.rigCol >= 0) { track(&st.delta[lay.rigCol], 6); posePlus(x.rig, &st.delta[lay.rigCol], y.rig, false); }
    if (lay.tdCol >= 0) {
        track(&st.delta[lay.tdCol], 1);
        y.td = std::fmax(w.tdLower, std::fmin(w.tdUpper, x.td + st.delta[lay.tdCol]));
    }
    for (int l = 0; l < w.L; ++l)
        if (lay.lmCol[l] >= 0) {
            track(&st.delta[lay.lmCol[l]], 1);
            y.rho[l] = std::fmax(w.minRho, x.rho[l] + st.delta[lay.lmCol[l]]);
        }
    st.maxIncrement = maxInc;
    st.candidate = y;
    double tc = 0;
    const bool costOk = !invalid && costAt(w, lay, y, tc);
    st.trialCost = tc;
    st.trialValid = costOk && std::isfinite(tc);
    return st;
}

// ------------------------------------------------------------------------------------
// Block-sparse, multi-threaded LM step (the timed CPU baseline). The same arithmetic as
// linearize + solveStep, with the landmark Schur complement formed landmark by landmark on the
// landmark's support columns only (anchor and observer poses, rig, td), as Ceres'
// SchurEliminator does for the reference CPU path (backend_visual_batch: cpu,
// BackendOptimizer.cpp:900-989, DENSE_SCHUR), and the landmark work split over std::thread
// ranges. Sums are combined in thread order, so the result depends only on the thread count.
// Checked against the dense solveStep in tests/test_oracle.py.
// ------------------------------------------------------------------------------------
struct Prepared {  // topology shared by every step of one window
    Window w;
    Layout lay;
    Bounds abs;
    std::vector<int> lmOff, lmFac;  // factors of each landmark (CSR)
};

Prepared prepare(const vib_window& win) {
    Prepared p;
    p.w = load(win);
    p.lay = layoutOf(p.w);
    p.abs = boundsOf(p.w, p.lay);
    p.lmOff.assign(p.w.L + 1, 0);
    for (int f = 0; f < p.w.F; ++f) p.lmOff[p.w.obsLm[f] + 1]++;
    for (int l = 0; l < p.w.L; ++l) p.lmOff[l + 1] += p.lmOff[l];
    std::vector<int> cur(p.lmOff.begin(), p.lmOff.end() - 1);
    p.lmFac.assign(p.w.F, 0);
    for (int f = 0; f < p.w.F; ++f) p.lmFac[cur[p.w.obsLm[f]]++] = f;
    return p;
}

struct LmTerms {  // one landmark's coupling on its support columns
    int begin = 0, count = 0;  // range in the thread's column / value lists
    double d = 0, g = 0;
};

struct ThreadPart {
    std::vector<double> H, Sc, g, gr;  // visual block, Schur correction, gradients (nr)
    std::vector<int> col;
    std::vector<double> val;
    std::vector<LmTerms> lm;  // landmarks of this thread's range, in order
    double cost = 0, pred = 0, maxInc = 0;
    bool valid = true, invalid = false;
};

template <class Fn>
void parallelRanges(int n, int threads, Fn fn) {
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) {
        const int a = (int)((int64_t)n * t / threads), b = (int)((int64_t)n * (t + 1) / threads);
        if (threads == 1) fn(t, a, b);
        else pool.emplace_back(fn, t, a, b);
    }
    for (auto& th : pool) th.join();
}

// Upper-triangle row-major Cholesky S = U^T U in place, right-looking (row updates vectorize).
bool choleskyUpper(std::vector<double>& A, int n) {
    for (int k = 0; k < n; ++k) {
        double* rk = &A[(size_t)k * n];
        const double d = rk[k];
        if (!(d > 0) || !std::isfinite(d)) return false;
        const double piv = std::sqrt(d);
        rk[k] = piv;
        for (int j = k + 1; j < n; ++j) rk[j] /= piv;
        for (int i = k + 1; i < n; ++i) {
            double* ri = &A[(size_t)i * n];
            const double u = rk[i];
            if (u == 0) continue;
            for (int j = i; j < n; ++j) ri[j] -= u * rk[j];
        }
    }
    return true;
}
void choleskyUpperSolve(const std::vector<double>& U, int n, std::vector<double>& b) {
    for (int i = 0; i < n; ++i) {  // U^T y = b
        double s = b[i];
        for (int k = 0; k < i; ++k) s -= U[(size_t)k * n + i] * b[k];
        b[i] = s / U[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {  // U x = y
        double s = b[i];
        for (int k = i + 1; k < n; ++k) s -= U[(size_t)i * n + k] * b[k];
        b[i] = s / U[(size_t)i * n + i];
    }
}

bool retainedCost(const Window& w, const Layout& lay, const State& x, double& cost) {
    for (int e = 0; e < w.E; ++e) {
        const Frozen imu = unpackImu(&w.imu[(size_t)VIB_IMU_STRIDE * e]);
        const int s0 = w.imuFrom[e], s1 = w.imuTo[e];
        double r[15];
        if (!imuEvaluate(imu, &x.pose[7 * s0], &x.sb[9 * s0], &x.pose[7 * s1], &x.sb[9 * s1], r, nullptr)) return false;
        for (int i = 0; i < 15; ++i) cost += 0.5 * r[i] * r[i];
    }
    if (w.priorRows) {
        std::vector<double> r;
        if (!priorEvaluate(w, lay, x, r, nullptr)) return false;
        for (double v : r) cost += 0.5 * v * v;
    }
    return true;
}

int blockStep(const Prepared& P, const State& x, double lambda, int threads, std::vector<double>& delta,
              double* metrics) {
    const Window& w = P.w;
    const Layout& lay = P.lay;
    const int nr = lay.nr;
    threads = std::max(1, std::min(threads, std::max(1, w.L)));
    std::vector<ThreadPart> parts(threads);
    // linearize the visual factors and eliminate each landmark on its support (one pass)
    parallelRanges(w.L, threads, [&](int t, int l0, int l1) {
        ThreadPart& tp = parts[t];
        tp.H.assign((size_t)nr * nr, 0); tp.Sc.assign((size_t)nr * nr, 0);
        tp.g.assign(nr, 0); tp.gr.assign(nr, 0);
        std::vector<int> local(nr, -1);
        double plusA[42], plusB[42], plusR[42];
        int cols[20];
        double v0[20], v1[20];
        for (int l = l0; l < l1; ++l) {
            LmTerms lt;
            lt.begin = (int)tp.col.size();
            const int rhoCol = lay.lmCol[l];
            for (int q = P.lmOff[l]; q < P.lmOff[l + 1]; ++q) {
                const int f = P.lmFac[q], a = w.lmFrame[l], b = w.obsFrame[f];
                const int ca = lay.poseCol[a], cb = a == b ? -1 : lay.poseCol[b];
                double r[2];
                VisualJ J;
                if (!visualEvaluate(w, f, &x.pose[7 * a], &x.pose[7 * b], x.rig, x.td, x.rho[l], r, &J, ca >= 0,
                                    cb >= 0, lay.rigCol >= 0, lay.tdCol >= 0, rhoCol >= 0)) { tp.valid = false; continue; }
                const double sq = r[0] * r[0] + r[1] * r[1];
                if (!std::isfinite(sq)) { tp.valid = false; continue; }
                tp.cost += robustCost(w.huber, sq);
                const double wgt = huberWeight(w.huber, sq);
                int n = 0;
                auto add = [&](int col, const double* amb, int size, const double* lift, int width) {
                    for (int c = 0; c < width; ++c) {
                        double s0 = 0, s1 = 0;
                        for (int k = 0; k < size; ++k) {
                            s0 += amb[k] * lift[k * width + c];
                            s1 += amb[size + k] * lift[k * width + c];
                        }
                        cols[n] = col + c; v0[n] = wgt * s0; v1[n] = wgt * s1; ++n;
                    }
                };
                if (ca >= 0) { plusJacobian(&x.pose[7 * a], plusA); add(ca, J.anchor, 7, plusA, 6); }
                if (cb >= 0) { plusJacobian(&x.pose[7 * b], plusB); add(cb, J.observer, 7, plusB, 6); }
                if (lay.rigCol >= 0) { plusJacobian(x.rig, plusR); add(lay.rigCol, J.rig, 7, plusR, 6); }
                const double one = 1;
                if (lay.tdCol >= 0) add(lay.tdCol, J.td, 1, &one, 1);
                const double rr0 = wgt * r[0], rr1 = wgt * r[1];
                for (int i = 0; i < n; ++i) {
                    tp.g[cols[i]] += v0[i] * rr0 + v1[i] * rr1;
                    double* Hi = &tp.H[(size_t)cols[i] * nr];
                    for (int j = 0; j < n; ++j) Hi[cols[j]] += v0[i] * v0[j] + v1[i] * v1[j];
                }
                if (rhoCol < 0) continue;
                const double jr0 = wgt * J.rho[0], jr1 = wgt * J.rho[1];
                lt.d += jr0 * jr0 + jr1 * jr1;
                lt.g += jr0 * rr0 + jr1 * rr1;
                for (int i = 0; i < n; ++i) {
                    int& li = local[cols[i]];
                    if (li < 0) { li = (int)tp.col.size(); tp.col.push_back(cols[i]); tp.val.push_back(0); }
                    tp.val[li] += jr0 * v0[i] + jr1 * v1[i];
                }
            }
            lt.count = (int)tp.col.size() - lt.begin;
            for (int i = lt.begin; i < lt.begin + lt.count; ++i) local[tp.col[i]] = -1;
            if (rhoCol >= 0) {  // S -= c c^T / damped(d), g_r -= c g / damped(d) on the support
                const double inv = 1 / damped(lt.d, lambda);
                const int* c = &tp.col[lt.begin];
                const double* v = &tp.val[lt.begin];
                for (int i = 0; i < lt.count; ++i) {
                    double* Si = &tp.Sc[(size_t)c[i] * nr];
                    const double vi = inv * v[i];
                    for (int j = 0; j < lt.count; ++j) Si[c[j]] -= vi * v[j];
                    tp.gr[c[i]] -= vi * lt.g;
                }
            }
            tp.lm.push_back(lt);
        }
    });
    std::vector<double> H((size_t)nr * nr, 0), g(nr, 0), Sc((size_t)nr * nr, 0), gr(nr, 0);
    double cost = 0;
    bool valid = true;
    for (const ThreadPart& tp : parts) {
        for (size_t i = 0; i < H.size(); ++i) { H[i] += tp.H[i]; Sc[i] += tp.Sc[i]; }
        for (int i = 0; i < nr; ++i) { g[i] += tp.g[i]; gr[i] += tp.gr[i]; }
        cost += tp.cost;
        valid = valid && tp.valid;
    }
    if (!retainedInto(w, lay, x, H, g, cost)) valid = false;
    if (!valid || !std::isfinite(cost)) return VIB_ERR_INVALID_RESIDUAL;
    std::vector<double> S(H), rhs(nr);
    for (size_t i = 0; i < S.size(); ++i) S[i] += Sc[i];
    for (int i = 0; i < nr; ++i) {
        S[(size_t)i * nr + i] += lambda * std::fmax(H[(size_t)i * nr + i], 1e-6);
        rhs[i] = -(g[i] + gr[i]);
    }
    if (!choleskyUpper(S, nr)) return VIB_ERR_NOT_POSITIVE_DEFINITE;
    std::vector<double> ds(rhs);
    choleskyUpperSolve(S, nr, ds);
    for (int i = 0; i < nr; ++i) {
        const double v = i == lay.tdCol ? x.td : 0.0;
        if (std::isfinite(ds[i])) ds[i] = std::fmax(P.abs.lower[i] - v, std::fmin(P.abs.upper[i] - v, ds[i]));
    }
    double pred = 0;
    for (int i = 0; i < nr; ++i) {
        double hd = 0;
        for (int j = 0; j < nr; ++j) hd += H[(size_t)i * nr + j] * ds[j];
        pred += -ds[i] * (g[i] + 0.5 * hd);
    }
    delta.assign(nr + lay.ne, 0);
    std::copy(ds.begin(), ds.end(), delta.begin());
    // back-substitution of the inverse depths, then the candidate
    parallelRanges(w.L, threads, [&](int t, int l0, int l1) {
        ThreadPart& tp = parts[t];
        tp.pred = 0;
        for (int l = l0, q = 0; l < l1; ++l, ++q) {
            const int i = lay.lmCol[l];
            if (i < 0) continue;
            const LmTerms& lt = tp.lm[q];
            double cd = 0;
            for (int k = lt.begin; k < lt.begin + lt.count; ++k) cd += tp.val[k] * ds[tp.col[k]];
            const double dl = -(lt.g + cd) / damped(lt.d, lambda);
            const double lo = P.abs.lower[i] - x.rho[l], hi = P.abs.upper[i] - x.rho[l];
            delta[i] = std::isfinite(dl) ? std::fmax(lo, std::fmin(hi, dl)) : dl;
            tp.pred += -delta[i] * (lt.g + cd + 0.5 * lt.d * delta[i]);
        }
    });
    for (const ThreadPart& tp : parts) pred += tp.pred;
    State y = x;
    bool invalid = false;
    double maxInc = 0;
    auto track = [&](const double* d, int width) {
        for (int k = 0; k < width; ++k) {
            if (!std::isfinite(d[k])) invalid = true;
            maxInc = std::fmax(maxInc, std::fabs(d[k]));
        }
    };
    for (int k = 0; k < w.K; ++k) {
        if (lay.poseCol[k] >= 0) { track(&delta[lay.poseCol[k]], 6); posePlus(&x.pose[7 * k], &delta[lay.poseCol[k]], &y.pose[7 * k], false); }
        if (lay.sbCol[k] >= 0) {
            track(&delta[lay.sbCol[k]], w.sbWidth[k]);
            for (int j = 0; j < w.sbWidth[k]; ++j) y.sb[9 * k + j] = x.sb[9 * k + j] + delta[lay.sbCol[k] + j];
        }
    }
    if (lay.rigCol >= 0) { track(&delta[lay.rigCol], 6); posePlus(x.rig, &delta[lay.rigCol], y.rig, false); }
    if (lay.tdCol >= 0) { track(&delta[lay.tdCol], 1); y.td = std::fmax(w.tdLower, std::fmin(w.tdUpper, x.td + delta[lay.tdCol])); }
    for (int l = 0; l < w.L; ++l)
        if (lay.lmCol[l] >= 0) { track(&delta[lay.lmCol[l]], 1); y.rho[l] = std::fmax(w.minRho, x.rho[l] + delta[lay.lmCol[l]]); }
    // trial cost (residual-only pass)
    double tc = 0;
    bool costOk = !invalid;
    if (costOk) {
        parallelRanges(w.L, threads, [&](int t, int l0, int l1) {
            ThreadPart& tp = parts[t];
            tp.cost = 0;
            tp.invalid = false;
            for (int l = l0; l < l1 && !tp.invalid; ++l)
                for (int q = P.lmOff[l]; q < P.lmOff[l + 1]; ++q) {
                    const int f = P.lmFac[q], a = w.lmFrame[l], b = w.obsFrame[f];
                    double r[2];
                    if (!visualEvaluate(w, f, &y.pose[7 * a], &y.pose[7 * b], y.rig, y.td, y.rho[l], r, nullptr, false,
                                        false, false, false, false)) { tp.invalid = true; break; }
                    const double sq = r[0] * r[0] + r[1] * r[1];
                    if (!std::isfinite(sq)) { tp.invalid = true; break; }
                    tp.cost += robustCost(w.huber, sq);
                }
        });
        for (const ThreadPart& tp : parts) { tc += tp.cost; costOk = costOk && !tp.invalid; }
        costOk = costOk && retainedCost(w, lay, y, tc);
    }
    metrics[0] = cost; metrics[1] = pred; metrics[2] = tc; metrics[3] = maxInc;
    metrics[4] = costOk && std::isfinite(tc) ? 1 : 0;
    return VIB_OK;
}

void writeState(const State& s, const Window& w, vib_window* out) {
    std::copy(s.pose.begin(), s.pose.end(), out->pose);
    std::copy(s.sb.begin(), s.sb.end(), out->speed_bias);
    std::copy(s.rho.begin(), s.rho.end(), out->lm_rho);
    std::copy_n(s.rig, 7, out->rig);
    *out->td = s.td;
    (void)w;
}

// Native LM loop (BackendOptimizer.cpp:167-307, resident branch).
int lmSolve(Window& w, const vib_solve_options& o, vib_solve_report& rep, State& finalState) {
    const auto t0 = std::chrono::steady_clock::now();
    const Layout lay = layoutOf(w);
    const Bounds abs = boundsOf(w, lay);
    State x = stateOf(w);
    rep = {};
    rep.effective_parameters = lay.nr + lay.ne;
    rep.retained_dimension = lay.nr;
    double lambda = o.initial_lambda > 0 ? o.initial_lambda : 1e-4, increase = 2, cost = 0;
    int accepted = 0;
    bool newPoint = true, converged = false;
    Linear L;
    int it = 0;
    for (int i = 0; i < o.max_iterations; ++i) {
        it = i + 1;
        if (newPoint) {
            L = linearize(w, lay, x);
            if (!L.valid) { rep.status = VIB_ERR_INVALID_RESIDUAL; return rep.status; }
            const double linCost = L.cost;
            if (!std::isfinite(linCost) || (i > 0 && std::fabs(linCost - cost) > 1e-8 * std::max(1.0, cost))) {
                rep.status = VIB_ERR_COST_MISMATCH; return rep.status;
            }
            if (i == 0) rep.initial_cost = cost = linCost;
        }
        Step st = solveStep(w, lay, L, x, abs, lambda);
        if (!st.ok) { rep.status = VIB_ERR_NOT_POSITIVE_DEFINITE; return rep.status; }
        const double predicted = st.predicted;
        double nextCost = 0, gain = -1;
        if (st.trialValid && st.maxIncrement < 1e-10) { converged = true; break; }
        if (st.trialValid && predicted > 0 && std::isfinite(predicted)) {
            nextCost = st.trialCost;
            gain = (cost - nextCost) / predicted;
        }
        if (gain > 1e-3) {
            x = st.candidate;
            ++accepted;
            converged = cost - nextCost <= 1e-6 * std::max(1.0, cost);
            cost = nextCost;
            lambda = std::max(1e-12, lambda * std::max(1.0 / 3, 1 - std::pow(2 * gain - 1, 3)));
            increase = 2; newPoint = true;
            if (converged) break;
        } else {
            lambda *= increase; increase *= 2; newPoint = false;
        }
    }
    rep.iterations = it;
    rep.accepted_steps = accepted;
    rep.converged = converged;
    rep.final_cost = cost;
    rep.final_lambda = lambda;
    rep.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (!accepted && !converged) { rep.status = VIB_ERR_NO_ACCEPTABLE_STEP; return rep.status; }
    finalState = x;
    rep.status = VIB_OK;
    return VIB_OK;
}

// ------------------------------------------------------------------------------------
// Symmetric eigensolver (cyclic Jacobi; Golub & Van Loan, Matrix Computations, 8.5).
// Eigenvalues ascending, eigenvectors as columns of V (row-major n x n).
// ------------------------------------------------------------------------------------
bool jacobiEigen(std::vector<double> A, int n, std::vector<double>& evals, std::vector<double>& V) {
    V.assign((size_t)n * n, 0);
    for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0, total = 0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                const double v = A[(size_t)i * n + j] * A[(size_t)i * n + j];
                total += v;
                if (i != j) off += v;
            }
        if (off <= 1e-30 * total || off == 0) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = A[(size_t)p * n + q];
                if (apq == 0) continue;
                const double app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q];
                const double tau = (aqq - app) / (2 * apq);
                const double t = (tau >= 0 ? 1 : -1) / (std::fabs(tau) + std::sqrt(1 + tau * tau));
                const double c = 1 / std::sqrt(1 + t * t), s = t * c;
                for (int k = 0; k < n; ++k) {
                    const double akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
                    A[(size_t)k * n + p] = c * akp - s * akq;
                    A[(size_t)k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    const double apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
                    A[(size_t)p * n + k] = c * apk - s * aqk;
                    A[(size_t)q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
                    V[(size_t)k * n + p] = c * vkp - s * vkq;
                    V[(size_t)k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    evals.resize(n);
    for (int i = 0; i < n; ++i) evals[i] = A[(size_t)i * n + i];
    // sort ascending
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    std::sort(idx.begin(), idx.end(), [&](int a, int b) { return evals[a] < evals[b]; });
    std::vector<double> e2(n), V2((size_t)n * n);
    for (int j = 0; j < n; ++j) {
        e2[j] = evals[idx[j]];
        for (int i = 0; i < n; ++i) V2[(size_t)i * n + j] = V[(size_t)i * n + idx[j]];
    }
    evals = e2; V = V2;
    for (double v : evals) if (!std::isfinite(v)) return false;
    return true;
}

// Marginalizer::psdSolve (Marginalizer.cpp:240-283). H n x n, rhs n x cols (row-major).
bool psdSolve(const std::vector<double>& H, int n, const std::vector<double>& rhs, int cols,
              std::vector<double>& x, double absCut, double relCut) {
    x.assign((size_t)n * cols, 0);
    if (!n) return true;
    bool diagonal = true;
    for (int i = 0; i < n && diagonal; ++i)
        for (int j = 0; j < n; ++j) if (i != j && H[(size_t)i * n + j] != 0.0) { diagonal = false; break; }
    if (diagonal) {
        double mx = -kInf, mn = kInf;
        for (int i = 0; i < n; ++i) { mx = std::fmax(mx, H[(size_t)i * n + i]); mn = std::fmin(mn, H[(size_t)i * n + i]); }
        const double cutoff = std::max(absCut, relCut * std::max(0.0, mx));
        if (mn < -10.0 * cutoff) return false;
        for (int i = 0; i < n; ++i) {
            const double h = H[(size_t)i * n + i];
            for (int c = 0; c < cols; ++c) x[(size_t)i * cols + c] = h > cutoff ? rhs[(size_t)i * cols + c] / h : 0;
        }
        return true;
    }
    std::vector<double> sym((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) sym[(size_t)i * n + j] = 0.5 * (H[(size_t)i * n + j] + H[(size_t)j * n + i]);
    std::vector<double> ev, U;
    if (!jacobiEigen(sym, n, ev, U)) return false;
    const double cutoff = std::max(absCut, relCut * std::max(0.0, ev[n - 1]));
    if (ev[0] < -10.0 * cutoff) return false;
    std::vector<double> proj((size_t)n * cols, 0);
    for (int i = 0; i < n; ++i)
        for (int c = 0; c < cols; ++c) {
            double s = 0;
            for (int k = 0; k < n; ++k) s += U[(size_t)k * n + i] * rhs[(size_t)k * cols + c];
            proj[(size_t)i * cols + c] = ev[i] > cutoff ? s / ev[i] : 0;
        }
    for (int i = 0; i < n; ++i)
        for (int c = 0; c < cols; ++c) {
            double s = 0;
            for (int k = 0; k < n; ++k) s += U[(size_t)i * n + k] * proj[(size_t)k * cols + c];
            x[(size_t)i * cols + c] = s;
        }
    return true;
}

// ------------------------------------------------------------------------------------
// OLD marginalization (Marginalizer.cpp:284-474 with SlidingWindowManager.cpp:627-652
// elimination set, preserveOldestVisualInformation and landmarkFirstMarginalization on).
// ------------------------------------------------------------------------------------
struct MargResult {
    int status = VIB_OK;
    int rank = 0, nm = 0, nr = 0, landmarks = 0, consumed = 0;
    double cutoff = 0;
    std::vector<int> keyKind;
    std::vector<int64_t> keyId;
    std::vector<double> ref, A, r0;
    std::vector<double> Hprior, bprior;  // retained Schur system before the eigen cut
    std::vector<double> Hfull, bfull;    // whole system of the consumed factors (n x n, n)
    std::vector<int> elimMask;           // 1 for eliminated columns of the whole system
    int cols = 0;
};

// Schur elimination of slot `victim` (pose, speed/bias) and the landmarks of elimLm into a new
// prior, from the consumed visual factors `visual`, IMU edges `imus` and the previous prior
// (always included once, Marginalizer.cpp:311-316) (Marginalizer::compute, :284-474).
MargResult marginalizeSet(const Window& w, int victim, const std::vector<char>& elimLm,
                          const std::vector<int>& visual, const std::vector<int>& imus) {
    MargResult out;
    out.consumed = (int)visual.size() + (int)imus.size() + (w.priorRows ? 1 : 0);
    // keys touched, in BlockKey order: poses(id), sbs(id), rig, td, landmarks(id)
    std::set<int> poseSlots, sbSlots;
    bool rigUsed = false, tdUsed = false;
    std::vector<int> lmKeys;
    for (int f : visual) {
        const int l = w.obsLm[f];
        poseSlots.insert(w.lmFrame[l]); poseSlots.insert(w.obsFrame[f]);
        rigUsed = tdUsed = true;
    }
    for (int l = 0; l < w.L; ++l) if (elimLm[l]) lmKeys.push_back(l);
    for (int e : imus) { poseSlots.insert(w.imuFrom[e]); poseSlots.insert(w.imuTo[e]); sbSlots.insert(w.imuFrom[e]); sbSlots.insert(w.imuTo[e]); }
    std::vector<int> pKind; std::vector<int64_t> pId;
    for (int b = 0; b < w.priorBlocks; ++b) {
        const int k = w.priorKind[b];
        if (k == VIB_KEY_POSE) poseSlots.insert(slotOfFrameId(w, w.priorId[b]));
        else if (k == VIB_KEY_SPEED_BIAS) sbSlots.insert(slotOfFrameId(w, w.priorId[b]));
        else if (k == VIB_KEY_RIG) rigUsed = true;
        else tdUsed = true;
    }
    // column offsets (frame ids increase with slot, so slot order == id order)
    std::vector<int> poseCol(w.K, -1), sbCol(w.K, -1), lmCol(w.L, -1);
    int n = 0, rigCol = -1, tdCol = -1;
    for (int s : poseSlots) { poseCol[s] = n; n += 6; }
    for (int s : sbSlots) { sbCol[s] = n; n += 9; }
    if (rigUsed) { rigCol = n; n += 6; }
    if (tdUsed) { tdCol = n; n += 1; }
    for (int l : lmKeys) { lmCol[l] = n; n += 1; }
    std::vector<double> H((size_t)n * n, 0), b(n, 0);
    const State x = stateOf(w);
    std::vector<double> plus(42);
    // visual rows (robustified, all blocks treated as variables)
    for (int f : visual) {
        const int l = w.obsLm[f], a = w.lmFrame[l], ob = w.obsFrame[f];
        const bool same = a == ob;
        double r[2]; VisualJ J;
        if (!visualEvaluate(w, f, &x.pose[7 * a], &x.pose[7 * ob], x.rig, x.td, x.rho[l], r, &J, true, true, true, true, true)) {
            out.status = VIB_ERR_INVALID_RESIDUAL; return out;
        }
        const double sq = r[0] * r[0] + r[1] * r[1];
        const double wgt = huberWeight(w.huber, sq);
        std::vector<int> cols; std::vector<double> v0, v1;
        auto addBlock = [&](int col, const double* amb, int size, const double* lift, int width) {
            for (int c = 0; c < width; ++c) {
                double s0 = 0, s1 = 0;
                for (int k = 0; k < size; ++k) { s0 += amb[k] * lift[k * width + c]; s1 += amb[size + k] * lift[k * width + c]; }
                cols.push_back(col + c); v0.push_back(wgt * s0); v1.push_back(wgt * s1);
            }
        };
        plusJacobian(&x.pose[7 * a], plus.data()); addBlock(poseCol[a], J.anchor, 7, plus.data(), 6);
        if (!same) { plusJacobian(&x.pose[7 * ob], plus.data()); addBlock(poseCol[ob], J.observer, 7, plus.data(), 6); }
        plusJacobian(x.rig, plus.data()); addBlock(rigCol, J.rig, 7, plus.data(), 6);
        const double one = 1;
        addBlock(tdCol, J.td, 1, &one, 1);
        addBlock(lmCol[l], J.rho, 1, &one, 1);
        const double rr0 = wgt * r[0], rr1 = wgt * r[1];
        for (size_t i = 0; i < cols.size(); ++i) {
            b[cols[i]] += v0[i] * rr0 + v1[i] * rr1;
            for (size_t j = 0; j < cols.size(); ++j) H[(size_t)cols[i] * n + cols[j]] += v0[i] * v0[j] + v1[i] * v1[j];
        }
    }
    for (int e : imus) {
        const Frozen imu = unpackImu(&w.imu[(size_t)VIB_IMU_STRIDE * e]);
        const int s0 = w.imuFrom[e], s1 = w.imuTo[e];
        double r[15], J[4][15 * 9];
        if (!imuEvaluate(imu, &x.pose[7 * s0], &x.sb[9 * s0], &x.pose[7 * s1], &x.sb[9 * s1], r, J)) {
            out.status = VIB_ERR_INVALID_RESIDUAL; return out;
        }
        const int col[4] = {poseCol[s0], sbCol[s0], poseCol[s1], sbCol[s1]};
        const int width[4] = {6, 9, 6, 9};
        for (int bi = 0; bi < 4; ++bi)
            for (int ci = 0; ci < width[bi]; ++ci) {
                double gs = 0;
                for (int i = 0; i < 15; ++i) gs += J[bi][i * 9 + ci] * r[i];
                b[col[bi] + ci] += gs;
                for (int bj = 0; bj < 4; ++bj)
                    for (int cj = 0; cj < width[bj]; ++cj) {
                        double s = 0;
                        for (int i = 0; i < 15; ++i) s += J[bi][i * 9 + ci] * J[bj][i * 9 + cj];
                        H[(size_t)(col[bi] + ci) * n + col[bj] + cj] += s;
                    }
            }
    }
    if (w.priorRows) {
        // Build a layout in which every prior block is a variable at its marginalization column.
        Window wp = w;
        for (int k = 0; k < w.K; ++k) { wp.poseFree[k] = 1; wp.sbWidth[k] = 9; }
        wp.flags |= VIB_RIG_FREE | VIB_TD_FREE;
        Layout lay;
        lay.poseCol = poseCol; lay.sbCol = sbCol; lay.rigCol = rigCol; lay.tdCol = tdCol; lay.nr = n;
        lay.lmCol.assign(w.L, -1);
        std::vector<double> r, J;
        if (!priorEvaluate(wp, lay, x, r, &J)) { out.status = VIB_ERR_INVALID_RESIDUAL; return out; }
        for (int i = 0; i < w.priorRows; ++i) {
            const double* row = &J[(size_t)i * n];
            for (int a = 0; a < n; ++a) {
                if (row[a] == 0) continue;
                b[a] += row[a] * r[i];
                for (int c = 0; c < n; ++c) H[(size_t)a * n + c] += row[a] * row[c];
            }
        }
    }
    // partition
    std::vector<int> m, rr;
    std::vector<int> ranges, states;  // indices into m
    for (int i = 0; i < n; ++i) {
        bool elim = false;
        if (poseCol[victim] >= 0 && i >= poseCol[victim] && i < poseCol[victim] + 6) elim = true;
        if (sbCol[victim] >= 0 && i >= sbCol[victim] && i < sbCol[victim] + 9) elim = true;
        for (int l : lmKeys) if (lmCol[l] == i) elim = true;
        if (elim) {
            bool isLm = false;
            for (int l : lmKeys) if (lmCol[l] == i) isLm = true;
            (isLm ? ranges : states).push_back((int)m.size());
            m.push_back(i);
        } else rr.push_back(i);
    }
    const int nm = (int)m.size(), nrr = (int)rr.size();
    out.nm = nm; out.nr = nrr; out.landmarks = (int)lmKeys.size();
    out.Hfull = H; out.bfull = b;
    out.elimMask.assign(n, 0);
    for (int i : m) out.elimMask[i] = 1;
    std::vector<double> Hmm((size_t)nm * nm), Hmr((size_t)nm * nrr), Hrr((size_t)nrr * nrr), bm(nm), br(nrr);
    for (int i = 0; i < nm; ++i) {
        bm[i] = b[m[i]];
        for (int j = 0; j < nm; ++j) Hmm[(size_t)i * nm + j] = H[(size_t)m[i] * n + m[j]];
        for (int j = 0; j < nrr; ++j) Hmr[(size_t)i * nrr + j] = H[(size_t)m[i] * n + rr[j]];
    }
    for (int i = 0; i < nrr; ++i) {
        br[i] = b[rr[i]];
        for (int j = 0; j < nrr; ++j) Hrr[(size_t)i * nrr + j] = H[(size_t)rr[i] * n + rr[j]];
    }
    const double absCut = 1e-12, relCut = 1e-9;
    // landmark-first scalar elimination (Marginalizer.cpp:373-409)
    const int nl = (int)ranges.size(), ns = (int)states.size();
    if (nl) {
        std::vector<double> Hll((size_t)nl * nl);
        bool diag = true;
        for (int i = 0; i < nl; ++i)
            for (int j = 0; j < nl; ++j) {
                Hll[(size_t)i * nl + j] = Hmm[(size_t)ranges[i] * nm + ranges[j]];
                if (i != j && Hll[(size_t)i * nl + j] != 0.0) diag = false;
            }
        if (diag) {
            const int cols = ns + nrr;
            std::vector<double> coupling((size_t)nl * cols), rhsL((size_t)nl * (cols + 1)), solvedL;
            for (int i = 0; i < nl; ++i) {
                for (int j = 0; j < ns; ++j) coupling[(size_t)i * cols + j] = Hmm[(size_t)ranges[i] * nm + states[j]];
                for (int j = 0; j < nrr; ++j) coupling[(size_t)i * cols + ns + j] = Hmr[(size_t)ranges[i] * nrr + j];
                for (int j = 0; j < cols; ++j) rhsL[(size_t)i * (cols + 1) + j] = coupling[(size_t)i * cols + j];
                rhsL[(size_t)i * (cols + 1) + cols] = bm[ranges[i]];
            }
            if (!psdSolve(Hll, nl, rhsL, cols + 1, solvedL, absCut, relCut)) { out.status = VIB_ERR_INDEFINITE; return out; }
            std::vector<double> reduced((size_t)cols * cols, 0), gradient(cols);
            for (int i = 0; i < ns; ++i) {
                gradient[i] = bm[states[i]];
                for (int j = 0; j < ns; ++j) reduced[(size_t)i * cols + j] = Hmm[(size_t)states[i] * nm + states[j]];
                for (int j = 0; j < nrr; ++j) {
                    reduced[(size_t)i * cols + ns + j] = Hmr[(size_t)states[i] * nrr + j];
                    reduced[(size_t)(ns + j) * cols + i] = Hmr[(size_t)states[i] * nrr + j];
                }
            }
            for (int i = 0; i < nrr; ++i) {
                gradient[ns + i] = br[i];
                for (int j = 0; j < nrr; ++j) reduced[(size_t)(ns + i) * cols + ns + j] = Hrr[(size_t)i * nrr + j];
            }
            // subtract: reduced -= coupling^T * solved(:, :cols); gradient -= coupling^T * solved(:, cols)
            for (int a = 0; a < cols; ++a) {
                for (int c = 0; c < cols; ++c) {
                    double s = 0;
                    for (int k = 0; k < nl; ++k) s += coupling[(size_t)k * cols + a] * solvedL[(size_t)k * (cols + 1) + c];
                    reduced[(size_t)a * cols + c] -= s;
                }
                double s = 0;
                for (int k = 0; k < nl; ++k) s += coupling[(size_t)k * cols + a] * solvedL[(size_t)k * (cols + 1) + cols];
                gradient[a] -= s;
            }
            // Hmm/Hmr/Hrr/bm/br now refer to the state part only
            std::vector<double> nHmm((size_t)ns * ns), nHmr((size_t)ns * nrr), nHrr((size_t)nrr * nrr), nbm(ns), nbr(nrr);
            for (int i = 0; i < ns; ++i) {
                nbm[i] = gradient[i];
                for (int j = 0; j < ns; ++j) nHmm[(size_t)i * ns + j] = reduced[(size_t)i * cols + j];
                for (int j = 0; j < nrr; ++j) nHmr[(size_t)i * nrr + j] = reduced[(size_t)i * cols + ns + j];
            }
            for (int i = 0; i < nrr; ++i) {
                nbr[i] = gradient[ns + i];
                for (int j = 0; j < nrr; ++j) nHrr[(size_t)i * nrr + j] = reduced[(size_t)(ns + i) * cols + ns + j];
            }
            Hmm = nHmm; Hmr = nHmr; Hrr = nHrr; bm = nbm; br = nbr;
        }
    }
    const int nmm = (int)bm.size();
    // joint PSD solve of the remaining eliminated block
    std::vector<double> rhs((size_t)nmm * (nrr + 1)), solved;
    for (int i = 0; i < nmm; ++i) {
        for (int j = 0; j < nrr; ++j) rhs[(size_t)i * (nrr + 1) + j] = Hmr[(size_t)i * nrr + j];
        rhs[(size_t)i * (nrr + 1) + nrr] = bm[i];
    }
    if (!psdSolve(Hmm, nmm, rhs, nrr + 1, solved, absCut, relCut)) { out.status = VIB_ERR_INDEFINITE; return out; }
    for (int a = 0; a < nrr; ++a) {
        for (int c = 0; c < nrr; ++c) {
            double s = 0;
            for (int k = 0; k < nmm; ++k) s += Hmr[(size_t)k * nrr + a] * solved[(size_t)k * (nrr + 1) + c];
            Hrr[(size_t)a * nrr + c] -= s;
        }
        double s = 0;
        for (int k = 0; k < nmm; ++k) s += Hmr[(size_t)k * nrr + a] * solved[(size_t)k * (nrr + 1) + nrr];
        br[a] -= s;
    }
    std::vector<double> Hs((size_t)nrr * nrr);
    for (int i = 0; i < nrr; ++i)
        for (int j = 0; j < nrr; ++j) Hs[(size_t)i * nrr + j] = 0.5 * (Hrr[(size_t)i * nrr + j] + Hrr[(size_t)j * nrr + i]);
    out.Hprior = Hs; out.bprior = br;
    std::vector<double> ev, U;
    if (!jacobiEigen(Hs, nrr, ev, U)) { out.status = VIB_ERR_EIGEN_FAILED; return out; }
    const double cutoff = std::max(absCut, relCut * std::max(0.0, ev[nrr - 1]));
    out.cutoff = cutoff;
    if (ev[0] < -10.0 * cutoff) { out.status = VIB_ERR_INDEFINITE; return out; }
    for (int i = 0; i < nrr; ++i) if (ev[i] > cutoff) ++out.rank;
    out.cols = nrr;
    for (int i = 0; i < nrr; ++i) {
        if (!(ev[i] > cutoff)) continue;
        const double root = std::sqrt(ev[i]);
        double ub = 0;
        for (int k = 0; k < nrr; ++k) { out.A.push_back(root * U[(size_t)k * nrr + i]); ub += U[(size_t)k * nrr + i] * br[k]; }
        out.r0.push_back(ub / root);
    }
    // retained keys in column order
    for (int s : poseSlots) if (s != victim) { out.keyKind.push_back(VIB_KEY_POSE); out.keyId.push_back(w.frameId[s]);
        for (int i = 0; i < 7; ++i) out.ref.push_back(x.pose[7 * s + i]); }
    for (int s : sbSlots) if (s != victim) { out.keyKind.push_back(VIB_KEY_SPEED_BIAS); out.keyId.push_back(w.frameId[s]);
        for (int i = 0; i < 9; ++i) out.ref.push_back(x.sb[9 * s + i]); }
    if (rigUsed) { out.keyKind.push_back(VIB_KEY_RIG); out.keyId.push_back(0); for (int i = 0; i < 7; ++i) out.ref.push_back(x.rig[i]); }
    if (tdUsed) { out.keyKind.push_back(VIB_KEY_TD); out.keyId.push_back(0); out.ref.push_back(x.td); }
    return out;
}

// OLD: the victim is slot 0 with every landmark anchored in or observed by it
// (SlidingWindowManager.cpp:624-650 with preserveOldestVisualInformation).
MargResult marginalizeOld(const Window& w) {
    const int victim = 0;
    std::vector<char> elimLm(w.L, 0);
    for (int l = 0; l < w.L; ++l) if (w.lmFrame[l] == victim) elimLm[l] = 1;
    for (int f = 0; f < w.F; ++f) if (w.obsFrame[f] == victim) elimLm[w.obsLm[f]] = 1;
    std::vector<int> visual;
    for (int f = 0; f < w.F; ++f) {
        const int l = w.obsLm[f];
        if (elimLm[l] || w.obsFrame[f] == victim || w.lmFrame[l] == victim) visual.push_back(f);
    }
    std::vector<int> imus;
    for (int e = 0; e < w.E; ++e) if (w.imuFrom[e] == victim || w.imuTo[e] == victim) imus.push_back(e);
    return marginalizeSet(w, victim, elimLm, visual, imus);
}

// ------------------------------------------------------------------------------------
// IMU preintegration (ImuError::redoPreintegration, ImuError.cpp:254-462) and
// PseudoInverse::symmSqrtU (PseudoInverse.hpp:116-145).
// ------------------------------------------------------------------------------------
double sinc(double x) {  // ode.hpp:15-27
    if (std::fabs(x) > 1e-6) return std::sin(x) / x;
    const double x2 = x * x, x4 = x2 * x2, x6 = x2 * x2 * x2;
    return 1.0 - x2 / 6.0 + x4 / 120.0 - x6 / 5040.0;
}
M3 rightJacobian(const V3& phiVec) {  // Transformation.hpp:47-62
    const double phi = norm(phiVec);
    M3 ret = M3::identity();
    const M3 px = skew(phiVec), px2 = px * px;
    if (phi < 1.0e-4) ret = ret + (-0.5) * px + (1.0 / 6.0) * px2;
    else {
        const double phi2 = phi * phi, phi3 = phi2 * phi;
        ret = ret + (-(1.0 - std::cos(phi)) / phi2) * px + ((phi - std::sin(phi)) / phi3) * px2;
    }
    return ret;
}
template <int N>
M<N, N> mulABt(const M<N, N>& A, const M<N, N>& B) { return A * transpose(B); }

int preintegrate(const double* samples, int count, double t0, double t1, const double* sbLin,
                 const vib_imu_params& p, const double* gyroStart, const double* gyroEnd, double* out,
                 double* covOut = nullptr) {
    if (count < 2 || !(t1 > t0)) return VIB_ERR_INVALID_ARGUMENT;
    auto T = [&](int i) { return samples[i * 7]; };
    if (!(T(0) <= t0) || !(T(count - 1) >= t1)) return VIB_ERR_INVALID_ARGUMENT;
    double time = t0;
    Q Dq(1, 0, 0, 0);
    M3 Cint, Cdint, cross_, dadbg, dvdbg, dpdbg;
    V3 accInt, accDint;
    M<15, 15> dP[4];
    const V3 bg(sbLin + 3), ba(sbLin + 6);
    bool started = false;
    int steps = 0;
    for (int i = 0; i + 1 < count; ++i) {
        V3 w0(&samples[i * 7 + 1]), a0(&samples[i * 7 + 4]), w1(&samples[(i + 1) * 7 + 1]), a1(&samples[(i + 1) * 7 + 4]);
        double nexttime = T(i + 1);
        double dt = nexttime - time;
        if (t1 < nexttime) {
            const double interval = nexttime - T(i);
            nexttime = t1;
            dt = nexttime - time;
            const double r = dt / interval;
            w1 = (1.0 - r) * w0 + r * w1;
            a1 = (1.0 - r) * a0 + r * a1;
        }
        if (dt <= 0.0) continue;
        if (!started) {
            started = true;
            const double r = dt / (nexttime - T(i));
            w0 = r * w0 + (1.0 - r) * w1;
            a0 = r * a0 + (1.0 - r) * a1;
        }
        double gsat = 1.0, asat = 1.0;
        for (int k = 0; k < 3; ++k) {
            if (std::fabs(w0[k]) > p.g_max || std::fabs(w1[k]) > p.g_max) gsat = 100;
            if (std::fabs(a0[k]) > p.a_max || std::fabs(a1[k]) > p.a_max) asat = 100;
        }
        const V3 wtrue = 0.5 * (w0 + w1) - bg;
        const double th = norm(wtrue) * 0.5 * dt;
        const double sth = sinc(th), cth = std::cos(th);
        const V3 dqv = (sth * 0.5 * dt) * wtrue;
        const Q dq(cth, dqv.x, dqv.y, dqv.z);
        const Q Dq1 = Dq * dq;
        const M3 C = Dq.matrix(), C1 = Dq1.matrix();
        const V3 atrue = 0.5 * (a0 + a1) - ba;
        const M3 CC = C + C1;
        const M3 Cint1 = Cint + (0.5 * dt) * CC;
        const V3 accInt1 = accInt + (0.5 * dt) * (CC * atrue);
        Cdint = Cdint + dt * Cint + (0.25 * dt * dt) * CC;
        accDint = accDint + dt * accInt + (0.25 * dt * dt) * (CC * atrue);
        const M3 Jr = rightJacobian(dt * wtrue);
        dadbg = dadbg + dt * (C1 * Jr);
        const M3 cross1 = dq.conjugate().matrix() * cross_ + dt * Jr;
        const M3 ax = skew(atrue);
        const M3 mix = C * ax * cross_ + C1 * ax * cross1;
        const M3 dvdbg1 = dvdbg + (0.5 * dt) * mix;
        dpdbg = dpdbg + dt * dvdbg + (0.25 * dt * dt) * mix;
        M<15, 15> F = M<15, 15>::identity();
        auto put = [&](int r0, int c0, const M3& m) { for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) F(r0 + a, c0 + b) = m(a, b); };
        put(0, 3, -1.0 * skew(dt * accInt + (0.25 * dt * dt) * (CC * atrue)));
        put(0, 6, diag3(dt));
        put(0, 9, dt * dvdbg + (0.25 * dt * dt) * mix);
        put(0, 12, -1.0 * (dt * Cint) + (0.25 * dt * dt) * CC);
        put(3, 9, -dt * C1);
        put(6, 3, -1.0 * skew((0.5 * dt) * (CC * atrue)));
        put(6, 9, (0.5 * dt) * mix);
        put(6, 12, (-0.5 * dt) * CC);
        M<15, 15> Kq[4];
        for (int a = 0; a < 3; ++a) {
            Kq[0](3 + a, 3 + a) = gsat * dt;
            Kq[1](0 + a, 0 + a) = 0.5 * dt * dt * dt * asat * asat * asat;
            Kq[1](6 + a, 6 + a) = asat * dt;
            Kq[2](9 + a, 9 + a) = dt;
            Kq[3](12 + a, 12 + a) = dt;
        }
        for (int j = 0; j < 4; ++j) dP[j] = mulABt(F * dP[j], F) + Kq[j];
        Dq = Dq1; Cint = Cint1; accInt = accInt1; cross_ = cross1; dvdbg = dvdbg1;
        time = nexttime;
        ++steps;
        if (nexttime == t1) break;
    }
    if (steps <= 0) return VIB_ERR_INVALID_ARGUMENT;
    const double sig[4] = {p.sigma_g_c, p.sigma_a_c, p.sigma_gw_c, p.sigma_aw_c};
    std::vector<double> P(225, 0);
    for (int j = 0; j < 4; ++j) {
        const M<15, 15> sym = 0.5 * dP[j] + 0.5 * transpose(dP[j]);
        for (int i = 0; i < 225; ++i) P[i] += sym.a[i] * sig[j] * sig[j];
    }
    if (covOut) std::copy(P.begin(), P.end(), covOut);
    std::vector<double> ev, U;
    if (!jacobiEigen(P, 15, ev, U)) return VIB_ERR_EIGEN_FAILED;
    const double eps = std::numeric_limits<double>::epsilon();
    const double tol = std::max(eps, eps * 15 * ev[14]);
    double* o = out;
    for (int i = 0; i < VIB_IMU_STRIDE; ++i) o[i] = 0;
    o[0] = t1 - t0; o[1] = p.g;
    for (int i = 0; i < 3; ++i) o[2 + i] = p.p_BS[i];
    for (int i = 0; i < 4; ++i) o[5 + i] = p.q_BS[i];
    for (int i = 0; i < 3; ++i) { o[9 + i] = gyroStart[i]; o[12 + i] = gyroEnd[i]; }
    Dq.storeXyzw(o + 15);
    accInt.store(o + 19); accDint.store(o + 22);
    for (int i = 0; i < 6; ++i) o[25 + i] = sbLin[3 + i];
    for (int i = 0; i < 9; ++i) { o[31 + i] = dadbg.a[i]; o[40 + i] = dvdbg.a[i]; o[49 + i] = dpdbg.a[i]; o[58 + i] = Cint.a[i]; o[67 + i] = Cdint.a[i]; }
    for (int i = 0; i < 15; ++i) {
        const double s = std::sqrt(ev[i] > tol ? 1.0 / ev[i] : 1.0 / tol);
        for (int k = 0; k < 15; ++k) o[76 + i * 15 + k] = s * U[(size_t)k * 15 + i];
    }
    return VIB_OK;
}

}  // namespace
// ------------------------------------------------------------------------------------
// Two-frame lightweight solve (BackendOptimizer.cpp:476-590; the velocity-only speed/bias
// tangent of the scoped solve, :594,783-790). Free: pose and velocity of the two newest slots;
// everything else constant; only factors with a key among the free blocks.
// ------------------------------------------------------------------------------------
double imuNis(const Window& w, const State& x, int e) {
    const Frozen imu = unpackImu(&w.imu[(size_t)VIB_IMU_STRIDE * e]);
    const int s0 = w.imuFrom[e], s1 = w.imuTo[e];
    double r[15];
    if (!imuEvaluate(imu, &x.pose[7 * s0], &x.sb[9 * s0], &x.pose[7 * s1], &x.sb[9 * s1], r, nullptr))
        return std::numeric_limits<double>::quiet_NaN();
    double v = 0;
    for (double ri : r) v += ri * ri;
    return v;
}

int latestTwo(Window& w, const vib_solve_options& opt, int gate, double* nis, int& support, vib_solve_report& rep) {
    for (int i = 0; i < 4; ++i) nis[i] = std::numeric_limits<double>::quiet_NaN();
    support = 0;
    if (w.K < 2) return VIB_ERR_LIGHTWEIGHT_SUPPORT;
    const int previous = w.K - 2, latest = w.K - 1;
    auto isFree = [&](int k) { return k == previous || k == latest; };
    Window loc = w;
    loc.flags &= ~(VIB_RIG_FREE | VIB_TD_FREE);
    for (int k = 0; k < w.K; ++k) { loc.poseFree[k] = isFree(k); loc.sbWidth[k] = isFree(k) ? 3 : 0; }
    // visual factors: keys are anchor pose, observer pose, rig, td, landmark
    std::vector<int> lmMap(w.L, -1);
    loc.L = 0; loc.lmFrame.clear(); loc.lmCam.clear(); loc.lmBearing.clear(); loc.lmVel.clear();
    loc.lmTref.clear(); loc.rho.clear(); loc.lmFree.clear();
    loc.F = 0; loc.obsLm.clear(); loc.obsFrame.clear(); loc.obsCam.clear(); loc.obsBearing.clear();
    loc.obsVel.clear(); loc.obsTref.clear(); loc.obsSqrt.clear();
    std::set<int> current;
    for (int f = 0; f < w.F; ++f) {
        const int l = w.obsLm[f];
        if (!isFree(w.lmFrame[l]) && !isFree(w.obsFrame[f])) continue;
        if (lmMap[l] < 0) {
            lmMap[l] = loc.L++;
            loc.lmFrame.push_back(w.lmFrame[l]); loc.lmCam.push_back(w.lmCam[l]);
            for (int i = 0; i < 3; ++i) { loc.lmBearing.push_back(w.lmBearing[3 * l + i]); loc.lmVel.push_back(w.lmVel[3 * l + i]); }
            loc.lmTref.push_back(w.lmTref[l]); loc.rho.push_back(w.rho[l]); loc.lmFree.push_back(0);
        }
        ++loc.F;
        loc.obsLm.push_back(lmMap[l]); loc.obsFrame.push_back(w.obsFrame[f]); loc.obsCam.push_back(w.obsCam[f]);
        for (int i = 0; i < 3; ++i) {
            loc.obsBearing.push_back(w.obsBearing[3 * f + i]); loc.obsVel.push_back(w.obsVel[3 * f + i]);
            loc.obsSqrt.push_back(w.obsSqrt[3 * f + i]);
        }
        loc.obsTref.push_back(w.obsTref[f]);
        if (w.obsFrame[f] == latest) current.insert(l);
    }
    support = (int)current.size();
    if (support < 6) return VIB_ERR_LIGHTWEIGHT_SUPPORT;
    // IMU factors touching a free block: exactly the edge into `previous` and previous -> latest
    int oldEdge = -1, newEdge = -1;
    loc.E = 0; loc.imuFrom.clear(); loc.imuTo.clear(); loc.imu.clear();
    for (int e = 0; e < w.E; ++e) {
        if (!isFree(w.imuFrom[e]) && !isFree(w.imuTo[e])) continue;
        if (w.imuTo[e] == previous) oldEdge = loc.E;
        if (w.imuFrom[e] == previous && w.imuTo[e] == latest) newEdge = loc.E;
        ++loc.E;
        loc.imuFrom.push_back(w.imuFrom[e]); loc.imuTo.push_back(w.imuTo[e]);
        loc.imu.insert(loc.imu.end(), w.imu.begin() + (size_t)VIB_IMU_STRIDE * e, w.imu.begin() + (size_t)VIB_IMU_STRIDE * (e + 1));
    }
    if (loc.E != 2 || oldEdge < 0 || newEdge < 0) return VIB_ERR_IMU_BRIDGE;
    // the prior participates when it has a free key
    bool touches = false;
    for (int b = 0; b < w.priorBlocks; ++b)
        if ((w.priorKind[b] == VIB_KEY_POSE || w.priorKind[b] == VIB_KEY_SPEED_BIAS) && isFree(slotOfFrameId(w, w.priorId[b])))
            touches = true;
    if (!touches) { loc.priorRows = 0; loc.priorBlocks = 0; }
    const State x0 = stateOf(loc);
    nis[0] = imuNis(loc, x0, oldEdge);
    nis[2] = imuNis(loc, x0, newEdge);
    if (!std::isfinite(nis[0]) || !std::isfinite(nis[2])) return VIB_ERR_IMU_BRIDGE;
    vib_solve_options o = opt;
    o.max_iterations = std::min(3, std::max(1, opt.max_iterations));
    o.min_iterations = std::min(1, std::max(0, opt.min_iterations));
    State x;
    const int st = lmSolve(loc, o, rep, x);
    if (st != VIB_OK) return st;
    nis[1] = imuNis(loc, x, oldEdge);
    nis[3] = imuNis(loc, x, newEdge);
    if (!std::isfinite(nis[1]) || !std::isfinite(nis[3])) return VIB_ERR_IMU_BRIDGE;
    if (gate && nis[3] > 60.0) return VIB_ERR_IMU_GATE;
    for (int k : {previous, latest}) {
        for (int i = 0; i < 7; ++i) w.pose[7 * k + i] = x.pose[7 * k + i];
        for (int i = 0; i < 3; ++i) w.sb[9 * k + i] = x.sb[9 * k + i];
    }
    return VIB_OK;
}

}  // namespace oracle

using namespace oracle;

// ------------------------------------------------------------------------------------
// SECOND_NEW: the redundant-frame path (SlidingWindowManager.cpp:263-482, 587-680;
// BodyImuFactor.cpp:306-388).
// ------------------------------------------------------------------------------------
// Window without slot `drop` (-1: none), restricted to the kept landmarks, factors and IMU
// edges; slots above `drop` move down by one. The prior is copied unchanged.
Window filterWindow(const Window& w, int drop, const std::vector<char>& keepLm, const std::vector<char>& keepF,
                    const std::vector<char>& keepE) {
    Window o = w;
    auto slot = [&](int k) { return drop >= 0 && k > drop ? k - 1 : k; };
    if (drop >= 0) {
        o.K = w.K - 1;
        o.frameId.erase(o.frameId.begin() + drop);
        o.pose.erase(o.pose.begin() + 7 * drop, o.pose.begin() + 7 * drop + 7);
        o.sb.erase(o.sb.begin() + 9 * drop, o.sb.begin() + 9 * drop + 9);
        o.poseFree.erase(o.poseFree.begin() + drop);
        o.sbWidth.erase(o.sbWidth.begin() + drop);
    }
    std::vector<int> newLm(w.L, -1);
    o.lmFrame.clear(); o.lmCam.clear(); o.lmBearing.clear(); o.lmVel.clear(); o.lmTref.clear(); o.rho.clear(); o.lmFree.clear();
    for (int l = 0; l < w.L; ++l) {
        if (!keepLm[l]) continue;
        newLm[l] = (int)o.lmFrame.size();
        o.lmFrame.push_back(slot(w.lmFrame[l])); o.lmCam.push_back(w.lmCam[l]);
        for (int i = 0; i < 3; ++i) { o.lmBearing.push_back(w.lmBearing[3 * l + i]); o.lmVel.push_back(w.lmVel[3 * l + i]); }
        o.lmTref.push_back(w.lmTref[l]); o.rho.push_back(w.rho[l]); o.lmFree.push_back(w.lmFree[l]);
    }
    o.L = (int)o.lmFrame.size();
    o.obsLm.clear(); o.obsFrame.clear(); o.obsCam.clear(); o.obsBearing.clear(); o.obsVel.clear(); o.obsTref.clear(); o.obsSqrt.clear();
    for (int f = 0; f < w.F; ++f) {
        if (!keepF[f] || newLm[w.obsLm[f]] < 0 || w.obsFrame[f] == drop) continue;
        o.obsLm.push_back(newLm[w.obsLm[f]]); o.obsFrame.push_back(slot(w.obsFrame[f])); o.obsCam.push_back(w.obsCam[f]);
        for (int i = 0; i < 3; ++i) {
            o.obsBearing.push_back(w.obsBearing[3 * f + i]); o.obsVel.push_back(w.obsVel[3 * f + i]);
            o.obsSqrt.push_back(w.obsSqrt[3 * f + i]);
        }
        o.obsTref.push_back(w.obsTref[f]);
    }
    o.F = (int)o.obsLm.size();
    o.imuFrom.clear(); o.imuTo.clear(); o.imu.clear();
    for (int e = 0; e < w.E; ++e) {
        if (!keepE[e] || w.imuFrom[e] == drop || w.imuTo[e] == drop) continue;
        o.imuFrom.push_back(slot(w.imuFrom[e])); o.imuTo.push_back(slot(w.imuTo[e]));
        o.imu.insert(o.imu.end(), w.imu.begin() + (size_t)VIB_IMU_STRIDE * e, w.imu.begin() + (size_t)VIB_IMU_STRIDE * (e + 1));
    }
    o.E = (int)o.imuFrom.size();
    return o;
}

// Linear interpolation of the gyroscope at t (gyroAt, BodyImuFactor.cpp:29-38).
V3 gyroAt(const double* samples, int count, double t) {
    int next = 0;
    while (next < count && samples[7 * next] < t) ++next;
    if (next == 0) return V3(&samples[1]);
    if (next == count) return V3(&samples[7 * (count - 1) + 1]);
    const double t0 = samples[7 * (next - 1)], alpha = (t - t0) / (samples[7 * next] - t0);
    return (1 - alpha) * V3(&samples[7 * (next - 1) + 1]) + alpha * V3(&samples[7 * next + 1]);
}

// World pose of camera c of slot k: T_WB T_BR T_RC.
void cameraWorld(const Window& w, int k, int c, M3& R, V3& t) {
    const M3 Rb = Q::fromXyzw(&w.pose[7 * k + 3]).normalized().matrix();
    const M3 Re = Q::fromXyzw(w.rig + 3).normalized().matrix();
    const M3 Rc = fromPtr(&w.camR[9 * c]);
    R = Rb * Re * Rc;
    t = Rb * (Re * V3(&w.camT[3 * c]) + V3(w.rig)) + V3(&w.pose[7 * k]);
}

struct SecondNewResult {
    int status = VIB_OK;
    int imuMode = VIB_IMU_MERGE_NONE;
    int reset = 0, demoted = 0, retired = 0, dropped = 0;
    bool priorReused = false;
    MargResult marg;
    Window next;
};

SecondNewResult secondNew(const Window& w0, const double* samples, int count, double t0, double t1,
                          const vib_imu_params* params) {
    SecondNewResult out;
    Window w = w0;
    const int K = w.K;
    if (K < 3) { out.status = VIB_ERR_INVALID_ARGUMENT; return out; }
    const int v = K - 2, left = K - 3, right = K - 1;
    // IMU (prepareEviction, SlidingWindowManager.cpp:275-311)
    std::vector<int> incident;
    for (int e = 0; e < w.E; ++e) if (w.imuFrom[e] == v || w.imuTo[e] == v) incident.push_back(e);
    bool sbInPrior = false, priorTouches = false;
    for (int b = 0; b < w.priorBlocks; ++b)
        if ((w.priorKind[b] == VIB_KEY_SPEED_BIAS || w.priorKind[b] == VIB_KEY_POSE) && w.priorId[b] == w.frameId[v]) {
            priorTouches = true;
            if (w.priorKind[b] == VIB_KEY_SPEED_BIAS) sbInPrior = true;
        }
    out.imuMode = incident.empty() ? VIB_IMU_MERGE_NONE : VIB_IMU_MERGE_BRIDGE;
    std::vector<char> keepE(w.E, 1);
    std::vector<double> merged;
    if (incident.size() == 2 && !sbInPrior && samples && count > 0 && params) {
        int a = incident[0], b = incident[1];
        if (w.imuFrom[a] != left) std::swap(a, b);
        if (w.imuFrom[a] != left || w.imuTo[a] != v || w.imuFrom[b] != v || w.imuTo[b] != right) {
            out.status = VIB_ERR_INVALID_ARGUMENT; return out;  // "IMU merge endpoint/layout mismatch"
        }
        const double* da = &w.imu[(size_t)VIB_IMU_STRIDE * a];
        const double* db = &w.imu[(size_t)VIB_IMU_STRIDE * b];
        if (!(t1 > t0) || std::fabs((t1 - t0) - (da[0] + db[0])) > 1e-9) { out.status = VIB_ERR_INVALID_ARGUMENT; return out; }
        if (count > 10000) { out.status = VIB_ERR_INVALID_ARGUMENT; return out; }  // maxImuSamplesPerEpoch
        for (int i = 0; i < count; ++i) {
            for (int c = 0; c < 7; ++c) if (!std::isfinite(samples[7 * i + c])) { out.status = VIB_ERR_INVALID_ARGUMENT; return out; }
            if (i && !(samples[7 * i] > samples[7 * (i - 1)])) { out.status = VIB_ERR_INVALID_ARGUMENT; return out; }
        }
        keepE[a] = keepE[b] = 0;
        if (t1 - t0 > 10.0) {
            out.imuMode = VIB_IMU_MERGE_SUPPRESSED;  // maxImuIntervalSeconds: kept out of the solve
        } else {
            // incremental reuse keeps the left edge's reference bias while the estimate is within
            // 0.0003 rad/s (gyro) and 0.02 m/s^2 (accel) of it (BodyImuFactor.cpp:146-170)
            double sbLin[9];
            std::copy_n(&w.sb[9 * left], 9, sbLin);
            const V3 dbg = V3(&sbLin[3]) - V3(da + 25), dba = V3(&sbLin[6]) - V3(da + 28);
            const bool incremental = norm(dbg) <= 3e-4 && norm(dba) <= 0.02;
            if (incremental) for (int i = 0; i < 6; ++i) sbLin[3 + i] = da[25 + i];
            const V3 g0 = gyroAt(samples, count, t0), g1 = gyroAt(samples, count, t1);
            double gs[3], ge[3];
            g0.store(gs); g1.store(ge);
            merged.assign(VIB_IMU_STRIDE, 0);
            const int st = preintegrate(samples, count, t0, t1, sbLin, *params, gs, ge, merged.data());
            if (st != VIB_OK) { out.status = st; return out; }
            out.imuMode = incremental ? VIB_IMU_MERGE_INCREMENTAL : VIB_IMU_MERGE_REINTEGRATED;
        }
    }
    if (!merged.empty()) {
        w.imuFrom.push_back(left); w.imuTo.push_back(right);
        w.imu.insert(w.imu.end(), merged.begin(), merged.end());
        w.E += 1;
        keepE.push_back(1);
    }
    // landmarks anchored in the victim: reset to a surviving view, demote or retire (:312-457)
    std::vector<std::vector<int>> ofLm(w.L);
    for (int f = 0; f < w.F; ++f) ofLm[w.obsLm[f]].push_back(f);
    std::vector<char> keepLm(w.L, 1), keepF(w.F, 1), retired(w.L, 0);
    const State x = stateOf(w);
    for (int l = 0; l < w.L; ++l) {
        if (w.lmFrame[l] != v) continue;
        std::vector<int> surviving;
        bool eligible = sqnorm(V3(&w.lmVel[3 * l])) == 0;
        for (int f : ofLm[l]) {
            eligible = eligible && sqnorm(V3(&w.obsVel[3 * f])) == 0;
            if (w.obsFrame[f] != v) surviving.push_back(f);
        }
        if (surviving.empty() || !eligible || !w.lmFree[l]) { retired[l] = 1; continue; }
        if (surviving.size() == 1) {  // one surviving ray: demoted out of the window
            keepLm[l] = 0;
            for (int f : ofLm[l]) keepF[f] = 0;
            ++out.demoted;
            continue;
        }
        // preferred anchors: newer frames first, then lower camera index (all frames are keyframes here)
        std::stable_sort(surviving.begin(), surviving.end(), [&](int a, int b) {
            if (w.obsFrame[a] != w.obsFrame[b]) return w.obsFrame[a] > w.obsFrame[b];
            return w.obsCam[a] < w.obsCam[b];
        });
        M3 Ra; V3 ta;
        cameraWorld(w, v, w.lmCam[l], Ra, ta);
        const V3 world = Ra * (normalized(V3(&w.lmBearing[3 * l])) / w.rho[l]) + ta;
        int chosen = -1;
        double newRho = 0;
        for (int cand : surviving) {
            M3 Rc; V3 tc;
            cameraWorld(w, w.obsFrame[cand], w.obsCam[cand], Rc, tc);
            const V3 delta = transpose(Rc) * (world - tc);
            const V3 ray(&w.obsBearing[3 * cand]);
            const double range = dot(delta, ray);
            if (!(range > 0) || !std::isfinite(range) || 1 / range < w.minRho ||
                std::atan2(norm(cross(delta, ray)), range) > 0.03) continue;
            // the rebuilt factors must evaluate (the NIS threshold of the factors is disabled)
            Window t = w;
            t.lmFrame[l] = w.obsFrame[cand]; t.lmCam[l] = w.obsCam[cand];
            for (int i = 0; i < 3; ++i) { t.lmBearing[3 * l + i] = w.obsBearing[3 * cand + i]; t.lmVel[3 * l + i] = w.obsVel[3 * cand + i]; }
            t.lmTref[l] = w.obsTref[cand];
            bool valid = true;
            for (int f : surviving) {
                if (f == cand) continue;
                double r[2];
                if (!visualEvaluate(t, f, &x.pose[7 * t.lmFrame[l]], &x.pose[7 * w.obsFrame[f]], x.rig, x.td, 1 / range, r,
                                    nullptr, false, false, false, false, false)) { valid = false; break; }
            }
            if (valid) { chosen = cand; newRho = 1 / range; break; }
        }
        if (chosen < 0) { retired[l] = 1; continue; }
        w.lmFrame[l] = w.obsFrame[chosen]; w.lmCam[l] = w.obsCam[chosen];
        for (int i = 0; i < 3; ++i) { w.lmBearing[3 * l + i] = w.obsBearing[3 * chosen + i]; w.lmVel[3 * l + i] = w.obsVel[3 * chosen + i]; }
        w.lmTref[l] = w.obsTref[chosen];
        w.rho[l] = newRho;
        keepF[chosen] = 0;
        for (int f : ofLm[l]) if (w.obsFrame[f] == v) keepF[f] = 0;
        ++out.reset;
    }
    // the victim's image measurements of landmarks anchored elsewhere are discarded (:458-470)
    for (int f = 0; f < w.F; ++f)
        if (w.obsFrame[f] == v && w.lmFrame[w.obsLm[f]] != v && keepF[f]) { keepF[f] = 0; ++out.dropped; }
    // Schur elimination of the victim with the retired landmarks (:623-660), or prior reuse (:299-310)
    Window w1 = filterWindow(w, -1, keepLm, keepF, keepE);
    std::vector<char> ret1;
    for (int l = 0; l < w.L; ++l) if (keepLm[l]) ret1.push_back(retired[l]);
    std::vector<int> visual, imus;
    for (int f = 0; f < w1.F; ++f) if (ret1[w1.obsLm[f]]) visual.push_back(f);
    for (int e = 0; e < w1.E; ++e) if (w1.imuFrom[e] == v || w1.imuTo[e] == v) imus.push_back(e);
    for (char r : ret1) out.retired += r;
    std::vector<char> keepLm1(w1.L), keepF1(w1.F, 1), keepE1(w1.E, 1);
    for (int l = 0; l < w1.L; ++l) keepLm1[l] = !ret1[l];
    for (int f : visual) keepF1[f] = 0;
    for (int e : imus) keepE1[e] = 0;
    Window next = filterWindow(w1, v, keepLm1, keepF1, keepE1);
    if (visual.empty() && imus.empty() && !priorTouches) {
        out.priorReused = true;
        out.next = next;
        return out;
    }
    out.marg = marginalizeSet(w1, v, ret1, visual, imus);
    if (out.marg.status != VIB_OK) { out.status = out.marg.status; return out; }
    next.priorRows = out.marg.rank; next.priorBlocks = (int)out.marg.keyKind.size();
    next.priorKind = out.marg.keyKind; next.priorId = out.marg.keyId; next.priorRef = out.marg.ref;
    next.priorA = out.marg.A; next.priorR0 = out.marg.r0;
    out.next = next;
    return out;
}

// Copies a window into caller arrays (capacities checked by the caller through `sizes`).
void storeWindow(const Window& w, vib_window* o) {
    auto put = [](const auto& v, auto* p) { if (p && !v.empty()) std::copy(v.begin(), v.end(), const_cast<std::remove_const_t<std::remove_pointer_t<decltype(p)>>*>(p)); };
    o->num_frames = w.K; o->num_cameras = w.C; o->num_landmarks = w.L; o->num_observations = w.F; o->num_imu = w.E;
    o->flags = w.flags; o->huber_delta = w.huber; o->min_inverse_depth = w.minRho;
    o->td_lower = w.tdLower; o->td_upper = w.tdUpper;
    put(w.frameId, o->frame_id); put(w.pose, o->pose); put(w.sb, o->speed_bias);
    put(w.poseFree, o->pose_free); put(w.sbWidth, o->sb_width);
    std::copy_n(w.rig, 7, o->rig); *o->td = w.td;
    put(w.camR, o->cam_rotation); put(w.camT, o->cam_translation);
    put(w.lmFrame, o->lm_anchor_frame); put(w.lmCam, o->lm_anchor_cam); put(w.lmBearing, o->lm_bearing);
    put(w.lmVel, o->lm_velocity); put(w.lmTref, o->lm_tref); put(w.rho, o->lm_rho); put(w.lmFree, o->lm_free);
    put(w.obsLm, o->obs_landmark); put(w.obsFrame, o->obs_frame); put(w.obsCam, o->obs_cam);
    put(w.obsBearing, o->obs_bearing); put(w.obsVel, o->obs_velocity); put(w.obsTref, o->obs_tref); put(w.obsSqrt, o->obs_sqrt_info);
    put(w.imuFrom, o->imu_from); put(w.imuTo, o->imu_to); put(w.imu, o->imu_data);
    o->prior_rows = w.priorRows; o->prior_blocks = w.priorBlocks;
    put(w.priorKind, o->prior_key_kind); put(w.priorId, o->prior_key_id); put(w.priorRef, o->prior_ref);
    put(w.priorA, o->prior_A); put(w.priorR0, o->prior_r0);
}

extern "C" {

// Evaluate one visual factor: residual (2) and tangent Jacobian blocks at the window state:
// J layout (2 x 20 row-major): anchor(6) observer(6) rig(6) td(1) rho(1), unweighted (no Huber).
// Also returns the ambient Jacobians (2 x 30: anchor 7, observer 7, rig 7, td 1, rho 1) when
// `ambient` is non-null. Returns 1 if valid.
int oracle_visual_factor(const vib_window* win, int f, double* residual, double* J, double* ambient) {
    Window w = load(*win);
    const int l = w.obsLm[f], a = w.lmFrame[l], b = w.obsFrame[f];
    double r[2];
    VisualJ vj;
    std::fill(vj.anchor, vj.anchor + 14, 0.0); std::fill(vj.observer, vj.observer + 14, 0.0);
    std::fill(vj.rig, vj.rig + 14, 0.0); vj.td[0] = vj.td[1] = vj.rho[0] = vj.rho[1] = 0;
    const bool ok = visualEvaluate(w, f, &w.pose[7 * a], &w.pose[7 * b], w.rig, w.td, w.rho[l], r,
                                   J || ambient ? &vj : nullptr, true, true, true, true, true);
    if (!ok) return 0;
    residual[0] = r[0]; residual[1] = r[1];
    if (ambient) {
        for (int row = 0; row < 2; ++row) {
            double* o = ambient + row * 23;
            for (int i = 0; i < 7; ++i) { o[i] = vj.anchor[row * 7 + i]; o[7 + i] = vj.observer[row * 7 + i]; o[14 + i] = vj.rig[row * 7 + i]; }
            o[21] = vj.td[row]; o[22] = vj.rho[row];
        }
    }
    if (J) {
        double pa[42], pb[42], pr[42];
        plusJacobian(&w.pose[7 * a], pa); plusJacobian(&w.pose[7 * b], pb); plusJacobian(w.rig, pr);
        for (int row = 0; row < 2; ++row) {
            double* o = J + row * 20;
            for (int c = 0; c < 6; ++c) {
                double sa = 0, sb = 0, sr = 0;
                for (int k = 0; k < 7; ++k) {
                    sa += vj.anchor[row * 7 + k] * pa[k * 6 + c];
                    sb += (a == b ? 0.0 : vj.observer[row * 7 + k]) * pb[k * 6 + c];
                    sr += vj.rig[row * 7 + k] * pr[k * 6 + c];
                }
                o[c] = sa; o[6 + c] = sb; o[12 + c] = sr;
            }
            o[18] = vj.td[row]; o[19] = vj.rho[row];
        }
    }
    return 1;
}

// Residual only of the numeric reference path (WindowVisualFactor.cpp:411-430): uses a
// FromTwoVectors transport and plain log map; used to pin the analytic residual.
int oracle_visual_residual_reference(const vib_window* win, int f, double* out) {
    Window w = load(*win);
    const int l = w.obsLm[f], a = w.lmFrame[l], b = w.obsFrame[f];
    const double rho = w.rho[l], td = w.td;
    if (!std::isfinite(rho) || rho <= 0) return 0;
    V3 dummy;
    const V3 ba = timeBearing(V3(&w.lmBearing[3 * l]), V3(&w.lmVel[3 * l]), w.lmTref[l] - td, false, dummy);
    const V3 bb = timeBearing(V3(&w.obsBearing[3 * f]), V3(&w.obsVel[3 * f]), w.obsTref[f] - td, false, dummy);
    auto camPose = [&](int slot, int cam, M3& R, V3& t) {
        const M3 Rb = Q::fromXyzw(&w.pose[7 * slot + 3]).normalized().matrix();
        const M3 Re = Q::fromXyzw(w.rig + 3).normalized().matrix();
        const M3 Rc = fromPtr(&w.camR[9 * cam]);
        R = Rb * Re * Rc;
        t = Rb * (Re * V3(&w.camT[3 * cam]) + V3(w.rig)) + V3(&w.pose[7 * slot]);
    };
    M3 Ra, Rb; V3 ta, tb;
    camPose(a, w.lmCam[l], Ra, ta); camPose(b, w.obsCam[f], Rb, tb);
    const V3 pW = Ra * (ba / rho) + ta;
    const V3 pC = transpose(Rb) * (pW - tb);
    if (!finite(pC) || norm(pC) < 1e-10) return 0;
    const V3 pred = normalized(pC);
    const double c = std::fmin(1.0, std::fmax(-1.0, dot(bb, pred))), s = norm(cross(bb, pred));
    const V3 meas(&w.obsBearing[3 * f]);
    const V3 lg = s < 1e-8 ? pred - c * bb : (std::atan2(s, c) / s) * (pred - c * bb);
    // Quaternion FromTwoVectors(meas, bb) rotation (Eigen: shortest arc)
    const V3 v0 = normalized(meas), v1 = normalized(bb);
    const double cth = dot(v0, v1);
    const V3 axis = cross(v0, v1);
    const double s2 = std::sqrt((1 + cth) * 2);
    const Q qt(s2 * 0.5, axis.x / s2, axis.y / s2, axis.z / s2);
    const M3 transport = qt.matrix();
    const M<3, 2> B = sphereBasis(meas);
    const M<3, 2> TB = transport * B;
    M<2, 2> S; S(0, 0) = w.obsSqrt[3 * f]; S(1, 0) = w.obsSqrt[3 * f + 1]; S(1, 1) = w.obsSqrt[3 * f + 2];
    const M23 P = S * transpose(TB);
    mul23(P, lg, out);
    return 1;
}

// IMU factor residual (15) and tangent Jacobians (4 blocks of 15 x 9; pose blocks use 6 cols).
int oracle_imu_factor(const vib_window* win, int e, double* r, double* J) {
    Window w = load(*win);
    const Frozen imu = unpackImu(&w.imu[(size_t)VIB_IMU_STRIDE * e]);
    const int s0 = w.imuFrom[e], s1 = w.imuTo[e];
    double Jl[4][15 * 9];
    if (!imuEvaluate(imu, &w.pose[7 * s0], &w.sb[9 * s0], &w.pose[7 * s1], &w.sb[9 * s1], r, J ? Jl : nullptr)) return 0;
    if (J) for (int k = 0; k < 4; ++k) std::copy_n(Jl[k], 135, J + 135 * k);
    return 1;
}

// Normal equations at the window state: H (nr x nr full), cross (ne x nr), diag (ne),
// g (nr + ne), cost. Returns nr, or -1 if invalid. Pass null to query sizes.
int oracle_linearize(const vib_window* win, double* H, double* crossOut, double* diag, double* g, double* cost,
                     int* ne_out) {
    Window w = load(*win);
    const Layout lay = layoutOf(w);
    if (ne_out) *ne_out = lay.ne;
    if (!H) return lay.nr;
    const Linear L = linearize(w, lay, stateOf(w));
    if (!L.valid) return -1;
    std::copy(L.H.begin(), L.H.end(), H);
    if (crossOut) std::copy(L.cross.begin(), L.cross.end(), crossOut);
    if (diag) std::copy(L.diag.begin(), L.diag.end(), diag);
    std::copy(L.g.begin(), L.g.end(), g);
    *cost = L.cost;
    return lay.nr;
}

// One solveOnline step at the window state: delta (nr + ne), damped reduced S (nr x nr),
// rhs (nr), metrics [cost, predicted, trial cost, max increment, trial valid].
int oracle_solve_step(const vib_window* win, double lambda, double* delta, double* S, double* rhs,
                      double* metrics) {
    Window w = load(*win);
    const Layout lay = layoutOf(w);
    const Bounds abs = boundsOf(w, lay);
    const State x = stateOf(w);
    const Linear L = linearize(w, lay, x);
    if (!L.valid) return VIB_ERR_INVALID_RESIDUAL;
    Step st = solveStep(w, lay, L, x, abs, lambda);
    if (!st.ok) return VIB_ERR_NOT_POSITIVE_DEFINITE;
    std::copy(st.delta.begin(), st.delta.end(), delta);
    if (S) std::copy(st.S.begin(), st.S.end(), S);
    if (rhs) std::copy(st.rhs.begin(), st.rhs.end(), rhs);
    metrics[0] = L.cost; metrics[1] = st.predicted; metrics[2] = st.trialCost;
    metrics[3] = st.maxIncrement; metrics[4] = st.trialValid ? 1 : 0;
    return VIB_OK;
}

// The same step as oracle_solve_step, block-sparse on `threads` host threads; delta (nr + ne),
// metrics [cost, predicted, trial cost, max increment, trial valid].
int oracle_solve_step_threaded(const vib_window* win, double lambda, int threads, double* delta, double* metrics) {
    const Prepared P = prepare(*win);
    std::vector<double> d;
    const int st = blockStep(P, stateOf(P.w), lambda, threads, d, metrics);
    if (st == VIB_OK) std::copy(d.begin(), d.end(), delta);
    return st;
}

// Timed CPU baseline: `steps` block-sparse LM steps of one window on `threads` host threads
// (topology prepared once, as a resident window), wall milliseconds of each step in ms[].
int oracle_time_steps(const vib_window* win, double lambda, int threads, int steps, double* ms) {
    const Prepared P = prepare(*win);
    const State x = stateOf(P.w);
    std::vector<double> d;
    double m[5];
    for (int s = 0; s < steps; ++s) {
        const auto t0 = std::chrono::steady_clock::now();
        const int st = blockStep(P, x, lambda, threads, d, m);
        ms[s] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (st != VIB_OK) return st;
    }
    return VIB_OK;
}

// Timed CPU baseline over many windows: each of `threads` host threads takes whole windows
// (window i goes to thread i % threads) and runs one single-threaded block-sparse LM step of
// each, `rounds` times; returns the wall seconds of all rounds.
int oracle_time_windows(const vib_window* wins, int count, double lambda, int threads, int rounds, double* seconds) {
    std::vector<Prepared> P;
    P.reserve(count);
    for (int i = 0; i < count; ++i) P.push_back(prepare(wins[i]));
    std::vector<int> status(threads, VIB_OK);
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&, t] {
            std::vector<double> d;
            double m[5];
            for (int r = 0; r < rounds; ++r)
                for (int i = t; i < count; i += threads) {
                    const int st = blockStep(P[i], stateOf(P[i].w), lambda, 1, d, m);
                    if (st != VIB_OK) status[t] = st;
                }
        });
    for (auto& th : pool) th.join();
    *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (int st : status) if (st != VIB_OK) return st;
    return VIB_OK;
}

int oracle_cost(const vib_window* win, double* cost) {
    Window w = load(*win);
    const Layout lay = layoutOf(w);
    return costAt(w, lay, stateOf(w), *cost) ? VIB_OK : VIB_ERR_INVALID_RESIDUAL;
}

// Full native LM solve; writes the final state into the window's in/out arrays.
// Two-frame solve; nis = {old before, old after, new before, new after}. Writes the poses and
// velocities of the two newest slots back into win on success.
int oracle_solve_latest_two(vib_window* win, const vib_solve_options* opt, int gate, double* nis, int* support,
                            vib_solve_report* rep) {
    Window w = load(*win);
    int sup = 0;
    const int st = latestTwo(w, *opt, gate, nis, sup, *rep);
    if (support) *support = sup;
    if (st == VIB_OK) {
        std::copy(w.pose.begin(), w.pose.end(), win->pose);
        std::copy(w.sb.begin(), w.sb.end(), win->speed_bias);
    }
    return st;
}

int oracle_solve(vib_window* win, const vib_solve_options* opt, vib_solve_report* rep) {
    Window w = load(*win);
    State s;
    const int st = lmSolve(w, *opt, *rep, s);
    if (st == VIB_OK) writeState(s, w, win);
    return st;
}

// OLD marginalization. Outputs the retained Schur system (H n x n, b n) and the prior
// (A rank x n, r0 rank, keys, refs). Sizes via sizes[0..4] = {n, rank, blocks, refLen, landmarks}.
int oracle_marginalize_old(const vib_window* win, int* sizes, double* H, double* b, double* A, double* r0,
                           int* keyKind, int64_t* keyId, double* ref, double* cutoff) {
    Window w = load(*win);
    MargResult m = marginalizeOld(w);
    if (m.status != VIB_OK) return m.status;
    sizes[0] = m.nr; sizes[1] = m.rank; sizes[2] = (int)m.keyKind.size(); sizes[3] = (int)m.ref.size();
    sizes[4] = m.landmarks;
    if (H) std::copy(m.Hprior.begin(), m.Hprior.end(), H);
    if (b) std::copy(m.bprior.begin(), m.bprior.end(), b);
    if (A) std::copy(m.A.begin(), m.A.end(), A);
    if (r0) std::copy(m.r0.begin(), m.r0.end(), r0);
    if (keyKind) std::copy(m.keyKind.begin(), m.keyKind.end(), keyKind);
    if (keyId) std::copy(m.keyId.begin(), m.keyId.end(), keyId);
    if (ref) std::copy(m.ref.begin(), m.ref.end(), ref);
    if (cutoff) *cutoff = m.cutoff;
    return VIB_OK;
}

// The whole linear system of the consumed factors before any elimination (size in *n; H, b,
// mask may be null for a size query); the retained columns of mask == 0 are in the order of the
// prior keys.
int oracle_marginal_system(const vib_window* win, int* n, double* H, double* b, int* mask) {
    Window w = load(*win);
    MargResult m = marginalizeOld(w);
    if (m.status != VIB_OK) return m.status;
    *n = (int)m.bfull.size();
    if (H) std::copy(m.Hfull.begin(), m.Hfull.end(), H);
    if (b) std::copy(m.bfull.begin(), m.bfull.end(), b);
    if (mask) std::copy(m.elimMask.begin(), m.elimMask.end(), mask);
    return VIB_OK;
}

// SECOND_NEW marginalization. sizes (in: capacities {L, F, E, prior cells, prior rows, prior
// blocks}; out: {K, L, F, E, prior rows, prior cols, prior blocks, ref length}), report fields
// {status, prior_reused, imu_mode, reset, demoted, retired, dropped, rank}; `out` receives the
// window after the slide when every capacity suffices.
int oracle_marginalize_second_new(const vib_window* win, const double* samples, int count, double t0, double t1,
                                  const vib_imu_params* params, int* sizes, int* report, vib_window* out) {
    const Window w = load(*win);
    SecondNewResult r = secondNew(w, samples, count, t0, t1, params);
    report[0] = r.status; report[1] = r.priorReused; report[2] = r.imuMode; report[3] = r.reset;
    report[4] = r.demoted; report[5] = r.retired; report[6] = r.dropped; report[7] = r.marg.rank;
    if (r.status != VIB_OK) return r.status;
    const Window& n = r.next;
    int cols = 0;
    for (int k : n.priorKind) cols += k == VIB_KEY_POSE || k == VIB_KEY_RIG ? 6 : (k == VIB_KEY_SPEED_BIAS ? 9 : 1);
    const int cap[6] = {sizes[0], sizes[1], sizes[2], sizes[3], sizes[4], sizes[5]};
    sizes[0] = n.K; sizes[1] = n.L; sizes[2] = n.F; sizes[3] = n.E; sizes[4] = n.priorRows; sizes[5] = cols;
    sizes[6] = n.priorBlocks; sizes[7] = (int)n.priorRef.size();
    if (out && n.L <= cap[0] && n.F <= cap[1] && n.E <= cap[2] && (size_t)n.priorRows * cols <= (size_t)cap[3] &&
        n.priorRows <= cap[4] && n.priorBlocks <= cap[5])
        storeWindow(n, out);
    return VIB_OK;
}

// The whole linear system of the factors a SECOND_NEW slide consumes (as oracle_marginal_system;
// *n = 0 when the prior is reused).
int oracle_second_new_system(const vib_window* win, const double* samples, int count, double t0, double t1,
                             const vib_imu_params* params, int* n, double* H, double* b, int* mask) {
    const Window w = load(*win);
    SecondNewResult r = secondNew(w, samples, count, t0, t1, params);
    if (r.status != VIB_OK) return r.status;
    *n = r.priorReused ? 0 : (int)r.marg.bfull.size();
    if (r.priorReused) return VIB_OK;
    if (H) std::copy(r.marg.Hfull.begin(), r.marg.Hfull.end(), H);
    if (b) std::copy(r.marg.bfull.begin(), r.marg.bfull.end(), b);
    if (mask) std::copy(r.marg.elimMask.begin(), r.marg.elimMask.end(), mask);
    return VIB_OK;
}

int oracle_preintegrate(const double* samples, int count, double t0, double t1, const double* sb,
                        const vib_imu_params* p, const double* g0, const double* g1, double* out) {
    return preintegrate(samples, count, t0, t1, sb, *p, g0, g1, out);
}

// The same preintegration, also returning the 15x15 covariance P (row-major, state order
// [p, alpha, v, b_g, b_a]) that the whitening W is computed from.
int oracle_preintegrate_cov(const double* samples, int count, double t0, double t1, const double* sb,
                            const vib_imu_params* p, const double* g0, const double* g1, double* out, double* cov) {
    return preintegrate(samples, count, t0, t1, sb, *p, g0, g1, out, cov);
}

// FrameStore::observation for an undistorted pinhole (FrameStore.cpp:102-124,
// PinholeCamera.hpp:778-807).
int oracle_pinhole_observation(double fx, double fy, double cx, double cy, double u, double v, double sigma,
                               double* bearing, double* sqrtInfo) {
    const V3 ray((u - cx) / fx, (v - cy) / fy, 1.0);
    M<3, 2> J; J(0, 0) = 1 / fx; J(1, 1) = 1 / fy;
    const V3 b = normalized(ray);
    const double rn = norm(ray);
    const M<3, 2> D = (1.0 / rn) * ((M3::identity() - outer(b, b)) * J);
    const M<2, 3> B = transpose(sphereBasis(b));
    const M<2, 2> BD = B * D;
    const M<2, 2> cov = (sigma * sigma) * (BD * transpose(BD));
    const double l00 = std::sqrt(cov(0, 0));
    const double l10 = cov(1, 0) / l00;
    const double l11 = std::sqrt(cov(1, 1) - l10 * l10);
    if (!(l00 > 0) || !(l11 > 0)) return VIB_ERR_INVALID_ARGUMENT;
    b.store(bearing);
    sqrtInfo[0] = 1 / l00; sqrtInfo[1] = -l10 / (l00 * l11); sqrtInfo[2] = 1 / l11;
    return VIB_OK;
}

int oracle_eigen_sym(const double* A, int n, double* evals, double* V) {
    std::vector<double> a(A, A + (size_t)n * n), ev, U;
    if (!jacobiEigen(a, n, ev, U)) return VIB_ERR_EIGEN_FAILED;
    std::copy(ev.begin(), ev.end(), evals);
    std::copy(U.begin(), U.end(), V);
    return VIB_OK;
}

void oracle_pose_plus(const double* x, const double* d, double* y) { posePlus(x, d, y, false); }
void oracle_pose_minus(const double* x, const double* ref, double* d) { poseMinus(x, ref, d); }

}  // extern "C"
