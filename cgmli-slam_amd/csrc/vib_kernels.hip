// HIP kernels of the sliding-window Schur backend for gfx950 (CDNA4), FP64 path.
//
// One LM iteration over a batch of windows is the launch sequence
//   k_lin_visual   (chunks)  : S^2 factors -> visual normal block partials, per-factor
//                              J_a^T J_rho, per-landmark (d, g_rho, c_rig, c_td), cost
//   k_lin_retained (windows) : IMU + prior rows -> dense retained normal matrix and cost
//   k_reduce       (slices)  : fixed-order sum of the chunk partials
//   k_schur        (chunks)  : -c c^T / (d + lambda max(d,1e-6)) partials
//   k_assemble     (windows) : damped reduced camera system + rhs
//   k_cholesky     (windows) : FP64 Cholesky, triangular solves, refinement, state trial
//   k_backsub      (chunks)  : inverse-range back-substitution, model decrease, rho trial
//   k_trial_cost   (chunks)  : robust cost at the trial state
//   k_lm_update    (windows) : retained trial cost, gain, lambda schedule, accept
// mirroring VisualCudaBatch::linearizeSchur / solveOnline / launchTrial
// (src/backend/VisualCudaBatch.cu:433-604,715-721) and the host LM of
// src/backend/BackendOptimizer.cpp:167-307, with no atomics to global memory.
#include <hip/hip_runtime.h>

#include <cstdlib>

#include <type_traits>

#include "../../include/vib.h"
#include "vib_device.hpp"
#include "vib_math.hpp"

namespace vib {

// ---------------------------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ int tri_index(int i, int j, int nv) {  // i <= j, packed upper rows
    return i * nv - (i * (i - 1)) / 2 + (j - i);
}

typedef double d4 __attribute__((ext_vector_type(4)));  // v_mfma_f64_16x16x4f64 accumulator
typedef float f4 __attribute__((ext_vector_type(4)));

// Deterministic block reductions (fixed tree order).
// Pairwise tree over the workgroup (level s: element t < s takes t + s, s = n/2 .. 1). The
// levels s >= 64 go through LDS with a barrier each; the last six run in wave 0, lane t taking
// lane t + s's value by a shuffle, which is the same operation on the same operands.
template <bool MAX>
__device__ __forceinline__ double block_reduce(double v, double* red) {
    const int t = threadIdx.x, n = blockDim.x;
    red[t] = v;
    __syncthreads();
    int s = n / 2;
    for (; s >= 64; s >>= 1) {
        if (t < s) red[t] = MAX ? fmax(red[t], red[t + s]) : red[t] + red[t + s];
        __syncthreads();
    }
    if (t < 64) {
        double x = red[t];
        for (; s > 0; s >>= 1) {
            const double y = __shfl_down(x, s, 64);
            if (t < s) x = MAX ? fmax(x, y) : x + y;
        }
        if (t == 0) red[0] = x;
    }
    __syncthreads();
    const double r = red[0];
    __syncthreads();
    return r;
}
__device__ double block_sum(double v, double* red) { return block_reduce<false>(v, red); }
__device__ double block_max(double v, double* red) { return block_reduce<true>(v, red); }
// N values through the same tree, sharing its barriers: value q is a maximum if bit q of MAXMASK
// is set, else a sum; red holds N * blockDim.x doubles.
template <int N, unsigned MAXMASK>
__device__ __forceinline__ void block_reduce_n(double (&v)[N], double* red) {
    const int t = threadIdx.x, n = blockDim.x;
#pragma unroll
    for (int q = 0; q < N; ++q) red[q * n + t] = v[q];
    __syncthreads();
    int s = n / 2;
    for (; s >= 64; s >>= 1) {
        if (t < s)
#pragma unroll
            for (int q = 0; q < N; ++q) {
                double* r = red + q * n;
                r[t] = ((MAXMASK >> q) & 1u) ? fmax(r[t], r[t + s]) : r[t] + r[t + s];
            }
        __syncthreads();
    }
    if (t < 64) {
        double x[N];
#pragma unroll
        for (int q = 0; q < N; ++q) x[q] = red[q * n + t];
        for (; s > 0; s >>= 1)
#pragma unroll
            for (int q = 0; q < N; ++q) {
                const double y = __shfl_down(x[q], s, 64);
                if (t < s) x[q] = ((MAXMASK >> q) & 1u) ? fmax(x[q], y) : x[q] + y;
            }
        if (t == 0)
#pragma unroll
            for (int q = 0; q < N; ++q) red[q * n] = x[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < N; ++q) v[q] = red[q * n];
    __syncthreads();
}

// Orders LDS traffic between the lanes of one wavefront (single-wave phases of a workgroup).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ double wave_reduce_sum(double v) {
    for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// Sum over the kChunkLanes lanes of a landmark group (xor tree, deterministic).
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
    for (int m = kChunkLanes / 2; m > 0; m >>= 1) v += __shfl_xor(v, m, kChunkLanes);
    return v;
}

// ---------------------------------------------------------------------------------------------
// Frame/camera table (LDS): world rotation/translation of every (frame, camera).
// ---------------------------------------------------------------------------------------------
struct FrameTable {
    double Rb[kMaxFrames][9];
    double pb[kMaxFrames][3];
    double Rc[kMaxFrames][kMaxCameras][9];
    double tc[kMaxFrames][kMaxCameras][3];
    int valid[kMaxFrames];  // finite and |q|^2 > 1e-16
    int unit[kMaxFrames];   // ||q|^2 - 1| <= 1e-6 (required for a pose Jacobian)
    int pcol[kMaxFrames];   // pose column of every frame (-1: constant), read per factor
    int rig_valid, rig_unit;
};

__device__ void build_frame_table(const WinDev& w, const double* pose, const double* rig, FrameTable& ft) {
    const int t = threadIdx.x;
    const double* qe = rig + 3;
    if (t == 0) {
        bool ok = true;
        for (int i = 0; i < 7; ++i) ok = ok && isfinite(rig[i]);
        const double n2 = qe[0] * qe[0] + qe[1] * qe[1] + qe[2] * qe[2] + qe[3] * qe[3];
        ft.rig_valid = ok && n2 > 1e-16;
        ft.rig_unit = isfinite(n2) && fabs(n2 - 1) <= 1e-6;
    }
    if (t < kMaxFrames) ft.pcol[t] = t < w.K ? w.pose_col[t] : -1;
    if (t < w.K) {
        const double* x = pose + 7 * t;
        bool ok = true;
        for (int i = 0; i < 7; ++i) ok = ok && isfinite(x[i]);
        const double n2 = x[3] * x[3] + x[4] * x[4] + x[5] * x[5] + x[6] * x[6];
        ft.valid[t] = ok && n2 > 1e-16;
        ft.unit[t] = isfinite(n2) && fabs(n2 - 1) <= 1e-6;
        const M3<double> R = quat_to_matrix(x + 3);
        for (int i = 0; i < 9; ++i) ft.Rb[t][i] = R.m[i];
        ft.pb[t][0] = x[0]; ft.pb[t][1] = x[1]; ft.pb[t][2] = x[2];
    }
    if (t < w.K * w.C) {
        const int k = t / w.C, c = t % w.C;
        const double* x = pose + 7 * k;
        const M3<double> Rb = quat_to_matrix(x + 3);
        const M3<double> Re = quat_to_matrix(qe);
        M3<double> Rcam;
        for (int i = 0; i < 9; ++i) Rcam.m[i] = w.cam[12 * c + i];
        const V3<double> tcam = load3(w.cam + 12 * c + 9);
        const M3<double> Rwc = mul(mul(Rb, Re), Rcam);
        const V3<double> twc = mul(Rb, mul(Re, tcam) + load3(rig)) + load3(x);
        for (int i = 0; i < 9; ++i) ft.Rc[k][c][i] = Rwc.m[i];
        ft.tc[k][c][0] = twc.x; ft.tc[k][c][1] = twc.y; ft.tc[k][c][2] = twc.z;
    }
}

// ---------------------------------------------------------------------------------------------
// One S^2 visual factor (VisualSphereBatchMath.hpp:67-174) in tangent coordinates.
// ---------------------------------------------------------------------------------------------
struct LmData {
    V3<double> raw, vel;
    double tref, rho;
    int frame, cam;
};
struct FactorJ {
    double r[2];
    double Ja[12];   // 2 x 6 anchor tangent (observer = -Ja)
    double Jr[12];   // 2 x 6 rig tangent
    double Jt[2];
    double Jrho[2];
    double w;        // Huber weight
};

// One observation record (planes of stride F): bearing, bearing velocity, reference time, the
// square-root information (lower 2x2).
struct ObsData {
    V3<double> raw, vel;
    double tref, s00, s10, s11;
};
__device__ __forceinline__ ObsData load_obs(const double* ob, size_t F, int f) {
    ObsData o;
    o.raw = {ob[OB_BX * F + f], ob[OB_BY * F + f], ob[OB_BZ * F + f]};
    o.vel = {ob[OB_VX * F + f], ob[OB_VY * F + f], ob[OB_VZ * F + f]};
    o.tref = ob[OB_T * F + f];
    o.s00 = ob[OB_S00 * F + f]; o.s10 = ob[OB_S10 * F + f]; o.s11 = ob[OB_S11 * F + f];
    return o;
}

// One visual factor: observation `obs`, observer frame/camera fc.
template <bool WANT_J>
__device__ bool eval_factor(const WinDev& w, const FrameTable& ft, const LmData& lm, const ObsData& obs, int fc,
                            double td, bool want_td, FactorJ& o) {
    const int b = fc >> 8, cb = fc & 255;
    const int a = lm.frame, ca = lm.cam;
    const double rho = lm.rho;
    if (!(rho > 0) || !isfinite(rho) || !isfinite(td)) return false;
    if (!ft.valid[a] || !ft.valid[b] || !ft.rig_valid) return false;
    const bool same = a == b;
    const V3<double> braw = obs.raw, bvel = obs.vel;
    const double btref = obs.tref;
    const double s00 = obs.s00, s10 = obs.s10, s11 = obs.s11;
    const bool movingA = sqnorm(lm.vel) > 0, movingB = sqnorm(bvel) > 0;
    const bool tdd = WANT_J && want_td;
    V3<double> da = {0, 0, 0}, db = {0, 0, 0}, ba, bb;
    if (movingA) ba = time_bearing(lm.raw, lm.vel, lm.tref - td, tdd, da);
    else { const double in = inv_sqrt(sqnorm(lm.raw)); ba = {lm.raw.x * in, lm.raw.y * in, lm.raw.z * in}; }
    if (movingB) bb = time_bearing(braw, bvel, btref - td, tdd, db);
    else { const double in = inv_sqrt(sqnorm(braw)); bb = {braw.x * in, braw.y * in, braw.z * in}; }
    if (!isfinite(ba.x) || !isfinite(ba.y) || !isfinite(ba.z) || !isfinite(bb.x) || !isfinite(bb.y) || !isfinite(bb.z))
        return false;
    // camera chain
    M3<double> Rwa, Rwb;
    for (int i = 0; i < 9; ++i) { Rwa.m[i] = ft.Rc[a][ca][i]; Rwb.m[i] = ft.Rc[b][cb][i]; }
    const V3<double> twa = {ft.tc[a][ca][0], ft.tc[a][ca][1], ft.tc[a][ca][2]};
    const V3<double> twb = {ft.tc[b][cb][0], ft.tc[b][cb][1], ft.tc[b][cb][2]};
    const double irho = 1.0 / rho;
    const V3<double> pw = mul(Rwa, irho * ba) + twa;
    const V3<double> pc = mulT(Rwb, pw - twb);
    // measurement projection: P = S B^T (transported when the bearing moves)
    V3<double> c0, c1;
    sphere_basis(braw, c0, c1);
    double M[6] = {s00 * c0.x, s00 * c0.y, s00 * c0.z,
                   s10 * c0.x + s11 * c1.x, s10 * c0.y + s11 * c1.y, s10 * c0.z + s11 * c1.z};
    double P[6];
    double den = 2, iden = 0.5;
    double K[9];
    if (movingB) {
        den = 1 + dot(braw, bb);
        iden = 1.0 / den;
        const V3<double> k = cross(braw, bb);
        K[0] = 0; K[1] = -k.z; K[2] = k.y; K[3] = k.z; K[4] = 0; K[5] = -k.x; K[6] = -k.y; K[7] = k.x; K[8] = 0;
        double Tr[9];  // transport = I + K + K K / den
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                double kk = K[r * 3] * K[c] + K[r * 3 + 1] * K[3 + c] + K[r * 3 + 2] * K[6 + c];
                Tr[r * 3 + c] = (r == c ? 1.0 : 0.0) + K[r * 3 + c] + kk * iden;
            }
        for (int r = 0; r < 2; ++r)
            for (int c = 0; c < 3; ++c)  // P = M * Tr^T
                P[r * 3 + c] = M[r * 3] * Tr[c * 3] + M[r * 3 + 1] * Tr[c * 3 + 1] + M[r * 3 + 2] * Tr[c * 3 + 2];
    } else {
        for (int i = 0; i < 6; ++i) P[i] = M[i];
    }
    SphereEval<double> e;
    if (!sphere_residual(pc, bb, P, den, WANT_J, e)) return false;
    o.r[0] = e.r[0]; o.r[1] = e.r[1];
    const double sq = e.r[0] * e.r[0] + e.r[1] * e.r[1];
    o.w = huber_weight(w.huber, sq);
    if (!WANT_J) return true;
    const double wt = o.w;
    // D = reduce * Rwb^T (2 x 3)
    double D[6];
    for (int r = 0; r < 2; ++r)
        for (int k = 0; k < 3; ++k)
            D[r * 3 + k] = e.reduce[r * 3] * Rwb.m[k * 3] + e.reduce[r * 3 + 1] * Rwb.m[k * 3 + 1] +
                           e.reduce[r * 3 + 2] * Rwb.m[k * 3 + 2];
    // anchor pose: [D, -D [p_w]x] ; rows: [D_r, p_w x D_r]
    const int ac = ft.pcol[a], bc = ft.pcol[b];
    if (same) {
        for (int i = 0; i < 12; ++i) { o.Ja[i] = 0; o.Jr[i] = 0; }
    } else {
        if ((ac >= 0 && !ft.unit[a]) || (bc >= 0 && !ft.unit[b]) || (w.rig_col >= 0 && !ft.rig_unit)) return false;
        for (int r = 0; r < 2; ++r) {
            const V3<double> Dr = {D[r * 3], D[r * 3 + 1], D[r * 3 + 2]};
            const V3<double> rot = cross(pw, Dr);
            o.Ja[r * 6 + 0] = wt * Dr.x; o.Ja[r * 6 + 1] = wt * Dr.y; o.Ja[r * 6 + 2] = wt * Dr.z;
            o.Ja[r * 6 + 3] = wt * rot.x; o.Ja[r * 6 + 4] = wt * rot.y; o.Ja[r * 6 + 5] = wt * rot.z;
        }
        if (w.rig_col >= 0) {
            M3<double> Ra, Rb;
            for (int i = 0; i < 9; ++i) { Ra.m[i] = ft.Rb[a][i]; Rb.m[i] = ft.Rb[b][i]; }
            const V3<double> pa = {ft.pb[a][0], ft.pb[a][1], ft.pb[a][2]};
            const V3<double> pb = {ft.pb[b][0], ft.pb[b][1], ft.pb[b][2]};
            const V3<double> body = mulT(Ra, pw - pa);   // point in the anchor body frame
            const V3<double> q = mulT(Rb, pw - pb);      // point in the observer body frame
            for (int r = 0; r < 2; ++r) {
                const V3<double> Dr = {D[r * 3], D[r * 3 + 1], D[r * 3 + 2]};
                const V3<double> DRa = mulT(Ra, Dr), DRb = mulT(Rb, Dr);  // rows of D R_a, D R_b
                const V3<double> tr = DRa - DRb;
                const V3<double> rot = cross(DRb, q) - cross(DRa, body);
                o.Jr[r * 6 + 0] = wt * tr.x; o.Jr[r * 6 + 1] = wt * tr.y; o.Jr[r * 6 + 2] = wt * tr.z;
                o.Jr[r * 6 + 3] = wt * rot.x; o.Jr[r * 6 + 4] = wt * rot.y; o.Jr[r * 6 + 5] = wt * rot.z;
            }
        }
    }
    // rho: D Rwa (-ba / rho^2)
    {
        const V3<double> v = mul(Rwa, (-irho * irho) * ba);
        o.Jrho[0] = wt * (D[0] * v.x + D[1] * v.y + D[2] * v.z);
        o.Jrho[1] = wt * (D[3] * v.x + D[4] * v.y + D[5] * v.z);
        if (!isfinite(o.Jrho[0]) || !isfinite(o.Jrho[1])) return false;
    }
    o.Jt[0] = o.Jt[1] = 0;
    if (tdd && (movingA || movingB)) {
        // J = D Rwa (da/rho) + P Jb db + M dR^T log  (VisualSphereBatchMath.hpp:162-171)
        const V3<double> v = mul(Rwa, irho * da);
        double t1[2] = {D[0] * v.x + D[1] * v.y + D[2] * v.z, D[3] * v.x + D[4] * v.y + D[5] * v.z};
        // Jb db = f(-bb (n.db) - c db) + g u (n.db)
        const double ndb = dot(e.n, db);
        const V3<double> jbdb = e.f * ((-ndb) * bb - e.c * db) + (e.g * ndb) * e.u;
        double t2[2] = {P[0] * jbdb.x + P[1] * jbdb.y + P[2] * jbdb.z, P[3] * jbdb.x + P[4] * jbdb.y + P[5] * jbdb.z};
        double t3[2] = {0, 0};
        if (movingB) {
            const V3<double> kd = cross(braw, db);
            double dK[9] = {0, -kd.z, kd.y, kd.z, 0, -kd.x, -kd.y, kd.x, 0};
            const double mdb = dot(braw, db), mdd = mdb * (iden * iden);
            double dR[9];
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) {
                    double a1 = 0, a2 = 0, kk = 0;
                    for (int m = 0; m < 3; ++m) {
                        a1 += dK[r * 3 + m] * K[m * 3 + c];
                        a2 += K[r * 3 + m] * dK[m * 3 + c];
                        kk += K[r * 3 + m] * K[m * 3 + c];
                    }
                    dR[r * 3 + c] = dK[r * 3 + c] + (a1 + a2) * iden - kk * mdd;
                }
            // dR^T log
            const V3<double> lg = e.logv;
            const V3<double> dl = {dR[0] * lg.x + dR[3] * lg.y + dR[6] * lg.z, dR[1] * lg.x + dR[4] * lg.y + dR[7] * lg.z,
                                   dR[2] * lg.x + dR[5] * lg.y + dR[8] * lg.z};
            t3[0] = M[0] * dl.x + M[1] * dl.y + M[2] * dl.z;
            t3[1] = M[3] * dl.x + M[4] * dl.y + M[5] * dl.z;
        }
        o.Jt[0] = wt * (t1[0] + t2[0] + t3[0]);
        o.Jt[1] = wt * (t1[1] + t2[1] + t3[1]);
        if (!isfinite(o.Jt[0]) || !isfinite(o.Jt[1])) return false;
    }
    return true;
}

__device__ __forceinline__ LmData load_landmark(const WinDev& w, const double* rho, int l) {
    LmData d;
    const size_t L = (size_t)w.L;
    d.raw = {w.lmd[LM_BX * L + l], w.lmd[LM_BY * L + l], w.lmd[LM_BZ * L + l]};
    d.vel = {w.lmd[LM_VX * L + l], w.lmd[LM_VY * L + l], w.lmd[LM_VZ * L + l]};
    d.tref = w.lmd[LM_T * L + l];
    d.rho = rho[l];
    const int fc = w.lm_fc[l];
    d.frame = fc >> 8;
    d.cam = fc & 255;
    return d;
}

__device__ __forceinline__ bool window_active(const WinDev& w) { return w.lm->status == 0; }

constexpr int kAccSize = kMaxVisual * (kMaxVisual + 1) / 2 + kMaxVisual;
constexpr int kLinXcds = 8;  // XCDs of an MI355X: workgroup i of a launch runs on XCD i % 8

// ---------------------------------------------------------------------------------------------
// k_lin_visual: visual factors in (anchor frame, observer frame) pair order, one factor per lane
// (VisualCudaBatch.cu:46-136). Every run of one pair inside a wave reduces to the Gram matrix of
// its stacked weighted rows [J_anchor (6) | J_rig (6) | J_td | r] on v_mfma_f64_16x16x4f64
// (32 steps over the wave's 128 rows); the Gram matrix is scattered once into the packed visual
// block (anchor, observer, rig and td blocks follow from it with signs, J_observer = -J_anchor).
// Per-factor landmark terms go to fac_u / fac_lm in landmark order (k_lm_sums adds them up).
// ---------------------------------------------------------------------------------------------
constexpr int kGramCols = 16;
// Per-wave row staging, column-major: element (row, col) at col * kRowPitch + row. A lane writes
// its two rows of a column as one 16-byte store (consecutive lanes contiguous); the MFMA reads
// (16 columns x 4 rows per step) fall on distinct bank pairs since kRowPitch = 4 mod 32.
constexpr int kRowPitch = 132;
constexpr int kWaves = kBlock / 64;
constexpr int kGramDim = 14;  // [J_anchor (6) | J_rig (6) | J_td | r]
constexpr int kGramEntries = kGramDim * (kGramDim + 1) / 2;
constexpr int kGramsPerWave = 4;  // listed Gram matrices per wave and scatter pass
__device__ __forceinline__ int gram_entry(int i, int j) { return i * kGramDim - i * (i - 1) / 2 + (j - i); }

__device__ __forceinline__ void acc_add(double* acc, int nv, int ci, int cj, double v) {
    if (ci < 0 || cj < 0) return;
    acc[ci <= cj ? tri_index(ci, cj, nv) : tri_index(cj, ci, nv)] += v;
}

// Gram entry G[i][j] (i <= j) of pair (a, b) into the visual block and gradient. The elements
// one entry adds to are determined by the entry and the pair's blocks, so different entries of
// one Gram matrix never meet.
__device__ __forceinline__ void scatter_gram(double* acc, double* gacc, const WinDev& w, const FrameTable& ft, int a,
                                             int b, int i, int j, double g) {
    const int nv = w.nv;
    const bool same = a == b;
    const int ac = same ? -1 : ft.pcol[a], bc = same ? -1 : ft.pcol[b];
    const int rc = w.rig_col, tc = w.td_col;
    if (j < 6) {  // anchor/observer pose blocks
        acc_add(acc, nv, ac >= 0 ? ac + i : -1, ac >= 0 ? ac + j : -1, g);
        acc_add(acc, nv, bc >= 0 ? bc + i : -1, bc >= 0 ? bc + j : -1, g);
        if (ac >= 0 && bc >= 0) {
            acc_add(acc, nv, ac + i, bc + j, -g);
            if (i != j) acc_add(acc, nv, ac + j, bc + i, -g);
        }
    } else if (j < 12) {
        if (i < 6) {
            if (rc >= 0 && !same) {
                acc_add(acc, nv, ac >= 0 ? ac + i : -1, rc + j - 6, g);
                acc_add(acc, nv, bc >= 0 ? bc + i : -1, rc + j - 6, -g);
            }
        } else if (rc >= 0) {
            acc_add(acc, nv, rc + i - 6, rc + j - 6, g);
        }
    } else if (j == 12) {
        if (tc < 0) return;
        if (i < 6) {
            acc_add(acc, nv, ac >= 0 ? ac + i : -1, tc, g);
            acc_add(acc, nv, bc >= 0 ? bc + i : -1, tc, -g);
        } else if (i < 12) {
            if (rc >= 0 && !same) acc_add(acc, nv, rc + i - 6, tc, g);
        } else {
            acc_add(acc, nv, tc, tc, g);
        }
    } else if (j == 13) {  // gradient J^T r
        if (i < 6) {
            if (ac >= 0) gacc[ac + i] += g;
            if (bc >= 0) gacc[bc + i] -= g;
        } else if (i < 12) {
            if (rc >= 0 && !same) gacc[rc + i - 6] += g;
        } else if (i == 12) {
            if (tc >= 0) gacc[tc] += g;
        }
    }
}

// Per-factor coupling terms, stored at the factor's pair-order position p (k_lin_visual writes
// them contiguously; the landmark-ordered kernels read them through pf_pos). FP64: fac_u[k F + p]
// (6 planes), fac_lm[q F + p] (FacLm planes).
// Mixed precision: the same arrays hold each term as an FP32 pair (hi, lo) with
// hi = fl32(v), lo = fl32(v - hi), so hi + lo carries v to about 2^-48: u in float planes
// k (hi) and 6 + k (lo); the Hessian-side landmark terms d, c_rig (6) and c_td in float planes
// j (hi) and 8 + j (lo), j = fl32_plane(q); the landmark gradient g_rho in FP64 behind them
// (fac_lm + 8 F doubles). Both forms fit the FP64 allocation. The Schur product reads the hi
// planes only (FP32 rows, FP32 MFMA); the refinement operator and the back-substitution read
// hi + lo.
__device__ __forceinline__ int fl32_plane(int q) { return q == FL_D ? 0 : q - 1; }
constexpr int kFl32Planes = FL_PLANES - 1;  // FP32 Hessian-side planes of fac_lm
template <bool MIXED>
__device__ __forceinline__ double load_u(const WinDev& w, size_t F, int k, int p) {
    if constexpr (MIXED) return (double)((const float*)w.fac_u)[k * F + p];
    return w.fac_u[k * F + p];
}
template <bool MIXED>
__device__ __forceinline__ double load_u_exact(const WinDev& w, size_t F, int k, int p) {
    if constexpr (MIXED) {
        const float* u32 = (const float*)w.fac_u;
        return (double)u32[k * F + p] + (double)u32[(6 + k) * F + p];
    }
    return w.fac_u[k * F + p];
}
__device__ __forceinline__ void store_pair(float* planes, size_t F, int hi_plane, int lo_plane, size_t p, double v) {
    const float hi = (float)v;
    planes[hi_plane * F + p] = hi;
    planes[lo_plane * F + p] = (float)(v - (double)hi);
}

// IMU edges and prior of a window (k_retained below). Their linearization runs as extra
// workgroups of the visual linearization, so it overlaps the visual work in one launch. (The
// trial cost stays a launch of its own: inside k_trial_cost its 280 registers would halve that
// kernel's occupancy.)
// LDS doubles of one retained item's wave: the prior's chart and residual, or an IMU edge's
// whitening, whitened residual and raw Jacobian
constexpr int kRetainedScratch = 2 * kMaxRetained > 240 + 450 ? 2 * kMaxRetained : 240 + 450;
template <int MODE>
__device__ __forceinline__ void retained_body(const WinDev& w, int y, int lane, double* scratch);
// The retained items as extra workgroups of a chunk kernel: item i = (window i / per,
// edge or prior i % per), one per wave.
template <int MODE>
__device__ __forceinline__ void retained_items(const WinDev* wins, int first_block, int items, int per,
                                               double* scratch_of_wave) {
    const int item = ((int)blockIdx.x - first_block) * (kBlock / 64) + (int)threadIdx.x / 64;
    if (item < items) retained_body<MODE>(wins[item / per], item % per, threadIdx.x % 64, scratch_of_wave);
}

// Workgroups 0 .. rb-1 run the retained items (linearization, mode 0), items / per as above;
// they come first so that their longer evaluation overlaps the visual chunks instead of
// trailing them. rb is a multiple of the XCD count, so visual chunk j (workgroup rb + j) keeps
// the XCD j % 8 of the work-list order.
template <bool MIXED>
__global__ __launch_bounds__(kBlock) void k_lin_visual(const WinDev* wins, const Chunk* lin, int rb, int items, int per) {
    __shared__ double acc[kAccSize];
    __shared__ FrameTable ft;
    __shared__ double rows[kWaves][kGramCols * kRowPitch];  // per wave: 2 rows per lane x 16 columns
    __shared__ int more[kWaves];
    __shared__ double glist[kWaves][kGramsPerWave][kGramEntries];
    __shared__ int gkey[kWaves][kGramsPerWave];
    __shared__ int ngram[kWaves];
    __shared__ double red[2 * kBlock];
    if ((int)blockIdx.x < rb) {  // the row staging of the wave holds the prior's scratch
        static_assert(kGramCols * kRowPitch >= kRetainedScratch, "a wave's row staging holds a retained item");
        retained_items<0>(wins, 0, items, per, rows[threadIdx.x / 64]);
        return;
    }
    const Chunk ch = lin[blockIdx.x - rb];
    if (ch.local < 0) return;  // an empty position of the XCD-ordered work list
    const WinDev& w = wins[ch.win];
    if (!window_active(w) || !w.lm->new_point) return;
    const int nv = w.nv, tri = w.tri;
    for (int i = threadIdx.x; i < tri + nv; i += kBlock) acc[i] = 0;
    build_frame_table(w, w.pose, w.rig, ft);
    __syncthreads();
    double* gacc = acc + tri;
    const double td = *w.td;
    const bool want_td = w.td_col >= 0;
    const int t = threadIdx.x, lane = t % 64, wave = t / 64;
    double* V = rows[wave];
    const size_t F = (size_t)w.F;
    double cost = 0;
    int invalid = 0;
    const int nrounds = (ch.lm_end - ch.lm_begin + kBlock - 1) / kBlock;  // the same for every wave
    // scatter threads: Gram entry t = (gi, gj), summed over consecutive listed matrices of one
    // pair (ckey) before it is added to the accumulator
    int gi = 0, gj = t < kGramEntries ? t : 0, ckey = -1;
    double cval = 0;
    while (gj >= kGramDim - gi) {  // packed upper index -> (gi, gj); ends for every t < 105
        gj -= kGramDim - gi;
        ++gi;
    }
    gj += gi;
    // a factor's records (observation, anchor bearing, ids); each round's are loaded during the
    // previous round's Gram phase, so their global latency overlaps it
    struct LinIn {
        ObsData ob;
        V3<double> raw, vel;
        double tref;
        int ids, lmi;
    };
    auto load_in = [&](int q) {
        LinIn in{};
        if (q < ch.lm_end) {
            in.ob = load_obs(w.pf_ob, F, q);
            in.raw = {w.pf_an[LM_BX * F + q], w.pf_an[LM_BY * F + q], w.pf_an[LM_BZ * F + q]};
            in.vel = {w.pf_an[LM_VX * F + q], w.pf_an[LM_VY * F + q], w.pf_an[LM_VZ * F + q]};
            in.tref = w.pf_an[LM_T * F + q];
            in.ids = w.pf_ids[q];
            in.lmi = w.pf_lm[q];
        }
        return in;
    };
    LinIn next_in = load_in(ch.lm_begin + 64 * wave + lane);
    for (int round = 0; round < nrounds; ++round) {
        const int p = ch.lm_begin + round * kBlock + 64 * wave + lane;
        const LinIn cur = next_in;
        int key = -1;
#pragma unroll
        for (int c = 0; c < kGramCols; ++c) {
            V[c * kRowPitch + 2 * lane] = 0.0;
            V[c * kRowPitch + 2 * lane + 1] = 0.0;
        }
        if (p < ch.lm_end) {
            const int ids = cur.ids;
            LmData lm;
            lm.raw = cur.raw;
            lm.vel = cur.vel;
            lm.tref = cur.tref;
            lm.rho = w.rho[cur.lmi];
            lm.frame = ids >> 24;
            lm.cam = (ids >> 16) & 255;
            const int fc = ids & 0xffff;
            FactorJ J;
            bool ok = eval_factor<true>(w, ft, lm, cur.ob, fc, td, want_td, J);
            const double sq = ok ? J.r[0] * J.r[0] + J.r[1] * J.r[1] : 0.0;
            ok = ok && isfinite(sq);
            double u[6] = {0, 0, 0, 0, 0, 0}, fl[FL_PLANES] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            if (ok) {
                cost += huber_cost(w.huber, sq);
                const double r0 = J.w * J.r[0], r1 = J.w * J.r[1];
                const bool same = lm.frame == (fc >> 8);
                double* v0 = V + 2 * lane;  // rows 2 lane (v0[c kRowPitch]) and 2 lane + 1 (v0[... + 1])
#pragma unroll
                for (int k = 0; k < 6; ++k) {
                    v0[k * kRowPitch] = J.Ja[k]; v0[k * kRowPitch + 1] = J.Ja[6 + k];
                    v0[(6 + k) * kRowPitch] = J.Jr[k]; v0[(6 + k) * kRowPitch + 1] = J.Jr[6 + k];
                }
                v0[12 * kRowPitch] = J.Jt[0]; v0[12 * kRowPitch + 1] = J.Jt[1];
                v0[13 * kRowPitch] = r0; v0[13 * kRowPitch + 1] = r1;
                key = lm.frame * 32 + (fc >> 8);
                if (!same)
                    for (int k = 0; k < 6; ++k) u[k] = J.Ja[k] * J.Jrho[0] + J.Ja[6 + k] * J.Jrho[1];
                fl[FL_D] = J.Jrho[0] * J.Jrho[0] + J.Jrho[1] * J.Jrho[1];
                fl[FL_G] = J.Jrho[0] * r0 + J.Jrho[1] * r1;
                if (!same)
                    for (int k = 0; k < 6; ++k) fl[FL_CR0 + k] = J.Jr[k] * J.Jrho[0] + J.Jr[6 + k] * J.Jrho[1];
                fl[FL_CT] = J.Jt[0] * J.Jrho[0] + J.Jt[1] * J.Jrho[1];
            } else {
                invalid = 1;
            }
            if constexpr (MIXED) {
                float* u32 = (float*)w.fac_u;
                float* l32 = (float*)w.fac_lm;
                for (int k = 0; k < 6; ++k) store_pair(u32, F, k, 6 + k, p, u[k]);
                for (int q = 0; q < FL_PLANES; ++q)
                    if (q != FL_G) store_pair(l32, F, fl32_plane(q), kFl32Planes + fl32_plane(q), p, fl[q]);
                w.fac_lm[kFl32Planes * F + p] = fl[FL_G];
            } else {
                for (int k = 0; k < 6; ++k) w.fac_u[k * F + p] = u[k];
                for (int q = 0; q < FL_PLANES; ++q) w.fac_lm[q * F + p] = fl[q];
            }
        }
        wave_sync();
        if (round + 1 < nrounds) next_in = load_in(p + kBlock);
        // runs of one pair: Gram matrix on MFMA. A run occupies the lanes between its first and
        // last factor (pair order), so only the MFMA steps over its rows run; the rows of other
        // keys inside a boundary step are masked to zero. Every wave lists the packed upper
        // entries of up to kGramsPerWave Gram matrices in LDS, then the workgroup scatters the
        // lists, one thread per Gram entry and the lists in wave order: the entries of one
        // Gram matrix never share an accumulator element, so the scatter needs no atomics and
        // its sums do not depend on scheduling.
        // the wave's staged rows as MFMA operands: step st covers rows 4 st .. 4 st + 3 (the
        // factors of lanes 2 st and 2 st + 1); lane l supplies row 4 st + l / 16 (factor
        // 2 st + l / 32), column l % 16
        const int half = lane / 32;
        const double* vcol = V + (lane % 16) * kRowPitch + lane / 16;
        // passes of up to kGramsPerWave listed matrices per wave, until no wave has runs left (the
        // flags are read between the pass's two barriers, so every wave takes the same passes)
        unsigned long long pending = __ballot(key >= 0);
        for (;;) {
            int cnt = 0;
            while (pending && cnt < kGramsPerWave) {
                const int leader = __ffsll(pending) - 1;
                const int pk = __shfl(key, leader, 64);
                const unsigned long long run = __ballot(key == pk);
                const int last = 63 - __clzll(run);
                // groups of four steps; run membership of a step's rows comes from the
                // wave-uniform `run` mask, the exit is wave-uniform and the loads unconditional
                // (index clamped to the wave's 32 steps), so the chain has no divergent branch
                d4 G = {0.0, 0.0, 0.0, 0.0};
                const int s0 = leader / 2, s1 = last / 2;
#pragma unroll
                for (int g = 0; g < 8; ++g) {
                    const int st = s0 + 4 * g;
                    if (st > s1) break;
                    double x[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int sc = min(st + u, 31);
                        const double v = vcol[4 * sc];
                        const bool in = (st + u <= s1) & (((run >> (2 * sc + half)) & 1ull) != 0ull);
                        x[u] = in ? v : 0.0;
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) G = __builtin_amdgcn_mfma_f64_16x16x4f64(x[u], x[u], G, 0, 0, 0);
                }
                double* gl = glist[wave][cnt];
#pragma unroll
                for (int rr = 0; rr < 4; ++rr) {
                    const int i = 4 * rr + lane / 16, j = lane % 16;
                    if (i <= j && j < kGramDim) gl[gram_entry(i, j)] = G[rr];
                }
                if (lane == 0) gkey[wave][cnt] = pk;
                ++cnt;
                pending &= ~run;
            }
            if (lane == 0) {
                ngram[wave] = cnt;
                more[wave] = pending != 0ull;
            }
            __syncthreads();
            bool again = false;
#pragma unroll
            for (int q = 0; q < kWaves; ++q) again = again || more[q] != 0;
            if (t < kGramEntries) {
                // listed matrices in (wave, list position) order; the next one's key and entry
                // are read from LDS while the current one is merged
                static_assert(kWaves * kGramsPerWave <= 32, "one bit per listed matrix");
                unsigned listed = 0;
#pragma unroll
                for (int q = 0; q < kWaves; ++q) listed |= ((1u << ngram[q]) - 1u) << (kGramsPerWave * q);
                const int* keys = &gkey[0][0];
                const double* vals = &glist[0][0][0] + t;
                int s = listed ? __builtin_ctz(listed) : -1;
                int knext = s >= 0 ? keys[s] : -1;
                double vnext = s >= 0 ? vals[s * kGramEntries] : 0.0;
                while (s >= 0) {
                    const int k2 = knext;
                    const double val = vnext;
                    listed &= listed - 1u;
                    s = listed ? __builtin_ctz(listed) : -1;
                    if (s >= 0) {
                        knext = keys[s];
                        vnext = vals[s * kGramEntries];
                    }
                    if (k2 != ckey) {
                        if (ckey >= 0) scatter_gram(acc, gacc, w, ft, ckey / 32, ckey % 32, gi, gj, cval);
                        ckey = k2;
                        cval = 0;
                    }
                    cval += val;
                }
            }
            __syncthreads();
            if (!again) break;
        }
        wave_sync();
    }
    if (t < kGramEntries && ckey >= 0) scatter_gram(acc, gacc, w, ft, ckey / 32, ckey % 32, gi, gj, cval);
    __syncthreads();
    double* part = w.lin_part + (size_t)ch.local * (tri + nv);
    for (int i = threadIdx.x; i < tri + nv; i += kBlock) part[i] = acc[i];
    double sums[2] = {cost, invalid ? 1.0 : 0.0};
    block_reduce_n<2, 0b10u>(sums, red);  // the cost, then the invalid flag (maximum)
    const double c = sums[0], inv = sums[1];
    if (threadIdx.x == 0) {
        w.lin_sc[(size_t)ch.local * CS_PLANES + CS_COST] = c;
        w.lin_sc[(size_t)ch.local * CS_PLANES + CS_INVALID] = inv;
    }
}

// ---------------------------------------------------------------------------------------------
// k_lm_sums: per-landmark d, g_rho, c_rig, c_td from the per-factor terms (fixed order).
// ---------------------------------------------------------------------------------------------
template <bool MIXED>
__global__ __launch_bounds__(kBlock) void k_lm_sums(const WinDev* wins, const Chunk* chunks) {
    const Chunk ch = chunks[blockIdx.x];
    const WinDev& w = wins[ch.win];
    if (!window_active(w) || !w.lm->new_point) return;
    const size_t L = (size_t)w.L, F = (size_t)w.F;
    for (int l = ch.lm_begin + threadIdx.x; l < ch.lm_end; l += blockDim.x) {
        double sum[FL_PLANES] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int f = w.lm_off[l]; f < w.lm_off[l + 1]; ++f) {
            const int p = w.pf_pos[f];
            if constexpr (MIXED) {
                const float* l32 = (const float*)w.fac_lm;
                for (int q = 0; q < FL_PLANES; ++q)
                    sum[q] += q == FL_G ? w.fac_lm[kFl32Planes * F + p]
                                        : (double)l32[fl32_plane(q) * F + p] +
                                              (double)l32[(kFl32Planes + fl32_plane(q)) * F + p];
            } else {
                for (int q = 0; q < FL_PLANES; ++q) sum[q] += w.fac_lm[q * F + p];
            }
        }
        for (int q = 0; q < FL_PLANES; ++q) w.lm_out[q * L + l] = sum[q];
    }
}

// ---------------------------------------------------------------------------------------------
// Retained factors: IMU (GpuRetainedFactorMath.hpp:53-140) and the square-root prior
// (VisualCudaBatch.cu:216-268, MarginalizationPrior.cpp:20-66).
// ---------------------------------------------------------------------------------------------
struct ImuSensor {
    V3<double> p, v, arm, lever;
    Qt<double> q;
};
__device__ ImuSensor imu_sensor(const double* d, const double* pose, const double* sb, int end) {
    const Qt<double> q = qnormalized(qload(pose + 3));
    ImuSensor s;
    s.q = qnormalized(qmul(q, qload(d + 5)));
    s.arm = qrotate(q, load3(d + 2));
    s.lever = cross(qrotate(s.q, load3(d + 9 + 3 * end) - load3(sb + 3)), s.arm);
    s.p = load3(pose) + s.arm;
    s.v = load3(sb) + s.lever;
    return s;
}
__device__ __forceinline__ V3<double> m3v(const double* M, V3<double> v) {  // row-major 3x3 * v
    return {M[0] * v.x + M[1] * v.y + M[2] * v.z, M[3] * v.x + M[4] * v.y + M[5] * v.z,
            M[6] * v.x + M[7] * v.y + M[8] * v.z};
}
__device__ __forceinline__ void skew_of(V3<double> v, double* S) {
    S[0] = 0; S[1] = -v.z; S[2] = v.y; S[3] = v.z; S[4] = 0; S[5] = -v.x; S[6] = -v.y; S[7] = v.x; S[8] = 0;
}
__device__ __forceinline__ void mm3(const double* A, const double* B, double* C) {
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) C[r * 3 + c] = A[r * 3] * B[c] + A[r * 3 + 1] * B[3 + c] + A[r * 3 + 2] * B[6 + c];
}

// Quantities of one IMU edge shared by all Jacobian columns (GpuRetainedFactorMath.hpp:71-101).
struct ImuShared {
    ImuSensor s0, s1;
    M3<double> R;
    double dt;
    V3<double> dp, dv;
    double Jrot[9], Jbias[9];
    double e[15];  // unwhitened residual
};

__device__ void imu_shared(const double* d, const double* pose0, const double* sb0, const double* pose1,
                           const double* sb1, bool jac, ImuShared& o) {
    o.s0 = imu_sensor(d, pose0, sb0, 0);
    o.s1 = imu_sensor(d, pose1, sb1, 1);
    o.R = qmatrix(qconj(o.s0.q));
    o.dt = d[0];
    const V3<double> grav = {0, 0, d[1]};
    const V3<double> bg = load3(sb0 + 3) - load3(d + 25), ba = load3(sb0 + 6) - load3(d + 28);
    o.dp = o.s0.p - o.s1.p + o.dt * o.s0.v - (0.5 * o.dt * o.dt) * grav;
    o.dv = o.s0.v - o.s1.v - o.dt * grav;
    const V3<double> phi = -1.0 * m3v(d + 31, bg);
    const Qt<double> corr = rotation_exp(phi);
    const Qt<double> Dq = qload(d + 15);
    const Qt<double> dq = qmul(corr, Dq);
    const Qt<double> rel = qmul(qconj(o.s1.q), o.s0.q);
    double* e = o.e;
    const V3<double> ep = mul(o.R, o.dp) + load3(d + 22) + m3v(d + 49, bg) - m3v(d + 67, ba);
    const V3<double> eq = 2.0 * qvec(qmul(dq, rel));
    const V3<double> ev = mul(o.R, o.dv) + load3(d + 19) + m3v(d + 40, bg) - m3v(d + 58, ba);
    e[0] = ep.x; e[1] = ep.y; e[2] = ep.z; e[3] = eq.x; e[4] = eq.y; e[5] = eq.z; e[6] = ev.x; e[7] = ev.y; e[8] = ev.z;
    for (int i = 0; i < 6; ++i) e[9 + i] = sb0[3 + i] - sb1[3 + i];
    if (!jac) return;
    double GA[9];
    {
        const V3<double> v = phi;  // rotationJacobian(-phi) * A, negated
        const V3<double> mv = -1.0 * v;
        const double a2 = sqnorm(mv), n = sqrt(a2);
        const double c1 = a2 < 1e-10 ? 0.5 - a2 / 24 : (1 - cos(n)) / a2;
        const double c2 = a2 < 1e-10 ? 1.0 / 6 - a2 / 120 : (n - sin(n)) / (a2 * n);
        double Sk[9], S2[9], Jl[9];
        skew_of(mv, Sk);
        mm3(Sk, Sk, S2);
        for (int i = 0; i < 9; ++i) Jl[i] = (i % 4 == 0 ? 1.0 : 0.0) + c1 * Sk[i] + c2 * S2[i];
        mm3(Jl, d + 31, GA);
        for (int i = 0; i < 9; ++i) GA[i] = -GA[i];
    }
    const Qt<double> tail = qmul(qmul(Dq, qconj(o.s1.q)), o.s0.q);
    const Qt<double> left = qmul(dq, qconj(o.s1.q));
    for (int j = 0; j < 3; ++j) {
        const Qt<double> ax = {j == 0 ? 1.0 : 0.0, j == 1 ? 1.0 : 0.0, j == 2 ? 1.0 : 0.0, 0.0};
        const V3<double> c1 = qvec(qmul(qmul(left, ax), o.s0.q));
        const Qt<double> dbq = {GA[j], GA[3 + j], GA[6 + j], 0.0};
        const V3<double> c2 = qvec(qmul(qmul(corr, dbq), tail));
        o.Jrot[j] = c1.x; o.Jrot[3 + j] = c1.y; o.Jrot[6 + j] = c1.z;
        o.Jbias[j] = c2.x; o.Jbias[3 + j] = c2.y; o.Jbias[6 + j] = c2.z;
    }
}

// Column c (0..29) of the unwhitened IMU Jacobian [pose0 6 | sb0 9 | pose1 6 | sb1 9].
__device__ void imu_raw_column(const double* d, const ImuShared& o, int c, double* col) {
    for (int i = 0; i < 15; ++i) col[i] = 0;
    const int k = c < 6 ? 0 : (c < 15 ? 1 : (c < 21 ? 2 : 3));
    const int cc = c - (k == 0 ? 0 : (k == 1 ? 6 : (k == 2 ? 15 : 21)));
    const bool first = k < 2;
    const double sign = first ? 1.0 : -1.0;
    ImuSensor s;  // the edge's sensor at this column's end (selected by value: no indexed copy)
    s.p = first ? o.s0.p : o.s1.p;
    s.v = first ? o.s0.v : o.s1.v;
    s.arm = first ? o.s0.arm : o.s1.arm;
    s.lever = first ? o.s0.lever : o.s1.lever;
    s.q = first ? o.s0.q : o.s1.q;
    const double* R = o.R.m;
    // column j < 3 of a row-major 3 x 3 matrix, by selects so that the matrix stays in registers
    auto mcol = [](const double* M, int j, double* out3) {
#pragma unroll
        for (int i = 0; i < 3; ++i) out3[i] = j == 0 ? M[3 * i] : (j == 1 ? M[3 * i + 1] : M[3 * i + 2]);
    };
    double v[3], tmp[9], Sk[9];
    if (k % 2 == 0) {
        if (cc < 3) {
            mcol(R, cc, v);
            for (int i = 0; i < 3; ++i) col[i] = sign * v[i];
        } else {
            const int j = cc - 3;
            skew_of(s.p, Sk); mm3(R, Sk, tmp); mcol(tmp, j, v);
            for (int i = 0; i < 3; ++i) col[i] = -sign * v[i];
            mcol(o.Jrot, j, v);
            for (int i = 0; i < 3; ++i) col[3 + i] = sign * v[i];
            skew_of(s.lever, Sk); mm3(R, Sk, tmp); mcol(tmp, j, v);
            for (int i = 0; i < 3; ++i) col[6 + i] = -sign * v[i];
            if (first) {
                double Sdp[9], Sl[9], M2[9];
                skew_of(o.dp, Sdp); skew_of(s.lever, Sl);
                for (int i = 0; i < 9; ++i) M2[i] = Sdp[i] - o.dt * Sl[i];
                mm3(R, M2, tmp); mcol(tmp, j, v);
                for (int i = 0; i < 3; ++i) col[i] += v[i];
                skew_of(o.dv, Sk); mm3(R, Sk, tmp); mcol(tmp, j, v);
                for (int i = 0; i < 3; ++i) col[6 + i] += v[i];
            }
        }
    } else {
        double lb[9];
        const M3<double> Rs = qmatrix(s.q);
        skew_of(s.arm, Sk); mm3(Sk, Rs.m, lb);
        if (cc < 3) {
            mcol(R, cc, v);
            for (int i = 0; i < 3; ++i) col[6 + i] = sign * v[i];
            if (first) for (int i = 0; i < 3; ++i) col[i] = o.dt * v[i];
        } else if (cc < 6) {
            const int j = cc - 3;
            mm3(R, lb, tmp); mcol(tmp, j, v);
            for (int i = 0; i < 3; ++i) col[6 + i] = sign * v[i];
            col[9 + j] = sign;
            if (first) {
                for (int i = 0; i < 3; ++i) col[i] = o.dt * v[i] + d[49 + i * 3 + j];
                mcol(o.Jbias, j, v);
                for (int i = 0; i < 3; ++i) col[3 + i] = v[i];
                for (int i = 0; i < 3; ++i) col[6 + i] += d[40 + i * 3 + j];
            }
        } else {
            const int j = cc - 6;
            col[12 + j] = sign;
            if (first) {
                for (int i = 0; i < 3; ++i) col[i] = -d[67 + i * 3 + j];
                for (int i = 0; i < 3; ++i) col[6 + i] = -d[58 + i * 3 + j];
            }
        }
    }
}

// Prior chart of block b at state x, with the central-difference chart derivative (6x6, h=1e-7)
// for SE(3) blocks (VisualCudaBatch.cu:216-246).
__device__ const double* prior_state(const WinDev& w, const PriorBlockDev& b, const double* pose, const double* sb,
                                     const double* rig, const double* td) {
    if (b.kind == 0) return pose + 7 * b.slot;
    if (b.kind == 1) return sb + 9 * b.slot;
    if (b.kind == 2) return rig;
    return td;
}

// Retained factors of one window, one wavefront per IMU edge plus one for the prior.
// mode 0 linearizes at the state (residuals, whitened Jacobians, chart derivatives, A^T r);
// mode 1 evaluates the cost at the trial state. Results go to per-edge slots that k_assemble /
// k_lm_update combine in a fixed order. grid = (windows, max_edges + 1), block = one wave.

// One retained item of a window (IMU edge y < E, or the prior at y == E) on one wave; scratch:
// 2 kMaxRetained doubles of the wave's own LDS (the prior's chart and residual). MODE 0
// linearizes (residual, Jacobians, prior chart derivative), MODE 1 evaluates the trial cost.
template <int MODE>
__device__ __forceinline__ void retained_body(const WinDev& w, int y, int lane, double* scratch) {
    constexpr int mode = MODE;
    if (!window_active(w)) return;
    if (mode == 0 && !w.lm->new_point) return;
    const int E = w.E;
    const double* pose = mode == 0 ? w.pose : w.cpose;
    const double* sb = mode == 0 ? w.sb : w.csb;
    const double* rig = mode == 0 ? w.rig : w.crig;
    const double* td = mode == 0 ? w.td : w.ctd;
    double* sc = w.ret_sc;
    if (y < E) {
        // the whitening W (15 x 15) and the raw Jacobian columns in the wave's LDS; the products
        // are split over the lanes (one per residual row, then one per Jacobian entry)
        const int s0 = w.imu_from[y], s1 = w.imu_to[y];
        const double* d = w.imu + (size_t)y * 301;
        double* Ws = scratch;
        double* rs = scratch + 225;
        double* cols = scratch + 240;  // column c of the raw Jacobian at cols[15 c ..]
        for (int k = lane; k < 225; k += 64) Ws[k] = d[76 + k];
        ImuShared o;
        imu_shared(d, pose + 7 * s0, sb + 9 * s0, pose + 7 * s1, sb + 9 * s1, mode == 0, o);
        wave_sync();
        int bad = 0;
        if (lane < 15) {
            double v = 0;
            for (int k = 0; k < 15; ++k) v += Ws[lane * 15 + k] * o.e[k];
            rs[lane] = v;
            bad = !isfinite(v);
            if (mode == 0) w.imu_r[(size_t)y * 15 + lane] = v;
        }
        if (mode == 0) {
            if (lane < 30) imu_raw_column(d, o, lane, cols + 15 * lane);
            wave_sync();
            double* J = w.imu_J + (size_t)y * 450;
            for (int q = lane; q < 450; q += 64) {
                const int i = q / 30, c = q % 30;
                double v = 0;
                for (int m = 0; m < 15; ++m) v += Ws[i * 15 + m] * cols[15 * c + m];
                J[q] = v;
                if (!isfinite(v)) bad = 1;
            }
        }
        wave_sync();
        for (int m = 32; m > 0; m >>= 1) bad |= __shfl_xor(bad, m, 64);
        if (lane == 0) {
            double c = 0;
            for (int i = 0; i < 15; ++i) c += 0.5 * rs[i] * rs[i];
            sc[mode * E + y] = c;
            sc[2 * E + mode * E + y] = bad ? 1.0 : 0.0;
        }
        return;
    }
    if (y != E) return;
    double* psc = sc + 4 * E;  // prior: [lin cost, trial cost, lin invalid, trial invalid]
    if (w.prior_rows <= 0) {
        if (lane == 0) { psc[mode] = 0; psc[2 + mode] = 0; }
        return;
    }
    double* chart = scratch;
    double* res = scratch + kMaxRetained;
    const int m = w.prior_rows, n = w.prior_cols, B = w.prior_blocks;
    for (int bi = lane; bi < B; bi += 64) {
        const PriorBlockDev b = w.pblk[bi];
        const double* x = prior_state(w, b, pose, sb, rig, td);
        const double* ref = w.pref + b.ref;
        const bool se3 = b.kind == 0 || b.kind == 2;
        if (se3) pose_minus(x, ref, chart + b.chart);
        else for (int c = 0; c < b.tsize; ++c) chart[b.chart + c] = x[c] - ref[c];
    }
    // chart derivatives of the SE(3) blocks, one lane per (block, tangent column)
    if (mode == 0)
        for (int it = lane; it < 6 * B; it += 64) {
            const int bi = it / 6, c = it % 6;
            const PriorBlockDev b = w.pblk[bi];
            if (!((b.kind == 0 || b.kind == 2) && b.width > 0)) continue;
            const double* x = prior_state(w, b, pose, sb, rig, td);
            const double* ref = w.pref + b.ref;
            const double h = 1e-7;
            double step[6], pp[7], pm[7], dp[6], dm[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) step[k] = k == c ? h : 0.0;
            pose_plus(x, step, pp);
#pragma unroll
            for (int k = 0; k < 6; ++k) step[k] = k == c ? -h : 0.0;
            pose_plus(x, step, pm);
            pose_minus(pp, ref, dp);
            pose_minus(pm, ref, dm);
            double* D = w.pD + (size_t)bi * 36;
#pragma unroll
            for (int r = 0; r < 6; ++r) D[r * 6 + c] = (dp[r] - dm[r]) / (2 * h);
        }
    wave_sync();
    double cost = 0;
    int bad = 0;
    for (int r = lane; r < m; r += 64) {
        const double* A = w.pA + (size_t)r * n;
        double v = w.pr0[r];
        for (int c = 0; c < n; ++c) v += A[c] * chart[c];
        res[r] = v;
        if (!isfinite(v)) bad = 1;
        cost += 0.5 * v * v;
    }
    wave_sync();
    if (mode == 0) {
        for (int c = lane; c < n; c += 64) {
            double v = 0;
            for (int r = 0; r < m; ++r) v += w.pA[(size_t)r * n + c] * res[r];
            w.patr[c] = v;
        }
    }
    cost = wave_reduce_sum(cost);
    for (int k = 32; k > 0; k >>= 1) bad |= __shfl_xor(bad, k, 64);
    if (lane == 0) {
        psc[mode] = cost;
        psc[2 + mode] = bad ? 1.0 : 0.0;
    }
}

// The retained items alone: grid (window, item), one wave each.
__global__ __launch_bounds__(64) void k_retained(const WinDev* wins, int mode) {
    __shared__ double scratch[kRetainedScratch];
    if (mode == 0) retained_body<0>(wins[blockIdx.x], blockIdx.y, threadIdx.x, scratch);
    else retained_body<1>(wins[blockIdx.x], blockIdx.y, threadIdx.x, scratch);
}

// ---------------------------------------------------------------------------------------------
// Fixed-order reduction of chunk partials: out[i] = sum_c part[c][i]; 16 entries per workgroup,
// chunks split over 16 lanes per entry and combined in lane order.
// grid.x = window, grid.y = slice
// ---------------------------------------------------------------------------------------------
constexpr int kRedEntries = 16;               // entries per workgroup (one 128-byte line per chunk)

// Entries [slice E, slice E + E) of window w's partial sum, which 0: visual block, 1: Schur
// correction; kBlock / E chunk lanes per entry, combined in lane order.
// Threads kBlock.. of a larger workgroup only take part in the barrier, so the summation order is
// the same in every launch that runs it.
template <int E>
__device__ __forceinline__ void reduce_slice(const WinDev& w, int which, double* out, int slice, double* red) {
    if (!window_active(w)) return;
    if (which == 0 && !w.lm->new_point) return;
    const int len = w.tri + w.nv;
    if (slice * E >= len) return;
    constexpr int lanes = kBlock / E;
    const bool act = threadIdx.x < kBlock;
    const int e = threadIdx.x % E, cl = threadIdx.x / E;
    const int i = slice * E + e;
    const double* part = which == 0 ? w.lin_part : w.schur_part;
    const int nparts = which == 0 ? w.nlin : w.nchunks;
    double s = 0;
    if (act && i < len) {
        // four partials' loads in flight at a time, added in the same order
        int c = cl;
        for (; c + 3 * lanes < nparts; c += 4 * lanes) {
            const double a0 = part[(size_t)c * len + i], a1 = part[(size_t)(c + lanes) * len + i];
            const double a2 = part[(size_t)(c + 2 * lanes) * len + i], a3 = part[(size_t)(c + 3 * lanes) * len + i];
            s += a0;
            s += a1;
            s += a2;
            s += a3;
        }
        for (; c < nparts; c += lanes) s += part[(size_t)c * len + i];
    }
    if (act) red[cl * E + e] = s;
    __syncthreads();
    if (act && cl == 0 && i < len) {
        double t = 0;
        for (int q = 0; q < lanes; ++q) t += red[q * E + e];
        out[i] = t;
    }
}

__global__ __launch_bounds__(kBlock) void k_reduce(const WinDev* wins, int which, double** outs) {
    __shared__ double red[kBlock];
    reduce_slice<kRedEntries>(wins[blockIdx.x], which, outs[2 * blockIdx.x + which], blockIdx.y, red);
}

// The visual-block partials are summed with 64 entries per workgroup (4 chunk lanes per entry)
// whether the sum runs inside the Schur launch or as a launch of its own: one summation order,
// so a window's result does not depend on the size of the batch it is solved in.
constexpr int kRedEntriesFused = 64;
__global__ __launch_bounds__(kBlock) void k_reduce_lin(const WinDev* wins, double** outs) {
    __shared__ double red[kBlock];
    reduce_slice<kRedEntriesFused>(wins[blockIdx.x], 0, outs[2 * blockIdx.x], blockIdx.y, red);
}

// ---------------------------------------------------------------------------------------------
// k_schur: -c c^T / damped(d) over each landmark's support (VisualCudaBatch.cu:145-161).
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ double damped_point(double d, double lambda) { return d + lambda * fmax(d, 1e-6); }

// The correction -sum_k c_k c_k^T / d_k is a product C^T diag(1/d) C of the chunk's dense
// coupling rows (c_k over the nv visual columns: anchor pose sum_f u_f, observer poses -u_f,
// rig, td), formed 64 landmarks at a time in LDS and multiplied on MFMA (v_mfma_f64_16x16x4f64,
// upper 16x16 tiles of the padded 128 x 128 block, 9 tiles per wave).
// FP64: row k holds sqrt(s_k) c_k and, in column nv, sqrt(s_k) g_k (s_k = 1 / damped(d_k) > 0,
// or the pseudo-inverse weight of a marginal system), so the tiles of C~^T C~ carry the
// gradient correction -sum_k s_k g_k c_k in column nv and the products need no scaling.
constexpr int kSchurRows = 64;                 // landmarks per LDS pass
constexpr int kSchurLd = 144;                  // LDS row stride (doubles; offsets rows by 128 B)
constexpr int kVisTiles = 8;                   // 128 / 16
constexpr int kSchurTiles = kVisTiles * (kVisTiles + 1) / 2;
// FP64: 512 threads (8 waves) per workgroup at 128 VGPRs: two workgroups fit a CU by LDS, so 16
// waves share its SIMDs and hide each other's load latency while the rows are built (256
// threads at 210 VGPRs ran 8 waves per CU). Mixed precision: 256 threads (its FP32 products
// need more registers per tile).
template <bool MIXED> constexpr int schur_block() { return MIXED ? 256 : 512; }
static_assert(kMaxVisual < kVisTiles * 16, "the gradient column fits the padded block");

// Mixed precision keeps the coupling rows in FP32 (37 KB of LDS) and multiplies them on
// v_mfma_f32_16x16x4f32; the FP32 products of each 16-landmark group are carried into FP64
// accumulators (the chunk partials, and their sum over chunks, stay FP64). Its gradient
// correction is summed in FP64 from the unscaled rows.
template <bool MIXED>
__device__ __forceinline__ void schur_body(const WinDev* wins, const Chunk* chunks) {
    using CT = typename std::conditional<MIXED, float, double>::type;
    constexpr int kSchurBlock = schur_block<MIXED>();
    constexpr int kSchurWaves = kSchurBlock / 64;
    constexpr int kSchurTilesPerWave = (kSchurTiles + kSchurWaves - 1) / kSchurWaves;
    constexpr int kSchurGroups = kSchurBlock / kChunkLanes;       // landmark groups per workgroup
    constexpr int kSchurRowsPerGroup = kSchurRows / kSchurGroups;  // rows one group builds per pass
    static_assert(kSchurRowsPerGroup * kSchurGroups == kSchurRows, "rows per pass");
    __shared__ CT crow[kSchurRows * kSchurLd];
    __shared__ CT sc[kSchurRows];
    __shared__ double qg[kSchurRows];
    __shared__ int pcol[kMaxFrames];
    const Chunk ch = chunks[blockIdx.x];
    const WinDev& w = wins[ch.win];
    if (!window_active(w)) return;
    const int nv = w.nv, tri = w.tri;
    const double lambda = w.lm->lambda;
    const int marginal = w.lm->marginal;
    const double marg_cut = w.lm->marg_cut;
    const int rig_col = w.rig_col, td_col = w.td_col;
    const int t = threadIdx.x, lane = t % 64, wave = t / 64;
    const int lgroup = t % kChunkLanes, group = t / kChunkLanes;
    const size_t L = (size_t)w.L, F = (size_t)w.F;
    if (t < kMaxFrames) pcol[t] = t < w.K ? w.pose_col[t] : -1;
    // this wave's output tiles (ti <= tj) of the padded block; FP64 also needs the tiles of
    // column nv (the gradient)
    const int last_col = MIXED ? nv - 1 : nv;
    int tti[kSchurTilesPerWave], ttj[kSchurTilesPerWave];
    d4 acc[kSchurTilesPerWave];
#pragma unroll
    for (int q = 0; q < kSchurTilesPerWave; ++q) {
        const int idx = wave + q * kSchurWaves;
        tti[q] = -1;
        ttj[q] = 0;
        acc[q] = d4{0.0, 0.0, 0.0, 0.0};
        if (idx < kSchurTiles) {
            int ti = 0, r = idx;
            while (r >= kVisTiles - ti) { r -= kVisTiles - ti; ++ti; }
            tti[q] = ti;
            ttj[q] = ti + r;
            if (16 * ttj[q] > last_col) tti[q] = -1;
        }
    }
    double gsum = 0;  // MIXED: gradient correction of column t (t < nv)
    // Dense coupling rows. Group g builds rows g + j groups (j < kSchurRowsPerGroup) of a pass from
    // three dependent global loads: the landmark record, the first factor of each lane (observer,
    // pair position), its coupling terms. In FP64 the passes are pipelined: a pass's landmark
    // records are loaded while the previous pass builds its rows, its factors during the previous
    // pass's MFMAs, so only the coupling-term round trip is left at the head of a pass. (Mixed
    // precision builds two rows per group; holding the next pass's loads would spill there.)
    constexpr bool kPipe = !MIXED;
    struct RowIn {
        int lmk, f1, fa, pa, fb, fp;
        bool live;
        double dl, gl;
    };
    auto load_records = [&](int base, RowIn* r) {
#pragma unroll
        for (int j = 0; j < kSchurRowsPerGroup; ++j) {
            const int l = base + group + j * kSchurGroups;
            const int lc = l < ch.lm_end ? l : ch.lm_begin;  // the loads stay inside the chunk
            r[j].lmk = l;
            r[j].live = l < ch.lm_end && w.lm_col[lc] >= 0;
            r[j].fa = w.lm_off[lc] + lgroup;
            r[j].f1 = r[j].live ? w.lm_off[lc + 1] : 0;  // a row that is not live has no factor to load
            r[j].pa = w.lm_fc[lc] >> 8;
            r[j].dl = w.lm_out[LO_D * L + lc];
            r[j].gl = w.lm_out[LO_G * L + lc];
        }
    };
    auto load_factors = [&](RowIn* r) {
#pragma unroll
        for (int j = 0; j < kSchurRowsPerGroup; ++j) {
            r[j].fb = -1;
            r[j].fp = 0;
            if (r[j].fa < r[j].f1) {
                r[j].fb = w.ob_fc[r[j].fa] >> 8;
                r[j].fp = w.pf_pos[r[j].fa];
            }
        }
    };
    RowIn cur[kSchurRowsPerGroup], nxt[kSchurRowsPerGroup];
    if constexpr (kPipe) {
        load_records(ch.lm_begin, cur);
        load_factors(cur);
    }
    for (int base = ch.lm_begin; base < ch.lm_end; base += kSchurRows) {
        if constexpr (!kPipe) {
            load_records(base, cur);
            load_factors(cur);
        }
        for (int i = t; i < kSchurRows * kSchurLd; i += kSchurBlock) crow[i] = CT(0);
        CT u[kSchurRowsPerGroup][6];  // FP32 hi parts in mixed precision (the products use them alone)
#pragma unroll
        for (int j = 0; j < kSchurRowsPerGroup; ++j)
#pragma unroll
            for (int k = 0; k < 6; ++k)
                u[j][k] = cur[j].fa < cur[j].f1 ? CT(load_u<MIXED>(w, F, k, cur[j].fp)) : CT(0);
        const bool more = kPipe && base + kSchurRows < ch.lm_end;
        if (more) load_records(base + kSchurRows, nxt);
        int lmk[kSchurRowsPerGroup], f1[kSchurRowsPerGroup], fa[kSchurRowsPerGroup], pa[kSchurRowsPerGroup];
        int fb[kSchurRowsPerGroup];
        bool live[kSchurRowsPerGroup];
        double dl[kSchurRowsPerGroup], gl[kSchurRowsPerGroup];
#pragma unroll
        for (int j = 0; j < kSchurRowsPerGroup; ++j) {
            lmk[j] = cur[j].lmk; f1[j] = cur[j].f1; fa[j] = cur[j].fa; pa[j] = cur[j].pa; fb[j] = cur[j].fb;
            live[j] = cur[j].live; dl[j] = cur[j].dl; gl[j] = cur[j].gl;
        }
        __syncthreads();  // rows cleared, pose columns staged
#pragma unroll
        for (int j = 0; j < kSchurRowsPerGroup; ++j) {
            const int r = group + j * kSchurGroups, l = lmk[j];
            // marginalization: pseudo-inverse of the diagonal landmark block (Marginalizer.cpp:373-409)
            double inv = 0;
            if (live[j]) inv = marginal ? (dl[j] > marg_cut ? 1.0 / dl[j] : 0.0) : 1.0 / damped_point(dl[j], lambda);
            const double rs = MIXED ? 1.0 : sqrt(inv);  // FP64 rows carry sqrt(s_k)
            if (lgroup == 0) {
                sc[r] = CT(inv);
                qg[r] = inv * gl[j];
            }
            if (!live[j]) continue;
            CT* row = crow + r * kSchurLd;
            const int a = pa[j], ac = pcol[a];
            if (fa[j] < f1[j]) {
                const int bc = fb[j] == a ? -1 : pcol[fb[j]];
#pragma unroll
                for (int k = 0; k < 6; ++k) {
                    const CT v = CT(rs * u[j][k]);
                    if (ac >= 0) atomicAdd(&row[ac + k], v);
                    if (bc >= 0) atomicAdd(&row[bc + k], -v);
                }
            }
            for (int f = fa[j] + kChunkLanes; f < f1[j]; f += kChunkLanes) {  // landmarks with more factors than lanes
                const int b = w.ob_fc[f] >> 8;
                const int bc = b == a ? -1 : pcol[b];
                const int p = w.pf_pos[f];
                for (int k = 0; k < 6; ++k) {
                    const CT v = CT(rs * load_u<MIXED>(w, F, k, p));
                    if (ac >= 0) atomicAdd(&row[ac + k], v);
                    if (bc >= 0) atomicAdd(&row[bc + k], -v);
                }
            }
            if (lgroup == 0) {
                if (rig_col >= 0)
                    for (int k = 0; k < 6; ++k) row[rig_col + k] = CT(rs * w.lm_out[(LO_CR0 + k) * L + l]);
                if (td_col >= 0) row[td_col] = CT(rs * w.lm_out[LO_CT * L + l]);
                if constexpr (!MIXED) row[nv] = CT(rs * gl[j]);
            }
        }
        __syncthreads();
        if (more) load_factors(nxt);
        // FP64: C~^T C~ on MFMA, A[i][k] = C~[k][16 ti + i] (lane i + 16 k), B[k][j] = C~[k][16 tj + j].
        // MIXED: C^T diag(s) C with the scale on the A operand.
#pragma unroll
        for (int q = 0; q < kSchurTilesPerWave; ++q) {
            if (tti[q] < 0) continue;
            const int ci = 16 * tti[q] + lane % 16, cj = 16 * ttj[q] + lane % 16;
            if constexpr (MIXED) {
                for (int g0 = 0; g0 < kSchurRows; g0 += 16) {
                    f4 part = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                    for (int k0 = g0; k0 < g0 + 16; k0 += 4) {
                        const int k = k0 + lane / 16;
                        part = __builtin_amdgcn_mfma_f32_16x16x4f32(sc[k] * crow[k * kSchurLd + ci], crow[k * kSchurLd + cj],
                                                                    part, 0, 0, 0);
                    }
#pragma unroll
                    for (int rr = 0; rr < 4; ++rr) acc[q][rr] += (double)part[rr];
                }
            } else {
#pragma unroll 4
                for (int k0 = 0; k0 < kSchurRows; k0 += 4) {
                    const int k = k0 + lane / 16;
                    acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(crow[k * kSchurLd + ci], crow[k * kSchurLd + cj], acc[q],
                                                                 0, 0, 0);
                }
            }
        }
        if constexpr (MIXED) {
            if (t < nv)
                for (int k = 0; k < kSchurRows; ++k) gsum += qg[k] * (double)crow[k * kSchurLd + t];
        }
        __syncthreads();
        if constexpr (kPipe)
#pragma unroll
            for (int j = 0; j < kSchurRowsPerGroup; ++j) cur[j] = nxt[j];
    }
    // chunk partial: packed upper triangle of -C^T diag(s) C, then -C^T (s g); output element
    // (row, col) of the tile in register rr: f64 MFMA row 4 rr + lane / 16, f32 MFMA row
    // 4 (lane / 16) + rr, column lane % 16 in both
    double* part = w.schur_part + (size_t)ch.local * (tri + nv);
#pragma unroll
    for (int q = 0; q < kSchurTilesPerWave; ++q) {
        if (tti[q] < 0) continue;
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
            const int ri = MIXED ? 4 * (lane / 16) + rr : 4 * rr + lane / 16;
            const int i = 16 * tti[q] + ri, j = 16 * ttj[q] + lane % 16;
            if (i <= j && j < nv) part[tri_index(i, j, nv)] = -acc[q][rr];
            else if (!MIXED && j == nv && i < nv) part[tri + i] = -acc[q][rr];
        }
    }
    if (MIXED && t < nv) part[tri + t] = -gsum;
}

// Workgroups nchunks.. sum the visual-block partials of the linearization (64 entries each,
// rslices per window), which the Schur products do not need: k_assemble reads both sums.
__global__ __launch_bounds__(schur_block<false>()) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_schur(const WinDev* wins, const Chunk* chunks, int nchunks, int rslices,
                                                      double** outs) {
    if ((int)blockIdx.x >= nchunks) {
        __shared__ double red[kBlock];
        const int r = blockIdx.x - nchunks, win = r / rslices;
        reduce_slice<kRedEntriesFused>(wins[win], 0, outs[2 * win], r % rslices, red);
        return;
    }
    schur_body<false>(wins, chunks);
}
__global__ __launch_bounds__(schur_block<true>()) __attribute__((amdgpu_waves_per_eu(2, 4))) void k_schur_mixed(const WinDev* wins, const Chunk* chunks, int nchunks,
                                                            int rslices, double** outs) {
    if ((int)blockIdx.x >= nchunks) {
        __shared__ double red[kBlock];
        const int r = blockIdx.x - nchunks, win = r / rslices;
        reduce_slice<kRedEntriesFused>(wins[win], 0, outs[2 * win], r % rslices, red);
        return;
    }
    schur_body<true>(wins, chunks);
}

// ---------------------------------------------------------------------------------------------
// k_assemble: full undamped normal matrix of the states (when newly linearized), then the damped
// reduced camera system S and rhs (prepareStateSolve, VisualCudaBatch.cu:163-172).
// red_lin / red_schur: reduced partials (tri + nv) per window.
// ---------------------------------------------------------------------------------------------
constexpr int kAssembleRows = 1;                       // rows per workgroup (320 workgroups at nr = 307)
constexpr int kAssembleSplit = kTile / kAssembleRows;  // workgroups per tile row

// Retained (IMU + prior) normal-matrix entry (i, j) from the per-edge Jacobians and the prior
// chart derivatives / Gram matrix, using the static column map. pv: row i of G D_i^T-side
// product, pv[c] = sum_x D_i[x][ii] G[chart_i + x][c] over the prior's chart columns c (formed
// once per row by prior_row), so the prior part of an entry is at most six products.
__device__ __forceinline__ double retained_entry(const WinDev& w, const int* ci, const int* cj, const double* pv) {
    double h = 0;
    for (int a = 0; a < 2; ++a) {
        if (ci[a] < 0) continue;
        const int e = ci[a] >> 8, li = ci[a] & 255;
        for (int b = 0; b < 2; ++b) {
            if (cj[b] < 0 || (cj[b] >> 8) != e) continue;
            const int lj = cj[b] & 255;
            const double* J = w.imu_J + (size_t)e * 450;
            double v = 0;
            for (int m = 0; m < 15; ++m) v += J[m * 30 + li] * J[m * 30 + lj];
            h += v;
        }
    }
    if (ci[2] >= 0 && cj[2] >= 0) {
        const int bj = cj[2] >> 8, jj = cj[2] & 255;
        const PriorBlockDev b2 = w.pblk[bj];
        const bool se2 = b2.kind == 0 || b2.kind == 2;
        if (pv) {
            if (se2) {
                double inner = 0;
                for (int yy = 0; yy < 6; ++yy) inner += pv[b2.chart + yy] * w.pD[bj * 36 + yy * 6 + jj];
                h += inner;
            } else {
                h += pv[b2.chart + jj];
            }
        } else {  // a prior wider than the row buffer: the entry's own products
            const int bi = ci[2] >> 8, ii = ci[2] & 255;
            const PriorBlockDev b1 = w.pblk[bi];
            const bool se1 = b1.kind == 0 || b1.kind == 2;
            const int n = w.prior_cols;
            for (int x = 0; x < b1.tsize; ++x) {
                const double d1 = se1 ? w.pD[bi * 36 + x * 6 + ii] : (x == ii ? 1.0 : 0.0);
                if (d1 == 0) continue;
                const double* G = w.pG + (size_t)(b1.chart + x) * n + b2.chart;
                double inner = 0;
                if (se2) {
                    for (int yy = 0; yy < 6; ++yy) inner += G[yy] * w.pD[bj * 36 + yy * 6 + jj];
                } else {
                    inner = G[jj];
                }
                h += d1 * inner;
            }
        }
    }
    return h;
}

// pv[c] = sum_x D_i[x][ii] G[chart_i + x][c] for the prior's chart columns c of the row whose
// column map is ci (a prior column); the workgroup's threads split the columns.
__device__ __forceinline__ void prior_row(const WinDev& w, const int* ci, double* pv) {
    const int bi = ci[2] >> 8, ii = ci[2] & 255;
    const PriorBlockDev b1 = w.pblk[bi];
    const bool se1 = b1.kind == 0 || b1.kind == 2;
    const int n = w.prior_cols;
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        double v = 0;
        for (int x = 0; x < b1.tsize; ++x) {
            const double d1 = se1 ? w.pD[bi * 36 + x * 6 + ii] : (x == ii ? 1.0 : 0.0);
            if (d1 != 0) v += d1 * w.pG[(size_t)(b1.chart + x) * n + c];
        }
        pv[c] = v;
    }
}

// Operand layout of a 16x16 tile: element (i, c) at (c / 4) * 64 + i + 16 (c % 4), so the 64
// doubles of column group c / 4 are the A or B operand of one MFMA (lane i + 16 k). It is the
// accumulator layout of the transposed tile; the factor tile store and the LDS panels use it.
__device__ __forceinline__ int opnd_pos(int i, int c) { return (c >> 2) * 64 + i + 16 * (c & 3); }

__global__ __launch_bounds__(kBlock) void k_assemble(WinDev* wins, double* const* reduced) {
    __shared__ double red[2 * kBlock];
    __shared__ double srow[kAssembleRows][kMaxRetained];  // the rows: undamped, then damped
    WinDev& w = wins[blockIdx.x];
    if (!window_active(w)) return;
    const int nr = w.nr, nv = w.nv;
    const int ti = blockIdx.y / kAssembleSplit, q = blockIdx.y % kAssembleSplit;
    if (kTile * ti >= nr) return;
    const int row0 = kTile * ti + kAssembleRows * q;
    const int row1 = min(nr, row0 + kAssembleRows);
    const VIB_G double* rl = (const VIB_G double*)reduced[2 * blockIdx.x + 0];
    const VIB_G double* rs = (const VIB_G double*)reduced[2 * blockIdx.x + 1];
    const int t = threadIdx.x, nt = blockDim.x;
    const bool fresh = w.lm->new_point;
    if (fresh) {
        static_assert(kAssembleRows == 1, "one prior row product per workgroup");
        // the prior products of this workgroup's row (a padding row of the last tile, row0 >= nr,
        // has no column map entry and no entries)
        __shared__ double prow[kMaxRetained];
        const double* pv = w.prior_cols <= kMaxRetained ? prow : nullptr;
        if (pv && row0 < nr && w.colmap[4 * row0 + 2] >= 0) prior_row(w, w.colmap + 4 * row0, prow);
        __syncthreads();
        for (int p = t; p < (row1 - row0) * nr; p += nt) {
            const int i = row0 + p / nr, j = p % nr;
            double h = retained_entry(w, w.colmap + 4 * i, w.colmap + 4 * j, pv);
            if (i < nv && j < nv) h += rl[i <= j ? tri_index(i, j, nv) : tri_index(j, i, nv)];
            if (w.hostH) h += w.hostH[(size_t)i * nr + j];
            w.Hfull[(size_t)i * nr + j] = h;
            srow[i - row0][j] = h;
        }
        for (int i = row0 + t; i < row1; i += nt) {
            const int* ci = w.colmap + 4 * i;
            double g = i < nv ? rl[w.tri + i] : 0.0;
            for (int a = 0; a < 2; ++a) {
                if (ci[a] < 0) continue;
                const int e = ci[a] >> 8, li = ci[a] & 255;
                const double* J = w.imu_J + (size_t)e * 450;
                const double* r = w.imu_r + (size_t)e * 15;
                for (int m = 0; m < 15; ++m) g += J[m * 30 + li] * r[m];
            }
            if (ci[2] >= 0) {
                const int bi = ci[2] >> 8, ii = ci[2] & 255;
                const PriorBlockDev b = w.pblk[bi];
                const bool se3 = b.kind == 0 || b.kind == 2;
                for (int x = 0; x < b.tsize; ++x)
                    g += (se3 ? w.pD[bi * 36 + x * 6 + ii] : (x == ii ? 1.0 : 0.0)) * w.patr[b.chart + x];
            }
            if (w.hostG) g += w.hostG[i];
            w.gfull[i] = g;
        }
        if (blockIdx.y == 0) {  // cost at the linearization point (fixed-order sums)
            double cost = 0, invalid = 0;
            for (int k = t; k < w.nlin; k += nt) {
                cost += w.lin_sc[(size_t)k * CS_PLANES + CS_COST];
                invalid = fmax(invalid, w.lin_sc[(size_t)k * CS_PLANES + CS_INVALID]);
            }
            for (int e = t; e < w.E; e += nt) {
                cost += w.ret_sc[e];
                invalid = fmax(invalid, w.ret_sc[2 * w.E + e]);
            }
            if (t == 0) {
                cost += w.ret_sc[4 * w.E];
                invalid = fmax(invalid, w.ret_sc[4 * w.E + 2]);
            }
            double sums[2] = {cost, invalid};
            block_reduce_n<2, 0b10u>(sums, red);
            cost = sums[0];
            invalid = sums[1];
            if (t == 0) {
                w.lm->lin_cost = cost;
                w.lm->lin_valid = !(invalid > 0) && isfinite(cost);
            }
        }
        __syncthreads();
    }
    // the damped row, once: to the saved system and to LDS, from where the tile store below
    // takes it in the factorization order
    const double lambda = w.lm->lambda;
    for (int p = t; p < (row1 - row0) * nr; p += nt) {
        const int i = row0 + p / nr, j = p % nr;
        const double h = fresh ? srow[i - row0][j] : w.Hfull[(size_t)i * nr + j];
        double v = h;
        if (i < nv && j < nv) v += rs[i <= j ? tri_index(i, j, nv) : tri_index(j, i, nv)];
        if (i == j) v += lambda * fmax(h, 1e-6);
        w.Scopy[(size_t)i * nr + j] = v;
        srow[i - row0][j] = v;
    }
    for (int i = row0 + t; i < row1; i += nt) w.rhs[i] = -(w.gfull[i] + (i < nv ? rs[w.tri + i] : 0.0));
    __syncthreads();
    // this workgroup's rows of the factor tile store, in the factorization order (speed/bias
    // columns first; padding rows and columns stay in place as identity) and operand layout
    const int nsb = nr - nv;
    for (int rho = 0; rho < kAssembleRows; ++rho) {
        const int r = row0 + rho;
        const int P = r >= nr ? r : (r < nv ? r + nsb : r - nv);
        for (int pc = t; pc < (P / kTile + 1) * kTile; pc += nt) {
            const int c = pc >= nr ? pc : (pc < nsb ? pc + nv : pc - nsb);
            w.fac[tile_base(P / kTile, pc / kTile) + opnd_pos(P % kTile, pc % kTile)] =
                (r < nr && c < nr) ? srow[rho][c] : (r == c ? 1.0 : 0.0);
        }
    }
}

// ---------------------------------------------------------------------------------------------
// k_cholesky: FP64 Cholesky of the damped reduced camera system on 16x16 tiles (right-looking),
// fused with the forward substitution, then the backward substitution, optional iterative
// refinement against the saved damped system, the step clamp, the model decrease of the states
// and the state trial x [+] delta (VisualCudaBatch.cu:163-189,323-340). One workgroup per window.
//
// The factor lives in the window's tile store (lower-triangle tiles in the accumulator layout of
// v_mfma_f64_16x16x4f64, written by k_assemble, padding rows = identity). Per tile column k:
//   A  every wave solves its panel tiles X = A L_kk^-T with MFMA against the inverse diagonal
//      tile (LDS, operand layout), stores them to the tile store; wave 0 first applies
//      L_kk^-1 to the right-hand side block.
//   B  waves 1.. update the trailing tiles T_ij -= X_i X_j^T with MFMA (operands from LDS, the
//      tile itself from the tile store; tiles of column k + 1 go to the next LDS panel).
//      Wave 0 updates the next diagonal tile, factors and inverts it in registers (lookahead)
//      and updates the right-hand side below the block.
// Two workgroup barriers per tile column.
// ---------------------------------------------------------------------------------------------
typedef __attribute__((address_space(1))) double gdouble;
constexpr int kTileDoubles = kTile * kTile;
constexpr int kCholThreads = 1024;
static_assert(kCholThreads / 64 >= 3, "the factorization needs a diagonal wave and two trailing waves");


// Value of v in lane src (src wave-uniform).
__device__ __forceinline__ double lane_bcast(double v, int src) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), src);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over the 16 lanes of the lane's row (quad xor 1, xor 2, half-row and row mirrors); every
// lane of the row gets the same bits.
__device__ __forceinline__ double row16_sum(double v) {
    v += __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false);   // quad_perm [1, 0, 3, 2]
    v += __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false);   // quad_perm [2, 3, 0, 1]
    v += __builtin_amdgcn_mov_dpp(v, 0x141, 0xF, 0xF, false);  // row_half_mirror
    v += __builtin_amdgcn_mov_dpp(v, 0x140, 0xF, 0xF, false);  // row_mirror
    return v;
}

// Value of v in lane Q of the lane's own 16-lane row (one v_mov_b64_dpp row_newbcast).
template <int Q>
__device__ __forceinline__ double row_bcast(double v) {
    return __builtin_amdgcn_mov_dpp(v, 0x150 + Q, 0xF, 0xF, false);
}

// Sum over the 64 lanes, the same bits in every lane: row sums by DPP, then the four rows'
// sums read from lanes 0, 16, 32, 48 and added in that order.
__device__ __forceinline__ double wave_sum_dpp(double v) {
    v = row16_sum(v);
    return (lane_bcast(v, 0) + lane_bcast(v, 16)) + (lane_bcast(v, 32) + lane_bcast(v, 48));
}

// 1 / sqrt(d) from the hardware estimate and two Newton steps.
__device__ __forceinline__ double rsqrt_nr(double d) {
    double q = __builtin_amdgcn_rsq(d);
    const double h = 0.5 * d;
    q = q * fma(-h * q, q, 1.5);
    q = q * fma(-h * q, q, 1.5);
    return q;
}

// Rank-1 update of columns C.. of the rows with column J of L: A[r][c] -= L[r][J] L[c][J].
template <int J, int C>
__device__ __forceinline__ void factor_update(double (&v)[kTile]) {
    if constexpr (C < kTile) {
        v[C] = fma(-v[J], row_bcast<C>(v[J]), v[C]);
        asm volatile("" : "+v"(v[C]));  // update in place now: the broadcast dies here
        factor_update<J, C + 1>(v);
    }
}

// Row Q of M = L^-1 (forward elimination): lane (r, g) of the wave holds columns 4 cc + g of
// row r of M in m[cc]; v[Q] is column Q of L (final) with 1 / L[Q][Q] on the diagonal lane.
template <int Q>
__device__ __forceinline__ void inverse_row(double (&m)[4], const double (&v)[kTile], int r) {
    const double sc = r == Q ? row_bcast<Q>(v[Q]) : 1.0;
#pragma unroll
    for (int cc = 0; cc < 4; ++cc)
        if (4 * cc <= Q) m[cc] *= sc;
    const double lq = r > Q ? v[Q] : 0.0;
#pragma unroll
    for (int cc = 0; cc < 4; ++cc)
        if (4 * cc <= Q) {
            m[cc] = fma(-lq, row_bcast<Q>(m[cc]), m[cc]);
            asm volatile("" : "+v"(m[cc]));
        }
}

// Step J of the right-looking factor (column J of L: pivot, scale, rank-1 update of the columns
// right of J) fused with row J - 1 of the inverse, which needs only column J - 1 of L: the two
// dependency chains are independent within a step and interleave. Lane r of a 16-lane row holds
// row r of the tile; every pivot must be positive and finite.
template <int J>
__device__ __forceinline__ void factor_inverse(double (&v)[kTile], double (&m)[4], int r, bool& ok) {
    if constexpr (J <= kTile) {
        if constexpr (J < kTile) {
            const double d = row_bcast<J>(v[J]);
            ok = ok && d > 0 && isfinite(d);
            const double is = rsqrt_nr(d);
            const double sc = v[J] * is;
            v[J] = r > J ? sc : (r == J ? is : 0.0);
            // rows above J hold zeros in column J; entries right of the diagonal are never read
            factor_update<J, J + 1>(v);
        }
        if constexpr (J > 0) inverse_row<J - 1>(m, v, r);
        factor_inverse<J + 1>(v, m, r, ok);
    }
}

// Cholesky factor and inverse of one SPD 16x16 tile; every 16-lane row of the wave holds the
// tile row-wise (v[c] = A[lane % 16][c], c <= lane % 16 used). The factor is formed redundantly
// in the four rows (the diagonal keeps 1 / L[j][j]), the inverse split over them by column
// (c % 4 = row index of the lane group); L^-1 goes to inv (LDS) and ginv (the tile's slot in the
// tile store, read by the substitutions) in operand layout. Returns false when a pivot is not
// positive and finite.
__device__ bool factor_diag_tile(double (&v)[kTile], double* inv, gdouble* ginv) {
    const int lane = threadIdx.x % 64, r = lane % kTile, g = lane / kTile;
    bool ok = true;
    double m[4];
#pragma unroll
    for (int cc = 0; cc < 4; ++cc) m[cc] = r == 4 * cc + g ? 1.0 : 0.0;
    factor_inverse<0>(v, m, r, ok);
#pragma unroll
    for (int cc = 0; cc < 4; ++cc) {
        inv[cc * 64 + lane] = m[cc];
        ginv[cc * 64 + lane] = m[cc];
    }
    return ok;
}

// Updates the tile (ti, tj) of the trailing matrix with the solved panel of column k,
// T -= X_ti X_tj^T (panel tiles indexed relative to k). The accumulator holds T in operand
// layout, which is the accumulator layout of T^T: the product is formed as T^T -= X_tj X_ti^T.
__device__ __forceinline__ d4 trailing_mfma(d4 acc, const double* panel, int ai, int aj, int lane) {
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
        const double a = -panel[(aj * 4 + kk) * 64 + lane];
        const double b = panel[(ai * 4 + kk) * 64 + lane];
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
    return acc;
}

// Accumulator registers of a tile in operand layout -> LDS tile (lane-linear, conflict free).
__device__ __forceinline__ void acc_to_panel(const d4& acc, double* tile, int lane) {
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) tile[rr * 64 + lane] = acc[rr];
}

constexpr int kRegSlots = 7;   // register-resident trailing tiles per trailing wave
constexpr int kGlobBatch = 3;  // tile-store trailing tiles in flight per wave

// Panel solve of tile column k: X_a = A_a L_kk^-T for the panel tiles a = 1.. (relative to k),
// written to the LDS panel (operand layout) and to the tile store. Wave 0 solves only X_1, the
// operand of the next diagonal tile, into LDS (its tile-store copy is written after the barrier
// by wave 1, so wave 0 has no stores in flight when it needs its prefetched diagonal tile); the
// other waves solve X_2.. round-robin, and the last wave first applies L_kk^-1 to the
// right-hand side block y_k.
__device__ __forceinline__ void solve_panel_tile(gdouble* Lt, double* panel, const double* inv, int k, int a, int lane,
                                                 bool to_store) {
    d4 acc = {0.0, 0.0, 0.0, 0.0};  // X^T = L_kk^-1 A_a^T: X in operand layout
#pragma unroll
    for (int kk = 0; kk < 4; ++kk)
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(inv[kk * 64 + lane], panel[(a * 4 + kk) * 64 + lane], acc, 0, 0, 0);
    acc_to_panel(acc, panel + a * kTileDoubles, lane);
    if (to_store) {
        const int base = tile_base(k + a, k);
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) Lt[base + rr * 64 + lane] = acc[rr];
    }
}

__device__ __forceinline__ void panel_solve(gdouble* Lt, double* panel, const double* inv, double* y, int k, int m,
                                            int wave, int lane, int nw, unsigned mk) {
    if (wave == 0) {
        if (m >= 1 && ((mk >> (k + 1)) & 1u)) solve_panel_tile(Lt, panel, inv, k, 1, lane, false);
        return;
    }
    if (wave == nw - 1) {
        double yr = 0;
        if (lane < kTile) {
#pragma unroll
            for (int c = 0; c < kTile; ++c) yr = fma(inv[opnd_pos(lane, c)], y[kTile * k + c], yr);
        }
        wave_sync();
        if (lane < kTile) y[kTile * k + lane] = yr;
    }
    for (int a = 1 + wave; a <= m; a += nw - 1)
        if ((mk >> (k + a)) & 1u) solve_panel_tile(Lt, panel, inv, k, a, lane, true);
}

// Tile-store copy of the LDS panel tile X_1 of column k.
__device__ __forceinline__ void store_panel_tile1(gdouble* Lt, const double* panel, int k, int lane) {
    const int base = tile_base(k + 1, k);
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) Lt[base + rr * 64 + lane] = panel[kTileDoubles + rr * 64 + lane];
}

// Lower Cholesky of the padded tile store Lt (T x T tiles) with the forward substitution of y
// (LDS, 16 T entries, padding zero) fused. pnl: two LDS panels; dtile: the diagonal tiles 1..
// (LDS, accumulator layout); dcur: the inverse diagonal tiles of the current and the next step
// (LDS). Wave 0 runs the diagonal tiles (update, factor, inverse: one step ahead of the trailing
// update); the other waves run the trailing update. The off-diagonal trailing tiles with the
// largest column index (register_tile order) stay in accumulator registers of the trailing
// waves for the whole factorization, the diagonal tiles in LDS; the other off-diagonal tiles are
// read from and written back to the tile store at every update. The inverse diagonal tiles
// replace the diagonal tiles of the tile store. The two roles run separate step loops with the
// same barrier sequence, so the resident tiles do not constrain the factor wave's registers.
__device__ __forceinline__ bool cholesky_forward(double* Lt_, int T, double* pnl, double* dtile, double* dcur, double* y,
                                                 int* flag, const uint32_t* mask) {
    gdouble* Lt = (gdouble*)Lt_;
    const int t = threadIdx.x, nt = blockDim.x, lane = t % 64, nw = nt / 64, nwt = nw - 1;
    const int wave = __builtin_amdgcn_readfirstlane(t / 64);
    constexpr int PANEL = kMaxTiles * kTileDoubles;
    for (int p = t; p < T * kTileDoubles; p += nt) {
        const int a = p / kTileDoubles, e = p % kTileDoubles;
        pnl[p] = Lt[tile_base(a, 0) + e];
        if (a > 0) dtile[p] = Lt[tile_base(a, a) + e];
    }
    __syncthreads();
    auto inv_of = [&](int k) { return dcur + (k & 1) * kTileDoubles; };
    if (wave == 0) {
        // diagonal tile k: factor and inverse
        auto factor = [&](double (&v)[kTile], int k) { return factor_diag_tile(v, inv_of(k), Lt + tile_base(k, k)); };
        {
            double v[kTile];
#pragma unroll
            for (int c = 0; c < kTile; ++c) v[c] = pnl[opnd_pos(lane % kTile, c)];
            if (!factor(v, 0) && lane == 0) *flag = 1;
        }
        __syncthreads();
        for (int k = 0; k < T; ++k) {
            if (*flag) return false;
            double* panel = pnl + (k & 1) * PANEL;
            double* next = pnl + ((k + 1) & 1) * PANEL;
            const int m = T - k - 1;
            d4 acc = {0.0, 0.0, 0.0, 0.0};
            if (m > 0) {  // the next diagonal tile is final since the last step: fetch it early
#pragma unroll
                for (int rr = 0; rr < 4; ++rr) acc[rr] = dtile[(k + 1) * kTileDoubles + rr * 64 + lane];
            }
            const unsigned mk = mask[k];
            panel_solve(Lt, panel, inv_of(k), y, k, m, wave, lane, nw, mk);
            __syncthreads();
            if (m > 0) {
                if ((mk >> (k + 1)) & 1u) acc = trailing_mfma(acc, panel, 1, 1, lane);
                acc_to_panel(acc, next, lane);
                wave_sync();
                double v[kTile];
#pragma unroll
                for (int c = 0; c < kTile; ++c) v[c] = next[opnd_pos(lane % kTile, c)];
                if (!factor(v, k + 1) && lane == 0) *flag = 1;
            }
            __syncthreads();
        }
        return true;
    }
    const int noff = (T - 1) * (T - 2) / 2;
    const int nreg = min(noff, kRegSlots * nwt);
    const int G = noff - nreg;
    d4 racc[kRegSlots];
    int rti[kRegSlots], rtj[kRegSlots];
#pragma unroll
    for (int s = 0; s < kRegSlots; ++s) {
        const int e = s * nwt + wave - 1;
        rti[s] = 0;
        rtj[s] = -1;
        if (e < nreg) {
            register_tile(T, e, rti[s], rtj[s]);
            const int base = tile_base(rti[s], rtj[s]);
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) racc[s][rr] = Lt[base + rr * 64 + lane];
        } else {
            racc[s] = d4{0.0, 0.0, 0.0, 0.0};
        }
    }
    // the loads complete here, once: a load still pending at the step loop would make the first
    // accumulator use of every step wait for all memory traffic in flight
#pragma unroll
    for (int s = 0; s < kRegSlots; ++s)
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) asm volatile("" ::"v"(racc[s][rr]));
    __syncthreads();
    for (int k = 0; k < T; ++k) {
        if (*flag) return false;
        double* panel = pnl + (k & 1) * PANEL;
        double* next = pnl + ((k + 1) & 1) * PANEL;
        const int m = T - k - 1;
        const unsigned mk = mask[k];  // tiles of panel column k that may be nonzero
        auto live = [&](int ti) { return ((mk >> ti) & 1u) != 0; };
        panel_solve(Lt, panel, inv_of(k), y, k, m, wave, lane, nw, mk);
        __syncthreads();
        if (wave == 1 && m >= 1 && live(k + 1)) store_panel_tile1(Lt, panel, k, lane);
        // right-hand side rows below the block
        for (int rr = (wave - 1) * 64 + lane; rr < kTile * m; rr += nwt * 64) {
            const int a = 1 + rr / kTile, i = rr % kTile;
            if (!live(k + a)) continue;
            double s = 0;
#pragma unroll
            for (int c = 0; c < kTile; ++c) s = fma(panel[a * kTileDoubles + opnd_pos(i, c)], y[kTile * k + c], s);
            y[kTile * (k + a) + i] -= s;
        }
        // tile-store items of this step: the off-diagonal tiles outside the register set with
        // column > k
        const int Fk = k * (T - 1) - k * (k + 1) / 2;
        const int nitems = max(0, G - Fk);
        d4 gacc[kGlobBatch];
        int gi[kGlobBatch], gj[kGlobBatch];
        int w0 = wave - 1;
        auto load_batch = [&]() {
#pragma unroll
            for (int b = 0; b < kGlobBatch; ++b) {
                const int w = w0 + b * nwt;
                gi[b] = -1;
                gj[b] = 0;
                if (w < nitems) {
                    store_tile(T, Fk + w, gi[b], gj[b]);
                    // untouched and not moving to the next panel: nothing to do
                    if (!(live(gi[b]) && live(gj[b])) && gj[b] != k + 1) {
                        gi[b] = -1;
                        continue;
                    }
                    const int base = tile_base(gi[b], gj[b]);
#pragma unroll
                    for (int rr = 0; rr < 4; ++rr) gacc[b][rr] = Lt[base + rr * 64 + lane];
                }
            }
        };
        auto finish_batch = [&]() {
#pragma unroll
            for (int b = 0; b < kGlobBatch; ++b) {
                if (gi[b] < 0) continue;
                if (live(gi[b]) && live(gj[b])) gacc[b] = trailing_mfma(gacc[b], panel, gi[b] - k, gj[b] - k, lane);
                if (gj[b] == k + 1) {
                    acc_to_panel(gacc[b], next + (gi[b] - k - 1) * kTileDoubles, lane);
                } else {
                    const int base = tile_base(gi[b], gj[b]);
#pragma unroll
                    for (int rr = 0; rr < 4; ++rr) Lt[base + rr * 64 + lane] = gacc[b][rr];
                }
            }
        };
        load_batch();
#pragma unroll
        for (int s = 0; s < kRegSlots; ++s) {
            if (rtj[s] <= k) continue;
            if (live(rti[s]) && live(rtj[s])) racc[s] = trailing_mfma(racc[s], panel, rti[s] - k, rtj[s] - k, lane);
            if (rtj[s] == k + 1) acc_to_panel(racc[s], next + (rti[s] - k - 1) * kTileDoubles, lane);
        }
        // diagonal tiles k + 2.. (LDS)
        for (int w = wave - 1; w < T - k - 2; w += nwt) {
            if (!live(k + 2 + w)) continue;
            double* dt = dtile + (k + 2 + w) * kTileDoubles;
            d4 acc;
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) acc[rr] = dt[rr * 64 + lane];
            acc = trailing_mfma(acc, panel, w + 2, w + 2, lane);
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) dt[rr * 64 + lane] = acc[rr];
        }
        finish_batch();
        for (w0 += kGlobBatch * nwt; w0 < nitems; w0 += kGlobBatch * nwt) {
            load_batch();
            finish_batch();
        }
        __syncthreads();
    }
    return true;
}

// acc += sum_c inv[(c, r)] u[c] with u[c] taken from lane c of the lane's 16-lane row.
template <int C>
__device__ __forceinline__ void bcast_dot(double& acc, const double* inv, double u, int r) {
    if constexpr (C < kTile) {
        acc = fma(inv[opnd_pos(C, r)], row_bcast<C>(u), acc);
        bcast_dot<C + 1>(acc, inv, u, r);
    }
}

// L^T x = y in place (x in LDS, 16 T entries) with the panel tiles from the tile store and the
// inverse diagonal tiles from dinv; one barrier per tile row. In round j wave 0 forms
//   x_j = L_jj^-T (y_j - L_{j+1,j}^T x_{j+1} - sum of the other waves' partials of round j + 1)
// and then L_{j,j-1}^T x_j for round j - 1 (its tile prefetched), while the other waves form
// the partials sum_{a >= j+1} L_{a,j-1}^T x_a of round j - 1, which need x only up to x_{j+1}.
// The next round's tiles are fetched after this round's work. part: two buffers (round parity)
// of 16 doubles per wave, then 16 doubles for wave 0's term.
constexpr int kBackWaves = kCholThreads / 64 - 1;                          // partial waves
constexpr int kBackSlots = (kMaxTiles - 2 + kBackWaves - 1) / kBackWaves;  // tiles per partial wave
constexpr int kBackPart = 2 * (kCholThreads / 64) * kTile + kTile;         // doubles of part
__device__ void cholesky_backward(const double* Lt_, int T, const double* dinv, double* x, double* part) {
    const gdouble* Lt = (const gdouble*)Lt_;
    const int t = threadIdx.x, wave = t / 64, lane = t % 64, nw = blockDim.x / 64, r = lane % 16;
    double* tterm = part + 2 * nw * kTile;
    d4 cur[kBackSlots], nxt[kBackSlots];  // partial waves: tiles (a, j - 1), a >= j + 1
    d4 tcur = {0.0, 0.0, 0.0, 0.0}, tnxt = {0.0, 0.0, 0.0, 0.0};  // wave 0: tile (j, j - 1)
    double tj = 0;                                                 // wave 0: (L_{j+1,j}^T x_{j+1})[r]
    auto load_round = [&](int j, d4 (&pw)[kBackSlots], d4& w0) {  // the tiles of round j
        if (j < 1) return;
        if (wave == 0) {
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) w0[rr] = Lt[tile_base(j, j - 1) + rr * 64 + lane];
        } else {
#pragma unroll
            for (int s = 0; s < kBackSlots; ++s) {
                const int a = j + wave + s * kBackWaves;
                if (a < T)
#pragma unroll
                    for (int rr = 0; rr < 4; ++rr) pw[s][rr] = Lt[tile_base(a, j - 1) + rr * 64 + lane];
            }
        }
    };
    load_round(T - 1, cur, tcur);
    for (int j = T - 1; j >= 0; --j) {
        if (wave == 0) {
            double u = x[kTile * j + r] - tj;
            if (j + 1 < T)
                for (int w = 1; w < nw; ++w) u -= part[(j & 1) * nw * kTile + w * kTile + r];
            const double* inv = dinv + j * kTileDoubles;
            double xr = 0;
            bcast_dot<0>(xr, inv, u, r);  // every 16-lane row holds u, so row broadcasts serve
            if (lane < kTile) x[kTile * j + lane] = xr;
            if (j >= 1) {  // (L_{j,j-1}^T x_j)[c]: the lane holds L[r][4 rr + q] (q = lane / 16)
#pragma unroll
                for (int rr = 0; rr < 4; ++rr) {
                    const double v = row16_sum(tcur[rr] * xr);
                    if (r == 0) tterm[4 * rr + lane / 16] = v;
                }
                wave_sync();
                tj = tterm[r];
            }
        } else if (j >= 1) {
            double p[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int s = 0; s < kBackSlots; ++s) {
                const int a = j + wave + s * kBackWaves;
                if (a < T) {
                    const double xa = x[kTile * a + r];
#pragma unroll
                    for (int rr = 0; rr < 4; ++rr) p[rr] = fma(cur[s][rr], xa, p[rr]);
                }
            }
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) {
                p[rr] = row16_sum(p[rr]);
                if (r == 0) part[((j - 1) & 1) * nw * kTile + wave * kTile + 4 * rr + lane / 16] = p[rr];
            }
        }
        load_round(j - 1, nxt, tnxt);  // after this round's work: its loads overlap the barrier
        __syncthreads();
#pragma unroll
        for (int s = 0; s < kBackSlots; ++s) cur[s] = nxt[s];
        tcur = tnxt;
    }
}

// Forward + backward solve with an existing factor (iterative refinement). The forward
// substitution runs tile column by tile column: wave 0 solves y_k with the inverse diagonal
// tile, then the waves update the rows below it (L_ak y_k; in the operand layout the lane holds
// L_ak[r][4 rr + q], r = lane % 16, q = lane / 16).
__device__ void cholesky_solve_again(const double* Lt_, int T, const double* dinv, double* x, double* part) {
    const gdouble* Lt = (const gdouble*)Lt_;
    const int t = threadIdx.x, wave = t / 64, lane = t % 64, nw = blockDim.x / 64;
    for (int k = 0; k < T; ++k) {
        if (wave == 0) {
            const double* inv = dinv + k * kTileDoubles;
            const double u = lane < kTile ? x[kTile * k + lane] : 0.0;
            double xr = 0;
#pragma unroll
            for (int c = 0; c < kTile; ++c) xr = fma(inv[opnd_pos(lane % kTile, c)], lane_bcast(u, c), xr);
            if (lane < kTile) x[kTile * k + lane] = xr;
        }
        __syncthreads();
        for (int a = k + 1 + wave; a < T; a += nw) {
            const int base = tile_base(a, k);
            double v = 0;
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) v = fma(Lt[base + rr * 64 + lane], x[kTile * k + 4 * rr + lane / 16], v);
            v += __shfl_xor(v, 16, 64);
            v += __shfl_xor(v, 32, 64);
            if (lane < kTile) x[kTile * a + lane] -= v;
        }
        __syncthreads();
    }
    cholesky_backward(Lt_, T, dinv, x, part);
}

// Sum over a wave, then over the waves of the block in fixed order (deterministic).
__device__ double block_sum_waves(double v, double* red) {
    const int t = threadIdx.x, wave = t / 64, lane = t % 64, nw = blockDim.x / 64;
    v = wave_reduce_sum(v);
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = 0;
    for (int w = 0; w < nw; ++w) s += red[w];
    __syncthreads();
    return s;
}

__device__ double block_max_waves(double v, double* red) {
    const int t = threadIdx.x, wave = t / 64, lane = t % 64, nw = blockDim.x / 64;
    for (int m = 32; m > 0; m >>= 1) v = fmax(v, __shfl_xor(v, m, 64));
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = red[0];
    for (int w = 1; w < nw; ++w) s = fmax(s, red[w]);
    __syncthreads();
    return s;
}

// The end of a step once the state step x (tangent order, LDS, N entries) is final: clamp to
// the relative bounds (only td is bounded among the states), the state model decrease (here only
// for a window without landmarks; k_backsub spreads it over its chunk blocks otherwise), the
// step norm and the state trial.
__device__ __forceinline__ void step_tail(WinDev& w, double* x, double* red) {
    const int n = w.nr, t = threadIdx.x, nt = blockDim.x, wave = t / 64, lane = t % 64, nw = nt / 64;
    if (t == 0 && w.td_col >= 0 && isfinite(x[w.td_col])) {
        const double tdv = *w.td;
        x[w.td_col] = fmax(w.td_lo - tdv, fmin(w.td_hi - tdv, x[w.td_col]));
    }
    __syncthreads();
    double pred = 0;
    if (w.nchunks == 0)
        for (int i = wave; i < n; i += nw) {
            double h = 0;
            for (int j = lane; j < n; j += 64) h += w.Hfull[(size_t)i * n + j] * x[j];
            h = wave_reduce_sum(h);
            if (lane == 0) pred += -x[i] * (w.gfull[i] + 0.5 * h);
        }
    const double ps = block_sum_waves(pred, red);
    double mx = 0, bad = 0;
    for (int i = t; i < n; i += nt) {
        w.dstate[i] = x[i];
        if (!isfinite(x[i])) bad = 1;
        mx = fmax(mx, fabs(x[i]));
    }
    const double maxinc = block_max_waves(mx, red);
    const double nonfinite = block_max_waves(bad, red);
    // state trial
    for (int k = t; k < w.K; k += nt) {
        const double* xp = w.pose + 7 * k;
        double* yp = w.cpose + 7 * k;
        if (w.pose_col[k] >= 0) pose_plus(xp, x + w.pose_col[k], yp);
        else for (int i = 0; i < 7; ++i) yp[i] = xp[i];
        for (int i = 0; i < 9; ++i) {
            const double v = w.sb[9 * k + i];
            w.csb[9 * k + i] = (w.sb_col[k] >= 0 && i < w.sb_w[k]) ? v + x[w.sb_col[k] + i] : v;
        }
    }
    if (t == 0) {
        if (w.rig_col >= 0) pose_plus(w.rig, x + w.rig_col, w.crig);
        else for (int i = 0; i < 7; ++i) w.crig[i] = w.rig[i];
        *w.ctd = w.td_col >= 0 ? fmax(w.td_lo, fmin(w.td_hi, *w.td + x[w.td_col])) : *w.td;
        w.lm->chol_ok = 1;
        w.lm->predicted = ps;
        w.lm->max_increment = maxinc;
        w.lm->trial_valid = nonfinite > 0 ? 0 : 1;
        w.lm->step_nonfinite = nonfinite > 0 ? 1 : 0;
    }
}

// refinement > 0 and defer == 0: iterative refinement against the saved damped system (Scopy),
// which is the exact operator in FP64. defer != 0 (mixed precision, whose Scopy holds FP32
// Schur products): the unrefined step goes to dstate and k_schur_apply / k_refine refine it
// against the exact operator and finish the step.
__global__ __launch_bounds__(kCholThreads) void k_cholesky(WinDev* wins, int refinement, int defer) {
    __shared__ double pnl[2 * kMaxTiles * kTileDoubles];
    __shared__ double dtile[kMaxTiles * kTileDoubles];
    __shared__ double dcur[2 * kTileDoubles];
    __shared__ double x[kMaxTiles * kTile];
    __shared__ double rres[kMaxTiles * kTile];
    __shared__ double part[kBackPart];
    __shared__ double red[kCholThreads / 64];
    __shared__ int flag;
    WinDev& w = wins[blockIdx.x];
    if (!window_active(w)) return;
    const int n = w.nr, t = threadIdx.x, nt = blockDim.x, wave = t / 64, lane = t % 64, nw = nt / 64;
    const int T = (n + kTile - 1) / kTile, N = kTile * T;
    // factorization order: speed/bias columns first (their IMU chain leaves most of their tiles
    // zero, which the factorization skips), then the visual columns
    const int nv = w.nv, nsb = n - nv;
    auto to_solve = [&](int c) { return c < nv ? c + nsb : c - nv; };
    auto to_tangent = [&](int p) { return p < nsb ? p + nv : p - nsb; };
    if (t == 0) flag = 0;
    for (int i = t; i < N; i += nt) x[i] = i < n ? w.rhs[to_tangent(i)] : 0.0;
    __syncthreads();
    const bool ok = cholesky_forward(w.fac, T, pnl, dtile, dcur, x, &flag, w.chol_mask);
    if (!ok) {
        if (t == 0) w.lm->chol_ok = 0;
        return;
    }
    // inverse diagonal tiles (diagonal slots of the tile store) -> LDS for the substitutions
    double* dinv = pnl;
    for (int p = t; p < T * kTileDoubles; p += nt) {
        const int a = p / kTileDoubles;
        dinv[p] = w.fac[tile_base(a, a) + p % kTileDoubles];
    }
    __syncthreads();
    cholesky_backward(w.fac, T, dinv, x, part);
    if (defer) {
        for (int c = t; c < n; c += nt) w.dstate[c] = x[to_solve(c)];
        if (t == 0) w.lm->chol_ok = 1;
        return;
    }
    for (int it = 0; it < refinement; ++it) {
        for (int i = wave; i < N; i += nw) {  // residual row i of the factorization order
            double v = 0;
            if (i < n) {
                const double* row = w.Scopy + (size_t)to_tangent(i) * n;
                for (int j = lane; j < n; j += 64) v += row[j] * x[to_solve(j)];
            }
            v = wave_reduce_sum(v);
            if (lane == 0) rres[i] = i < n ? w.rhs[to_tangent(i)] - v : 0.0;
        }
        __syncthreads();
        cholesky_solve_again(w.fac, T, dinv, rres, part);
        for (int i = t; i < n; i += nt) x[i] += rres[i];
        __syncthreads();
    }
    // the step back in the tangent layout
    for (int c = t; c < N; c += nt) rres[c] = c < n ? x[to_solve(c)] : 0.0;
    __syncthreads();
    step_tail(w, rres, red);
}

// ---------------------------------------------------------------------------------------------
// Refinement of the state step against the exact reduced operator (mixed precision).
// The reduced residual at x is
//   r = -g - (H + lambda D) x + sum_k c_k (g_k + c_k^T x) / damped(d_k)
// with H, g the undamped FP64 state block and gradient, D = diag(max(H_ii, 1e-6)), and c_k,
// d_k, g_k landmark k's coupling column, Hessian and gradient (VisualCudaBatch.cu:138-172).
// k_schur_apply forms the landmark sum per chunk from the couplings' exact (hi + lo) values,
// in FP64, as a chunk partial over the nv visual columns; k_refine sums the partials in fixed
// order, forms r, solves with the FP64 factor of the mixed system and adds the correction. Each
// pass contracts the step error by the factor's relative error (about 1e-4 on ill-conditioned
// windows), so two passes reach the FP64 step.
// ---------------------------------------------------------------------------------------------
// Workgroups nchunks.. form the state part of the residual, -g - (H + lambda D) x, sixteen rows
// each (hslices per window), into the window's scratch vector.
constexpr int kResidRows = kBlock / 16;
__device__ __forceinline__ void residual_rows(const WinDev& w, int slice) {
    const int n = w.nr, t = threadIdx.x, i = slice * kResidRows + t / 16, l16 = t % 16;
    if (i >= n) return;
    const double* row = w.Hfull + (size_t)i * n;
    const double* x = w.dstate;
    double v = 0;
#pragma unroll 4
    for (int j = l16; j < n; j += 16) v += row[j] * x[j];
#pragma unroll
    for (int m = 8; m > 0; m >>= 1) v += __shfl_xor(v, m, 16);
    if (l16 == 0) w.scratch[i] = -w.gfull[i] - v - w.lm->lambda * fmax(row[i], 1e-6) * x[i];
}

template <bool MIXED>
__global__ __launch_bounds__(kBlock) void k_schur_apply(const WinDev* wins, const Chunk* chunks, int nchunks,
                                                        int hslices) {
    constexpr int kZPitch = kMaxVisual + 1;
    __shared__ double zw[kBlock / 64][kZPitch];  // one accumulator per wave
    __shared__ double xs[kZPitch];
    __shared__ int pcol[kMaxFrames];
    if ((int)blockIdx.x >= nchunks) {
        const int r = blockIdx.x - nchunks;
        const WinDev& w = wins[r / hslices];
        if (window_active(w) && w.lm->chol_ok) residual_rows(w, r % hslices);
        return;
    }
    const Chunk ch = chunks[blockIdx.x];
    const WinDev& w = wins[ch.win];
    if (!window_active(w) || !w.lm->chol_ok) return;
    const int nv = w.nv, t = threadIdx.x, wave = t / 64;
    const double lambda = w.lm->lambda;
    const int lane = t % kChunkLanes, group = t / kChunkLanes;
    constexpr int groups = kBlock / kChunkLanes;
    const size_t L = (size_t)w.L, F = (size_t)w.F;
    for (int j = t; j < kZPitch; j += kBlock) xs[j] = j < nv ? w.dstate[j] : 0.0;
    if (t < kMaxFrames) pcol[t] = t < w.K ? w.pose_col[t] : -1;
    for (int j = t; j < (kBlock / 64) * kZPitch; j += kBlock) (&zw[0][0])[j] = 0.0;
    __syncthreads();
    double* z = zw[wave];
    constexpr int kCached = 2;  // factors per lane kept in registers between the dot and the scatter
    for (int base = ch.lm_begin; base < ch.lm_end; base += groups) {
        const int l = base + group;
        const bool free = l < ch.lm_end && w.lm_col[l] >= 0;
        int a = 0, ac = -1;
        double cd = 0;
        double uc[kCached][6];
        int bcc[kCached];
#pragma unroll
        for (int e = 0; e < kCached; ++e) bcc[e] = -2;  // -2: no factor (same-epoch or absent)
        if (free) {
            a = w.lm_fc[l] >> 8;
            ac = pcol[a];
            int e = 0;
            for (int f = w.lm_off[l] + lane; f < w.lm_off[l + 1]; f += kChunkLanes, ++e) {
                const int b = w.ob_fc[f] >> 8;
                if (b == a) continue;  // same-epoch factors carry no pose coupling
                const int bc = pcol[b], p = w.pf_pos[f];
                double u[6];
#pragma unroll
                for (int k = 0; k < 6; ++k) {
                    u[k] = load_u_exact<MIXED>(w, F, k, p);
                    cd += u[k] * ((ac >= 0 ? xs[ac + k] : 0.0) - (bc >= 0 ? xs[bc + k] : 0.0));
                }
#pragma unroll
                for (int q = 0; q < kCached; ++q)
                    if (q == e) {
                        bcc[q] = bc;
#pragma unroll
                        for (int k = 0; k < 6; ++k) uc[q][k] = u[k];
                    }
            }
        }
        cd = group_sum(cd);
        if (!free) continue;
        if (w.rig_col >= 0)
            for (int k = 0; k < 6; ++k) cd += w.lm_out[(LO_CR0 + k) * L + l] * xs[w.rig_col + k];
        if (w.td_col >= 0) cd += w.lm_out[LO_CT * L + l] * xs[w.td_col];
        const double wk = (w.lm_out[LO_G * L + l] + cd) / damped_point(w.lm_out[LO_D * L + l], lambda);
        // z += c_k w_k on the landmark's support; the lanes of one wave add in a fixed order
#pragma unroll
        for (int q = 0; q < kCached; ++q) {
            if (bcc[q] == -2) continue;
            for (int k = 0; k < 6; ++k) {
                const double v = wk * uc[q][k];
                if (ac >= 0) atomicAdd(&z[ac + k], v);
                if (bcc[q] >= 0) atomicAdd(&z[bcc[q] + k], -v);
            }
        }
        for (int f = w.lm_off[l] + lane + kCached * kChunkLanes; f < w.lm_off[l + 1]; f += kChunkLanes) {
            const int b = w.ob_fc[f] >> 8;
            if (b == a) continue;
            const int bc = pcol[b], p = w.pf_pos[f];
            for (int k = 0; k < 6; ++k) {
                const double v = wk * load_u_exact<MIXED>(w, F, k, p);
                if (ac >= 0) atomicAdd(&z[ac + k], v);
                if (bc >= 0) atomicAdd(&z[bc + k], -v);
            }
        }
        if (lane == 0) {
            if (w.rig_col >= 0)
                for (int k = 0; k < 6; ++k) atomicAdd(&z[w.rig_col + k], wk * w.lm_out[(LO_CR0 + k) * L + l]);
            if (w.td_col >= 0) atomicAdd(&z[w.td_col], wk * w.lm_out[LO_CT * L + l]);
        }
    }
    __syncthreads();
    if (t < nv) {
        double v = 0;
#pragma unroll
        for (int q = 0; q < kBlock / 64; ++q) v += zw[q][t];
        w.schur_part[(size_t)ch.local * (w.tri + nv) + t] = v;
    }
}

// last != 0: the final pass, which also finishes the step (clamp, model decrease, trial).
__global__ __launch_bounds__(kCholThreads) void k_refine(WinDev* wins, int last) {
    __shared__ double dinv[kMaxTiles * kTileDoubles];
    __shared__ double x[kMaxTiles * kTile];
    __shared__ double rres[kMaxTiles * kTile];
    __shared__ double zp[kCholThreads];
    __shared__ double part[kBackPart];
    __shared__ double red[kCholThreads / 64];
    WinDev& w = wins[blockIdx.x];
    if (!window_active(w) || !w.lm->chol_ok) return;
    const int n = w.nr, t = threadIdx.x, nt = blockDim.x;
    const int T = (n + kTile - 1) / kTile, N = kTile * T;
    const int nv = w.nv, nsb = n - nv, len = w.tri + nv;
    auto to_solve = [&](int c) { return c < nv ? c + nsb : c - nv; };
    for (int p = t; p < T * kTileDoubles; p += nt) {
        const int a = p / kTileDoubles;
        dinv[p] = w.fac[tile_base(a, a) + p % kTileDoubles];
    }
    for (int c = t; c < N; c += nt) x[c] = c < n ? w.dstate[c] : 0.0;
    // landmark sum z over the chunk partials: eight chunk lanes per column, each adding its
    // partials in chunk order; a lane issues the loads of up to kZBatch partials at once
    // (absent ones read as zero, which leaves the sum unchanged)
    {
        constexpr int kZBatch = 24;
        const int j = t / 8, q = t % 8;
        double v = 0;
        if (j < nv) {
            const double* src = w.schur_part + j;
            const int nch = w.nchunks;
            for (int c0 = q; c0 < nch; c0 += 8 * kZBatch) {
                double b[kZBatch];
#pragma unroll
                for (int e = 0; e < kZBatch; ++e) {
                    const int c = c0 + 8 * e;
                    b[e] = c < nch ? src[(size_t)c * len] : 0.0;
                }
#pragma unroll
                for (int e = 0; e < kZBatch; ++e) v += b[e];
            }
        }
        zp[t] = v;
    }
    __syncthreads();
    // reduced residual: the state part from k_schur_apply's row workgroups, plus z
    for (int i = t; i < n; i += nt) {
        double z = 0;
        if (i < nv)
#pragma unroll
            for (int q = 0; q < 8; ++q) z += zp[8 * i + q];
        rres[to_solve(i)] = w.scratch[i] + z;
    }
    for (int i = n + t; i < N; i += nt) rres[i] = 0.0;
    __syncthreads();
    cholesky_solve_again(w.fac, T, dinv, rres, part);
    for (int c = t; c < n; c += nt) x[c] += rres[to_solve(c)];
    __syncthreads();
    if (!last) {
        for (int c = t; c < n; c += nt) w.dstate[c] = x[c];
        return;
    }
    step_tail(w, x, red);
}

// ---------------------------------------------------------------------------------------------
// k_sym_eigen: A = Q diag(ev) Q^T for a batch of symmetric matrices, one workgroup per job
// (EigenJob: n x n, row-major, leading dimension ld; A is overwritten; the slides of many sessions
// share one launch). Q is returned transposed in V (row c = eigenvector c),
// ev unsorted. Three phases (Golub & Van Loan, Householder tridiagonalization and the symmetric
// QR step with implicit Wilkinson shift):
//   1. n-2 Householder reflections H_k = I - 2 v v^T reduce A to tridiagonal (d, e); each step
//      is a column norm, a matrix-vector product (one wave per row) and a rank-2 update,
//      two barriers; v is kept in row k of A's upper triangle;
//   2. Q^T = H_{n-3} ... H_0 accumulated into V;
//   3. implicit QR sweeps on the unreduced trailing block: one wave runs up to 64 sweeps on the
//      tridiagonal alone (deflation scans 64 entries at a time, one lane chases the bulge) and
//      logs their Givens rotations (global memory), then every thread applies the logged
//      rotations to its column of V^T (rows k, k+1 of V, coalesced); converged off-diagonals
//      (|e| <= eps (|d_k|+|d_k+1|)) deflate; at most 30 n sweeps.
// Serves the marginalization's pseudo-inverse of the eliminated state block and the prior's
// eigen cut and square root (Marginalizer::psdSolve / compute, Marginalizer.cpp:240-283,
// 430-466, whose GPU form is MarginalCudaWorkspace::eigen, VisualCudaBatch.cu:887-957).
// ---------------------------------------------------------------------------------------------
constexpr int kEigMax = 512;
constexpr int kEigThreads = 1024;
// NL > 0: the matrix, then the eigenvectors, are staged in the workgroup's LDS (n <= NL); every
// step of the three phases then reads and writes LDS instead of global memory. NL == 0: both
// stay in global memory (any n <= kEigMax).
constexpr int kEigLogSweeps = 64;  // QR sweeps logged per round of the rotation pass
// Doubles of the rotation log of an n x n problem (k_sym_eigen's `rot`).
__host__ __device__ inline size_t eigen_log_doubles(int n) { return 2 * (size_t)kEigLogSweeps * n + kEigLogSweeps; }
typedef __attribute__((address_space(3))) double ldouble;  // LDS
typedef __attribute__((address_space(3))) int lint;
// the matrix, then the eigenvectors, in LDS (NL > 0) or in global memory
template <int NL> struct EigenStore { typedef ldouble type; };
template <> struct EigenStore<0> { typedef VIB_G double type; };
template <int NL>
__device__ void sym_eigen_body(VIB_G double* A, VIB_G double* V, VIB_G double* ev, int n, int ld, VIB_G int* ok,
                               VIB_G double* rot, ldouble* lds, ldouble* d, ldouble* e, ldouble* q, lint& lo, lint& hi,
                               lint& state) {
    const int t = threadIdx.x, nt = blockDim.x, wave = t / 64, lane = t % 64, nw = nt / 64;
    typedef typename EigenStore<NL>::type ET;
    ET* Ap;
    if constexpr (NL > 0) Ap = lds;
    else Ap = A;
    const int lda = NL ? n : ld;
    if (NL) {
        for (int p = t; p < n * n; p += nt) Ap[p] = A[(size_t)(p / n) * ld + p % n];
        __syncthreads();
    }
    auto a = [&](int r, int c) -> ET& { return Ap[(size_t)r * lda + c]; };
    // 1. tridiagonalization. Every wave forms the reflector's scalars itself from column k (no
    //    barrier); v_i = (a(k+1+i, k) - alpha [i == 0]) / |x - alpha e1| is read from column k on
    //    the fly. Row k's upper part (dead after step k) receives v for phase 2.
    for (int k = 0; k + 2 < n; ++k) {
        const int m = n - k - 1;  // length of the reflected part (rows k+1 .. n-1)
        double ss = 0;
        for (int i = lane; i < m; i += 64) { const double x = a(k + 1 + i, k); ss += x * x; }
        ss = wave_sum_dpp(ss);
        const double x0 = a(k + 1, k), nx = sqrt(ss);
        const double alpha = x0 > 0 ? -nx : nx;
        // |x - alpha e1|^2 = 2 (ss - alpha x0)
        const double vn2 = 2 * (ss - alpha * x0);
        const double iv = vn2 > 0 ? 1 / sqrt(vn2) : 0.0;
        auto v = [&](int i) { return (a(k + 1 + i, k) - (i == 0 ? alpha : 0.0)) * iv; };
        // p = A22 v: thread i runs down column i of the symmetric A22 (the lanes of a wave read
        // consecutive entries of one row)
        for (int i = t; i < m; i += nt) {
            double s = 0;
#pragma unroll 8
            for (int j = 0; j < m; ++j) s += a(k + 1 + j, k + 1 + i) * v(j);
            q[i] = s;
        }
        __syncthreads();
        double kk = 0;
        for (int i = lane; i < m; i += 64) kk += v(i) * q[i];
        kk = wave_sum_dpp(kk);
        // A22 -= 2 (v w^T + w v^T), w = p - (v^T p) v
        for (int p = t; p < m * m; p += nt) {
            const int i = p / m, j = p % m;
            const double vi = v(i), vj = v(j);
            a(k + 1 + i, k + 1 + j) -= 2 * (vi * (q[j] - kk * vj) + (q[i] - kk * vi) * vj);
        }
        __syncthreads();
        // no reader of column k or row k is left in this step or the next ones
        if (t == 0) { d[k] = a(k, k); e[k] = vn2 > 0 ? alpha : x0; }
        for (int i = t; i < m; i += nt) a(k, k + 1 + i) = v(i);
    }
    if (t == 0) {
        if (n >= 2) { d[n - 2] = a(n - 2, n - 2); e[n - 2] = a(n - 1, n - 2); }
        d[n - 1] = a(n - 1, n - 1);
        e[n - 1] = 0;
    }
    __syncthreads();
    if (NL) {  // the reflectors (row k, columns k+1 ..) go back to global memory: LDS takes V next
        for (int p = t; p < n * n; p += nt) {
            const int r = p / n, c = p % n;
            if (c > r) A[(size_t)r * ld + c] = Ap[p];
        }
        __syncthreads();
    }
    ET* Vp;
    if constexpr (NL > 0) Vp = lds;
    else Vp = V;
    const int ldv = NL ? n : ld;
    // 2. Q^T = H_{n-3} ... H_0 in V: start from I, apply H_k on the right for k = n-3 .. 0
    //    (Q^T <- Q^T H_k acts on columns k+1 ..; H_0 last). Before H_k the product acts on rows
    //    and columns >= k+2 only, so rows <= k have zeros in columns > k and H_k changes rows
    //    k+1 .. n-1 alone.
    for (int i = t; i < n * n; i += nt) Vp[(size_t)(i / n) * ldv + i % n] = (i / n == i % n) ? 1.0 : 0.0;
    __syncthreads();
    for (int k = n - 3; k >= 0; --k) {
        const int m = n - k - 1;
        const VIB_G double* vk = A + (size_t)k * ld + k + 1;  // the reflector of step k
        for (int rr = wave; rr < m; rr += nw) {
            const ET* row = Vp + (size_t)(k + 1 + rr) * ldv + k + 1;
            double s = 0;
            for (int i = lane; i < m; i += 64) s += row[i] * vk[i];
            s = wave_sum_dpp(s);
            if (lane == 0) q[rr] = s;
        }
        __syncthreads();
        for (int p = t; p < m * m; p += nt) {
            const int rr = p / m, i = p % m;
            Vp[(size_t)(k + 1 + rr) * ldv + k + 1 + i] -= 2 * q[rr] * vk[i];
        }
        __syncthreads();
    }
    // 3. implicit symmetric QR on (d, e); V holds Q^T: a rotation G on (k, k+1) replaces rows
    //    k, k+1 of Q^T by (c row_k - s row_k+1, s row_k + c row_k+1). Wave 0 runs up to
    //    kEigLogSweeps sweeps on (d, e) alone, logging each sweep's block [l, m] and rotations in
    //    `rot`; then every thread applies the logged rotations, in order, to its column of V^T.
    //    Repeated until every off-diagonal has deflated (at most 30 n sweeps).
    VIB_G int2* span = reinterpret_cast<VIB_G int2*>(rot + 2 * (size_t)kEigLogSweeps * n);
    if (t == 0) { state = 0; hi = n - 1; lo = 0; }
    __syncthreads();
    for (int done = 0; done < 30 * n;) {
        if (wave == 0) {
            const double eps = 2.220446049250313e-16;
            auto small = [&](int i) { return fabs(e[i]) <= eps * (fabs(d[i]) + fabs(d[i + 1])); };
            int count = 0;
            while (count < kEigLogSweeps && done + count < 30 * n) {
                // deflation at the bottom: the trailing off-diagonals that are small are set to
                // zero; m = the last row of the unreduced block (64 entries per wave step)
                int m = hi;
                while (m > 0) {
                    const int i = m - 1 - lane;
                    const bool keep = i >= 0 && !small(i);
                    const unsigned long long b = __ballot(keep);
                    const int top = b ? m - (__ffsll(b) - 1) : m - 64;  // rows > top - 1 deflated
                    if (i >= 0 && i >= top) e[i] = 0;
                    m = b ? top : (m > 64 ? m - 64 : 0);
                    if (b) break;
                }
                if (lane == 0) hi = m;
                if (m == 0) {
                    if (lane == 0) state = 1;
                    break;
                }
                // l = first row of the block: the nearest small off-diagonal above m ends it
                int l = 0;
                for (int base = m - 2; base >= 0; base -= 64) {
                    const int i = base - lane;
                    const unsigned long long b = __ballot(i >= 0 && small(i));
                    if (b) {
                        l = base - (__ffsll(b) - 1) + 1;
                        break;
                    }
                }
                wave_sync();
                if (lane == 0) {
                    // Wilkinson shift from the trailing 2x2 block
                    const double dl = 0.5 * (d[m - 1] - d[m]), em = e[m - 1];
                    const double mu = d[m] - em * em / (dl + (dl >= 0 ? 1.0 : -1.0) * sqrt(dl * dl + em * em));
                    // the bulge chase; the running diagonal and off-diagonal entries stay in
                    // registers (ak = d[k], bk = e[k] as updated by the previous rotation), so a
                    // step loads only the untouched d[k+1], e[k+1] (e[m] is read and ignored)
                    VIB_G double* cs = rot + 2 * (size_t)count * n;
                    double ak = d[l], bk = e[l];
                    double x = ak - mu, z = bk;
                    for (int k = l; k < m; ++k) {
                        const double ck = d[k + 1];
                        const double en = e[k + 1];
                        const double rn = k + 1 < m ? en : 0.0;
                        const double h2 = x * x + z * z;
                        // 1 / sqrt(h2): the hardware estimate and two Newton steps, relative error
                        // a few ulp (h2 in the normal range), else the correctly rounded form
                        const double ih = h2 > 1e-280 && h2 < 1e280 ? rsqrt_nr(h2)
                                                                     : (h2 > 0 ? 1.0 / sqrt(h2) : 0.0);
                        const double c = h2 > 0 ? x * ih : 1.0, sn = h2 > 0 ? -z * ih : 0.0;
                        if (k > l) e[k - 1] = c * x - sn * z;  // (x, z) = (e[k-1], bulge)
                        const double dk = c * c * ak - 2 * c * sn * bk + sn * sn * ck;
                        const double dk1 = sn * sn * ak + 2 * c * sn * bk + c * c * ck;
                        const double ek = c * sn * (ak - ck) + (c * c - sn * sn) * bk;
                        d[k] = dk;
                        cs[2 * k] = c;
                        cs[2 * k + 1] = sn;
                        x = ek;
                        z = -sn * rn;  // the bulge
                        ak = dk1;
                        bk = c * rn;
                    }
                    d[m] = ak;
                    e[m - 1] = x;
                    span[count] = make_int2(l, m);
                }
                wave_sync();
                ++count;
            }
            if (lane == 0) lo = count;  // sweeps logged in this round
        }
        __syncthreads();
        const int count = lo;
        for (int r = t; r < n; r += nt)
            for (int sw = 0; sw < count; ++sw) {
                const int2 lm = span[sw];
                const VIB_G double* cs = rot + 2 * (size_t)sw * n;
                double x = Vp[(size_t)lm.x * ldv + r];
                int k = lm.x;
                // groups of 8 rotations: rows k+1 .. k+8 still hold their values from before
                // this sweep, so their loads (and the 8 rotations') are issued together
                for (; k + 8 <= lm.y; k += 8) {
                    double y[8], c[8], sn[8];
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        y[j] = Vp[(size_t)(k + 1 + j) * ldv + r];
                        c[j] = cs[2 * (k + j)];
                        sn[j] = cs[2 * (k + j) + 1];
                    }
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        Vp[(size_t)(k + j) * ldv + r] = c[j] * x - sn[j] * y[j];
                        x = sn[j] * x + c[j] * y[j];
                    }
                }
                for (; k < lm.y; ++k) {
                    const double y = Vp[(size_t)(k + 1) * ldv + r];
                    const double c = cs[2 * k], sn = cs[2 * k + 1];
                    Vp[(size_t)k * ldv + r] = c * x - sn * y;
                    x = sn * x + c * y;
                }
                Vp[(size_t)lm.y * ldv + r] = x;
            }
        __syncthreads();
        done += count;
        if (state) break;
        __syncthreads();
    }
    if (NL)
        for (int p = t; p < n * n; p += nt) V[(size_t)(p / n) * ld + p % n] = Vp[p];
    for (int i = t; i < n; i += nt) ev[i] = d[i];
    if (t == 0) *ok = state;
}

__global__ __launch_bounds__(kEigThreads) void k_sym_eigen(const EigenJob* jobs) {
    __shared__ double d[kEigMax], e[kEigMax], q[kEigMax];
    __shared__ int lo, hi, state;
    const EigenJob j = jobs[blockIdx.x];
    sym_eigen_body<0>(j.A, j.V, j.ev, j.n, j.ld, j.ok, j.rot, nullptr, (ldouble*)d, (ldouble*)e, (ldouble*)q,
                      *(lint*)&lo, *(lint*)&hi, *(lint*)&state);
}

// Matrices of up to kEigLds rows (every prior of a window of at most 20 frames: 19 poses, one
// speed/bias block, rig and time offset give 130) in LDS: 148 KB for the matrix, 3.3 KB of
// vectors.
constexpr int kEigLds = 136;
__global__ __launch_bounds__(kEigThreads) void k_sym_eigen_lds(const EigenJob* jobs) {
    __shared__ double lds[kEigLds * kEigLds];
    __shared__ double d[kEigLds], e[kEigLds], q[kEigLds];
    __shared__ int lo, hi, state;
    const EigenJob j = jobs[blockIdx.x];
    sym_eigen_body<kEigLds>(j.A, j.V, j.ev, j.n, j.ld, j.ok, j.rot, (ldouble*)lds, (ldouble*)d, (ldouble*)e,
                            (ldouble*)q, *(lint*)&lo, *(lint*)&hi, *(lint*)&state);
}

// ---------------------------------------------------------------------------------------------
// k_backsub: delta_k = -(g_k + c_k^T dx) / damped(d_k), clamped; model decrease; rho trial
// (VisualCudaBatch.cu:191-205).
// ---------------------------------------------------------------------------------------------
template <bool MIXED>
__global__ __launch_bounds__(kBlock) void k_backsub(const WinDev* wins, const Chunk* chunks) {
    __shared__ double red[3 * kBlock];
    __shared__ double dxv[kMaxVisual];  // the visual columns of the state step
    __shared__ int pcol[kMaxFrames];
    const Chunk ch = chunks[blockIdx.x];
    const WinDev& w = wins[ch.win];
    if (!window_active(w)) return;
    const double lambda = w.lm->lambda;
    const int lane = threadIdx.x % kChunkLanes, group = threadIdx.x / kChunkLanes;
    constexpr int groups = kBlock / kChunkLanes;
    const size_t L = (size_t)w.L, F = (size_t)w.F;
    const double* dx = w.dstate;
    // the step's visual columns and the pose columns in LDS: the coupling sums index them per
    // factor, one dependent global load fewer each
    for (int j = threadIdx.x; j < w.nv; j += kBlock) dxv[j] = dx[j];
    if (threadIdx.x < kMaxFrames) pcol[threadIdx.x] = (int)threadIdx.x < w.K ? w.pose_col[threadIdx.x] : -1;
    __syncthreads();
    double pred = 0, mx = 0;
    int bad = 0;
    // a landmark's records; the next pass's are loaded before this pass chases its factors
    struct LmIn {
        int col, fc, f0, f1;
        double rho, dd, gk;
    };
    auto load_lm = [&](int l) {
        LmIn in{-1, 0, 0, 0, 0.0, 0.0, 0.0};
        if (l < ch.lm_end) {
            in.col = w.lm_col[l];
            in.fc = w.lm_fc[l];
            in.f0 = w.lm_off[l];
            in.f1 = w.lm_off[l + 1];
            in.rho = w.rho[l];
            in.dd = w.lm_out[LO_D * L + l];
            in.gk = w.lm_out[LO_G * L + l];
        }
        return in;
    };
    LmIn next = load_lm(ch.lm_begin + group);
    for (int base = ch.lm_begin; base < ch.lm_end; base += groups) {
        const int l = base + group;
        const LmIn cur = next;
        next = load_lm(l + groups);
        const bool live = l < ch.lm_end;
        const bool free = live && cur.col >= 0;
        double cd = 0;
        if (free) {
            const int a = cur.fc >> 8;
            const int ac = pcol[a];
            for (int f = cur.f0 + lane; f < cur.f1; f += kChunkLanes) {
                const int b = w.ob_fc[f] >> 8;
                if (b == a) continue;
                const int bc = pcol[b];
                const int p = w.pf_pos[f];
                for (int k = 0; k < 6; ++k) {
                    const double u = load_u_exact<MIXED>(w, F, k, p);
                    cd += u * ((ac >= 0 ? dxv[ac + k] : 0.0) - (bc >= 0 ? dxv[bc + k] : 0.0));
                }
            }
        }
        cd = group_sum(cd);
        if (live && lane == 0) {
            const double rho = cur.rho;
            if (free) {
                if (w.rig_col >= 0)
                    for (int k = 0; k < 6; ++k) cd += w.lm_out[(LO_CR0 + k) * L + l] * dxv[w.rig_col + k];
                if (w.td_col >= 0) cd += w.lm_out[LO_CT * L + l] * dxv[w.td_col];
                const double dd = cur.dd, gk = cur.gk;
                const double d = -(gk + cd) / damped_point(dd, lambda);
                const double delta = isfinite(d) ? fmax(w.min_rho - rho, d) : d;
                pred += -delta * (gk + cd + 0.5 * dd * delta);
                if (!isfinite(delta)) bad = 1;
                mx = fmax(mx, fabs(delta));
                w.lm_delta[l] = delta;
                w.crho[l] = fmax(w.min_rho, rho + delta);
            } else {
                w.lm_delta[l] = 0;
                w.crho[l] = rho;
            }
        }
    }
    // this block's share of the state rows of -dx^T (g + 0.5 H dx), undamped H, one row per wave
    {
        const int n = w.nr, wave = threadIdx.x / 64, l64 = threadIdx.x % 64;
        const int r0 = (int)((long)ch.local * n / w.nchunks), r1 = (int)((long)(ch.local + 1) * n / w.nchunks);
        for (int i = r0 + wave; i < r1; i += kBlock / 64) {
            const double* row = w.Hfull + (size_t)i * n;
            double h = 0;
            for (int j = l64; j < n; j += 64) h += row[j] * dx[j];
            h = wave_reduce_sum(h);
            if (l64 == 0) pred += -dx[i] * (w.gfull[i] + 0.5 * h);
        }
    }
    double sums[3] = {pred, mx, bad ? 1.0 : 0.0};
    block_reduce_n<3, 0b110u>(sums, red);  // maxima: mx, bad
    const double ps = sums[0], m = sums[1], b = sums[2];
    if (threadIdx.x == 0) {
        double* sc = w.chunk_sc + (size_t)ch.local * CS_PLANES;
        sc[CS_PRED] = ps;
        sc[CS_MAXINC] = m;
        sc[CS_TINVALID] = b;
        sc[CS_NONFINITE] = b;
    }
}

// ---------------------------------------------------------------------------------------------
// k_trial_cost: robust visual cost at the trial state (VisualCudaBatch.cu:715-741), one factor
// per lane over the pair-ordered visual work list: every lane evaluates a factor, whatever the
// landmarks' observation counts (in landmark order, 8 lanes per landmark idle where a landmark
// has fewer observations). Per-range partials go to lin_sc; k_lm_update sums them in list order.
// ---------------------------------------------------------------------------------------------
// FUSED (small batches): workgroups 0 .. rb-1 evaluate the retained items at the trial state
// first (items / per as in the linearization), so the retained trial cost needs no launch of its
// own; rb is a multiple of the XCD count like there.
template <bool FUSED>
__global__ __launch_bounds__(kBlock) void k_trial_cost(const WinDev* wins, const Chunk* lin, int rb, int items, int per) {
    if constexpr (FUSED) {
        if ((int)blockIdx.x < rb) {
            __shared__ double rscratch[kWaves][kRetainedScratch];
            retained_items<1>(wins, 0, items, per, rscratch[threadIdx.x / 64]);
            return;
        }
    }
    __shared__ FrameTable ft;
    __shared__ double red[2 * kBlock];
    const Chunk ch = lin[blockIdx.x - rb];
    if (ch.local < 0) return;  // an empty position of the XCD-ordered work list
    const WinDev& w = wins[ch.win];
    if (!window_active(w)) return;
    build_frame_table(w, w.cpose, w.crig, ft);
    __syncthreads();
    const double td = *w.ctd;
    const size_t F = (size_t)w.F;
    double cost = 0;
    int invalid = 0;
    // the next factor's records are loaded before this one is evaluated
    struct TrialIn {
        ObsData ob;
        V3<double> raw, vel;
        double tref, rho;
        int ids;
    };
    // (the inverse range through the factor's landmark index, which is read one factor earlier
    // still, so the gather does not wait on the index load)
    auto load_in = [&](int q, int lmi) {
        TrialIn in{};
        in.ob = load_obs(w.pf_ob, F, q);
        in.raw = {w.pf_an[LM_BX * F + q], w.pf_an[LM_BY * F + q], w.pf_an[LM_BZ * F + q]};
        in.vel = {w.pf_an[LM_VX * F + q], w.pf_an[LM_VY * F + q], w.pf_an[LM_VZ * F + q]};
        in.tref = w.pf_an[LM_T * F + q];
        in.rho = w.crho[lmi];
        in.ids = w.pf_ids[q];
        return in;
    };
    int p = ch.lm_begin + (int)threadIdx.x;
    int lmi_next = p + kBlock < ch.lm_end ? w.pf_lm[p + kBlock] : 0;
    TrialIn next = p < ch.lm_end ? load_in(p, w.pf_lm[p]) : TrialIn{};
    for (; p < ch.lm_end; p += kBlock) {
        const TrialIn cur = next;
        if (p + kBlock < ch.lm_end) {
            next = load_in(p + kBlock, lmi_next);
            lmi_next = p + 2 * kBlock < ch.lm_end ? w.pf_lm[p + 2 * kBlock] : 0;
        }
        const int ids = cur.ids;
        LmData lm;
        lm.raw = cur.raw;
        lm.vel = cur.vel;
        lm.tref = cur.tref;
        lm.rho = cur.rho;
        lm.frame = ids >> 24;
        lm.cam = (ids >> 16) & 255;
        FactorJ J;
        if (!eval_factor<false>(w, ft, lm, cur.ob, ids & 0xffff, td, false, J)) { invalid = 1; continue; }
        const double sq = J.r[0] * J.r[0] + J.r[1] * J.r[1];
        if (!isfinite(sq)) { invalid = 1; continue; }
        cost += huber_cost(w.huber, sq);
    }
    double sums[2] = {cost, invalid ? 1.0 : 0.0};
    block_reduce_n<2, 0b10u>(sums, red);
    const double c = sums[0], b = sums[1];
    if (threadIdx.x == 0) {
        double* sc = w.lin_sc + (size_t)ch.local * CS_PLANES;
        sc[CS_TRIAL] = c;
        sc[CS_TINVALID] = b;
    }
}

// ---------------------------------------------------------------------------------------------
// k_lm_update: retained trial cost, then the native LM decision (BackendOptimizer.cpp:232-306).
// mode 0: LM step; 1: evaluate only (metrics, no state change); 2: accept the trial state.
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_lm_update(WinDev* wins, int mode) {
    __shared__ double red[5 * kBlock];
    WinDev& w = wins[blockIdx.x];
    LmState* lm = w.lm;
    const int t = threadIdx.x;
    if (mode != 2 && lm->status != 0) {
        if (t == 0) lm->accept_pending = 0;
        return;
    }
    if (mode == 2) {
        if (t == 0) lm->accept_pending = 1;
    } else {
        // per-chunk and retained scalars of the trial (fixed-order block sums)
        const int E = w.E;
        double pred = 0, trial = 0, mx = 0, tinv = 0, nonfinite = 0;
        for (int e = t; e < E; e += blockDim.x) {
            trial += w.ret_sc[E + e];
            tinv = fmax(tinv, w.ret_sc[3 * E + e]);
        }
        for (int k = t; k < w.nchunks; k += blockDim.x) {  // back-substitution (landmark chunks)
            const double* sc = w.chunk_sc + (size_t)k * CS_PLANES;
            pred += sc[CS_PRED];
            mx = fmax(mx, sc[CS_MAXINC]);
            tinv = fmax(tinv, sc[CS_TINVALID]);
            nonfinite = fmax(nonfinite, sc[CS_NONFINITE]);
        }
        for (int k = t; k < w.nlin; k += blockDim.x) {  // trial cost (visual factor ranges)
            const double* sc = w.lin_sc + (size_t)k * CS_PLANES;
            trial += sc[CS_TRIAL];
            tinv = fmax(tinv, sc[CS_TINVALID]);
        }
        double sums[5] = {pred, nonfinite, trial, mx, tinv};
        block_reduce_n<5, 0b11010u>(sums, red);  // maxima: nonfinite, mx, tinv
        pred = sums[0];
        nonfinite = sums[1];
        trial = sums[2];
        mx = sums[3];
        tinv = sums[4];
        if (t == 0) {
            int accept_flag = 0;
            pred += lm->predicted;
            mx = fmax(mx, lm->max_increment);
            trial += w.ret_sc[4 * E + 1];
            const bool trial_valid = !(tinv > 0) && !(w.ret_sc[4 * E + 3] > 0) && lm->trial_valid && isfinite(trial);
            const int i = lm->iteration;
            int status = 0;
            if (lm->new_point) {
                if (!lm->lin_valid) status = -3;
                else {
                    const double lc = lm->lin_cost;
                    if (!isfinite(lc) || (i > 0 && mode == 0 && fabs(lc - lm->cost) > 1e-8 * fmax(1.0, lm->cost))) status = -5;
                    else if (i == 0 || mode == 1) { lm->cost = lc; if (i == 0) lm->initial_cost = lc; }
                }
            }
            if (status == 0 && !lm->chol_ok) status = -4;
            if (nonfinite > 0) lm->step_nonfinite = 1;
            lm->predicted = pred;
            lm->trial_cost = trial;
            lm->max_increment = mx;
            lm->trial_valid = trial_valid;
            if (mode == 1) {
                lm->status = status;
            } else if (status != 0) {
                lm->status = status;
            } else {
                if (trial_valid && mx < 1e-10) {
                    lm->converged = 1;
                    lm->status = 1;
                } else {
                    double gain = -1;
                    if (trial_valid && pred > 0 && isfinite(pred)) gain = (lm->cost - trial) / pred;
                    lm->last_gain = gain;
                    if (gain > 1e-3) {
                        accept_flag = 1;
                        lm->accepted += 1;
                        const bool conv = lm->cost - trial <= 1e-6 * fmax(1.0, lm->cost);
                        lm->cost = trial;
                        const double c = 2 * gain - 1;
                        lm->lambda = fmax(1e-12, lm->lambda * fmax(1.0 / 3, 1 - c * c * c));
                        lm->increase = 2;
                        lm->new_point = 1;
                        if (conv) { lm->converged = 1; lm->status = 1; }
                    } else {
                        lm->lambda *= lm->increase;
                        lm->increase *= 2;
                        lm->new_point = 0;
                    }
                }
                lm->iteration = i + 1;
                if (lm->continuous) {  // relinearize every iteration, stop only at the budget
                    if (lm->status >= 0) lm->status = lm->iteration >= lm->max_iterations ? 1 : 0;
                    lm->new_point = 1;
                } else {
                    if (lm->status == 0 && lm->iteration >= lm->max_iterations) lm->status = 1;
                    if (lm->status == 1 && !lm->accepted && !lm->converged) lm->status = -6;
                }
            }
            lm->accept_pending = accept_flag;
        }
    }
}

// ---------------------------------------------------------------------------------------------
// k_accept: adopt the trial state after an accepted step (acceptResident, VisualCudaBatch.cu:702);
// slice 0 copies the states, every slice 4096 inverse depths.
// ---------------------------------------------------------------------------------------------
constexpr int kAcceptSlice = 4096;

__global__ __launch_bounds__(kBlock) void k_accept(WinDev* wins) {
    WinDev& w = wins[blockIdx.x];
    if (!w.lm->accept_pending) return;
    const int t = threadIdx.x;
    const int l0 = blockIdx.y * kAcceptSlice;
    if (l0 >= w.L && blockIdx.y > 0) return;
    if (blockIdx.y == 0) {
        for (int i = t; i < 7 * w.K; i += blockDim.x) w.pose[i] = w.cpose[i];
        for (int i = t; i < 9 * w.K; i += blockDim.x) w.sb[i] = w.csb[i];
        if (t < 7) w.rig[t] = w.crig[t];
        if (t == 0) *w.td = *w.ctd;
    }
    const int l1 = min(w.L, l0 + kAcceptSlice);
    for (int i = l0 + t; i < l1; i += blockDim.x) w.rho[i] = w.crho[i];
}


// ---------------------------------------------------------------------------------------------
// Prior Gram matrix G = A^T A (cols x cols) after a prior upload: one 256-thread workgroup per
// 16 x 16 tile of G (blockIdx.x = tile row * tiles + tile column, blockIdx.y = window), one
// output per thread. The two 16-column strips of A pass through LDS 16 rows at a time; each
// output sums its rows in row order.
// ---------------------------------------------------------------------------------------------
constexpr int kGramTile = 16;

__global__ __launch_bounds__(kBlock) void k_prior_gram(const WinDev* wins, double* const* G, int tiles) {
    __shared__ double a[kGramTile][kGramTile + 1], b[kGramTile][kGramTile + 1];
    const WinDev& w = wins[blockIdx.y];
    const int m = w.prior_rows, n = w.prior_cols;
    const int ti = blockIdx.x / tiles, tj = blockIdx.x % tiles;
    const int i0 = ti * kGramTile, j0 = tj * kGramTile;
    if (i0 >= n || j0 >= n) return;  // uniform over the workgroup
    const int t = threadIdx.x, r16 = t / kGramTile, c16 = t % kGramTile;
    const int i = i0 + r16, j = j0 + c16;
    double s = 0;
    for (int r0 = 0; r0 < m; r0 += kGramTile) {
        const int r = r0 + r16;
        a[r16][c16] = (r < m && i0 + c16 < n) ? w.pA[(size_t)r * n + i0 + c16] : 0.0;
        b[r16][c16] = (r < m && j0 + c16 < n) ? w.pA[(size_t)r * n + j0 + c16] : 0.0;
        __syncthreads();
        const int rn = min(kGramTile, m - r0);
        for (int q = 0; q < rn; ++q) s += a[q][r16] * b[q][c16];
        __syncthreads();
    }
    if (i < n && j < n) G[blockIdx.y][(size_t)i * n + j] = s;
}


// ---------------------------------------------------------------------------------------------
// Statistical gate (BackendOptimizer::statisticalGate, BackendOptimizer.cpp:394-472): the
// whitened visual residual before the robust loss, energy = |r|^2, at the current state.
// decision: 0 admitted, 1 over the gate, 2 evaluation failed / nonfinite (the reference aborts
// the solve), 3 not participating (a factor rejected earlier: zero square-root information).
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ bool factor_rejected(const double* ob, size_t F, int f) {
    return ob[OB_S00 * F + f] == 0.0 && ob[OB_S10 * F + f] == 0.0 && ob[OB_S11 * F + f] == 0.0;
}

__global__ __launch_bounds__(kBlock) void k_nis_gate(const WinDev* wins, const Chunk* chunks, double threshold,
                                                     double* energy, uint8_t* decision) {
    __shared__ FrameTable ft;
    const Chunk ch = chunks[blockIdx.x];
    const WinDev& w = wins[ch.win];
    build_frame_table(w, w.pose, w.rig, ft);
    __syncthreads();
    const double td = *w.td;
    const size_t F = (size_t)w.F;
    const int lane = threadIdx.x % kChunkLanes, group = threadIdx.x / kChunkLanes;
    constexpr int groups = kBlock / kChunkLanes;
    for (int base = ch.lm_begin; base < ch.lm_end; base += groups) {
        const int l = base + group;
        if (l >= ch.lm_end) continue;
        const LmData lm = load_landmark(w, w.rho, l);
        for (int f = w.lm_off[l] + lane; f < w.lm_off[l + 1]; f += kChunkLanes) {
            uint8_t d = 3;
            double sq = 0;
            if (!factor_rejected(w.ob, F, f)) {
                FactorJ J;
                if (!eval_factor<false>(w, ft, lm, load_obs(w.ob, F, f), w.ob_fc[f], td, false, J)) {
                    d = 2;
                } else {
                    sq = J.r[0] * J.r[0] + J.r[1] * J.r[1];
                    d = !isfinite(sq) ? 2 : (sq > threshold ? 1 : 0);
                    if (d == 2) sq = 0;
                }
            }
            energy[f] = sq;
            decision[f] = d;
        }
    }
}

// Rejection: the factors over the gate lose their square-root information in both observation
// orders, so no later kernel counts them (the reference's FactorLifecycle::Rejected).
__global__ __launch_bounds__(kBlock) void k_nis_apply(const WinDev* wins, const uint8_t* decision) {
    const WinDev& w = wins[0];
    const size_t F = (size_t)w.F;
    double* ob = writable(w.ob);
    double* pf = writable(w.pf_ob);
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < w.F; f += gridDim.x * blockDim.x) {
        if (decision[f] != 1) continue;
        const int p = w.pf_pos[f];
        ob[OB_S00 * F + f] = 0.0; ob[OB_S10 * F + f] = 0.0; ob[OB_S11 * F + f] = 0.0;
        pf[OB_S00 * F + p] = 0.0; pf[OB_S10 * F + p] = 0.0; pf[OB_S11 * F + p] = 0.0;
    }
}

// Residual statistics of the evaluated (decision 0 or 1) and admitted (decision 0) factors
// (ResidualStatistics, BackendTypes.hpp:209-214; accumulation and quantiles,
// BackendOptimizer.cpp:436-465), 9 doubles per set: factors, components, squared norm, over
// gate, max NIS, worst factor (first index of the maximum; -1 when empty), robust cost (Huber),
// median NIS (sorted[n / 2]), p95 NIS (sorted[ceil(0.95 n) - 1]). Workgroup 0 sums; workgroups
// 1..4 select one quantile each, exactly, by a most-significant-digit radix select over the bit
// patterns of the energies (non-negative doubles order like their bits).
constexpr int kStatSet = 9;
constexpr int kStatThreads = 1024;
__device__ __forceinline__ bool in_set(uint8_t d, int set) { return set == 0 ? d <= 1 : d == 0; }

__global__ __launch_bounds__(kStatThreads) void k_nis_stats(const double* energy, const uint8_t* decision, int F,
                                                            double huber, double* out) {
    __shared__ double red[kStatThreads];
    __shared__ int ridx[kStatThreads];
    __shared__ unsigned hist[2048];
    __shared__ unsigned scan[kStatThreads];
    __shared__ unsigned long long s_prefix;
    __shared__ long long s_rank;
    const int t = threadIdx.x;
    constexpr int U = 4;  // elements per thread and round, loaded before they are used
    constexpr int stride = kStatThreads * U;
    if (blockIdx.x == 0) {
        for (int set = 0; set < 2; ++set) {
            double n = 0, sum = 0, over = 0, robust = 0, mx = -1;
            int arg = -1;
            for (int base = 0; base < F; base += stride) {
                double e[U];
                uint8_t d[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int f = base + u * kStatThreads + t;
                    d[u] = f < F ? decision[f] : 3;
                    e[u] = f < F ? energy[f] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    if (!in_set(d[u], set)) continue;
                    n += 1; sum += e[u]; robust += huber_cost(huber, e[u]);
                    if (d[u] == 1) over += 1;
                    if (e[u] > mx) { mx = e[u]; arg = base + u * kStatThreads + t; }  // indices rise: first maximum kept
                }
            }
            n = block_sum(n, red);
            sum = block_sum(sum, red);
            over = block_sum(over, red);
            robust = block_sum(robust, red);
            red[t] = mx;
            ridx[t] = arg;
            __syncthreads();
            for (int s = kStatThreads / 2; s > 0; s >>= 1) {
                if (t < s) {
                    const double a = red[t], b = red[t + s];
                    const int ia = ridx[t], ib = ridx[t + s];
                    if (b > a || (b == a && ib >= 0 && (ia < 0 || ib < ia))) { red[t] = b; ridx[t] = ib; }
                }
                __syncthreads();
            }
            if (t == 0) {
                double* o = out + set * kStatSet;
                o[0] = n; o[1] = 2 * n; o[2] = sum; o[3] = over;
                o[4] = n > 0 ? red[0] : 0.0;
                o[5] = n > 0 ? (double)ridx[0] : -1.0;
                o[6] = robust;
            }
            __syncthreads();
        }
        return;
    }
    const int q = blockIdx.x - 1, set = q / 2, which = q % 2;
    double nd = 0;
    for (int f = t; f < F; f += kStatThreads) nd += in_set(decision[f], set) ? 1.0 : 0.0;
    nd = block_sum(nd, red);
    const long long n = (long long)nd;
    double* o = out + set * kStatSet + 7 + which;
    if (n == 0) {
        if (t == 0) *o = 0.0;
        return;
    }
    if (t == 0) {
        s_prefix = 0;
        s_rank = which == 0 ? n / 2 : (long long)ceil(0.95 * (double)n) - 1;
    }
    unsigned long long mask = 0;
    for (int pass = 0; pass < 6; ++pass) {  // digits: bits 53-63, 42-52, 31-41, 20-30, 9-19, 0-8
        const int sh = pass < 5 ? 53 - 11 * pass : 0;
        const int width = pass < 5 ? 11 : 9;
        const unsigned long long dmask = (1ull << width) - 1;
        for (int b = t; b < 2048; b += kStatThreads) hist[b] = 0;
        __syncthreads();
        const unsigned long long prefix = s_prefix;
        for (int base = 0; base < F; base += stride) {
            unsigned long long key[U];
            uint8_t d[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int f = base + u * kStatThreads + t;
                d[u] = f < F ? decision[f] : 3;
                key[u] = f < F ? (unsigned long long)__double_as_longlong(energy[f]) : 0ull;
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (in_set(d[u], set) && (key[u] & mask) == prefix) atomicAdd(&hist[(key[u] >> sh) & dmask], 1u);
        }
        __syncthreads();
        // the bin holding rank k: exclusive prefix of the counts (two bins per thread), the one
        // thread whose bins span k records the digit
        const unsigned c0 = hist[2 * t], c1 = hist[2 * t + 1];
        scan[t] = c0 + c1;
        __syncthreads();
        for (int off = 1; off < kStatThreads; off <<= 1) {
            const unsigned v = t >= off ? scan[t - off] : 0u;
            __syncthreads();
            scan[t] += v;
            __syncthreads();
        }
        const long long k = s_rank;
        const long long excl = (long long)scan[t] - (long long)(c0 + c1);
        __syncthreads();
        if (k >= excl && k < excl + (long long)c0) {
            s_prefix = prefix | ((unsigned long long)(2 * t) << sh);
            s_rank = k - excl;
        } else if (k >= excl + (long long)c0 && k < excl + (long long)(c0 + c1)) {
            s_prefix = prefix | ((unsigned long long)(2 * t + 1) << sh);
            s_rank = k - excl - c0;
        }
        mask |= dmask << sh;
        __syncthreads();
    }
    if (t == 0) *o = __longlong_as_double((long long)s_prefix);
}


// ---------------------------------------------------------------------------------------------
// k_whiten: pixel observations of undistorted pinhole cameras -> unit bearing and the
// lower-triangular square-root information (s00, s10, s11) of the bearing in the sphere basis
// (FrameStore::observation, FrameStore.cpp:102-124; PinholeCamera::backProject, the
// undistorted case). One thread per observation; the record written is the 48-byte
// (bearing, square-root information) pair of the window's observation planes.
// status: 0 ok, 1 outside the image, 2 back-projection failed, 3 singular covariance or
// invalid whitening, 4 invalid sigma or camera index.
// ---------------------------------------------------------------------------------------------
// One observation of k_whiten: writes the record and returns the status.
__device__ __forceinline__ int whiten_one(double2 uv, int c, double sg, const double* intr, int ncam, double& b0,
                                          double& b1, double& b2, double& s00, double& s10, double& s11) {
    int st = 0;
    b0 = b1 = b2 = s00 = s10 = s11 = 0;
    if (c < 0 || c >= ncam || !isfinite(sg) || !(sg > 0)) return 4;
    const double* k = intr + 6 * c;  // fx fy cx cy width height
    const double u = uv.x, v = uv.y;
    if (!(u >= 0) || !(v >= 0) || !(u < k[4]) || !(v < k[5])) return 1;
    const V3<double> ray = {(u - k[2]) / k[0], (v - k[3]) / k[1], 1.0};
    const double rn = sqrt(sqnorm(ray));
    if (!isfinite(rn) || rn < 1e-12 || !isfinite(1.0 / k[0]) || !isfinite(1.0 / k[1])) return 2;
    const V3<double> b = {ray.x / rn, ray.y / rn, ray.z / rn};
    V3<double> c0, c1;
    sphere_basis(b, c0, c1);
    // D = (I - b b^T) J / |ray|, J = d ray / d pixel = [1/fx 0; 0 1/fy; 0 0]; M = B^T D
    const double j0 = 1.0 / k[0], j1 = 1.0 / k[1];
    const V3<double> d0 = {(1 - b.x * b.x) * j0 / rn, (-b.y * b.x) * j0 / rn, (-b.z * b.x) * j0 / rn};
    const V3<double> d1 = {(-b.x * b.y) * j1 / rn, (1 - b.y * b.y) * j1 / rn, (-b.z * b.y) * j1 / rn};
    const double m00 = dot(c0, d0), m01 = dot(c0, d1), m10 = dot(c1, d0), m11 = dot(c1, d1);
    const double s2 = sg * sg;
    const double cov00 = s2 * (m00 * m00 + m01 * m01);
    const double cov10 = s2 * (m10 * m00 + m11 * m01);
    const double cov11 = s2 * (m10 * m10 + m11 * m11);
    const double l00 = sqrt(cov00), l10 = cov10 / l00, l11 = sqrt(cov11 - l10 * l10);
    if (!(l00 > 0) || !(l11 > 0)) return 3;
    const double a00 = 1 / l00, a10 = -l10 / (l00 * l11), a11 = 1 / l11;
    if (!isfinite(a00) || !isfinite(a10) || !isfinite(a11)) return 3;
    s00 = a00; s10 = a10; s11 = a11;
    b0 = b.x; b1 = b.y; b2 = b.z;
    return st;
}

constexpr int kWhitenPer = 2;  // observations per thread (their loads issued together)

__global__ __launch_bounds__(kBlock) void k_whiten(int count, const double* px, const int32_t* cam,
                                                   const double* intr, int ncam, const double* sigma,
                                                   double sigma_all, double* bearing, double* sqrt_info,
                                                   int32_t* status) {
    constexpr int nb = kWhitenPer * kBlock;  // observations of the block
    __shared__ double stage[6 * nb];         // the block's records, stored back as contiguous runs
    const int t = threadIdx.x, i0 = blockIdx.x * nb;
    double2 uv[kWhitenPer];
    int c[kWhitenPer];
    double sg[kWhitenPer];
#pragma unroll
    for (int r = 0; r < kWhitenPer; ++r) {
        const int i = i0 + r * kBlock + t;
        const bool in = i < count;
        uv[r] = in ? reinterpret_cast<const double2*>(px)[i] : make_double2(0.0, 0.0);
        c[r] = in && cam ? cam[i] : 0;
        sg[r] = in && sigma ? sigma[i] : sigma_all;
    }
#pragma unroll
    for (int r = 0; r < kWhitenPer; ++r) {
        const int i = i0 + r * kBlock + t, o = r * kBlock + t;
        double b0, b1, b2, s00, s10, s11;
        const int st = whiten_one(uv[r], c[r], sg[r], intr, ncam, b0, b1, b2, s00, s10, s11);
        if (i < count) status[i] = st;
        stage[3 * o] = b0; stage[3 * o + 1] = b1; stage[3 * o + 2] = b2;
        stage[3 * nb + 3 * o] = s00; stage[3 * nb + 3 * o + 1] = s10; stage[3 * nb + 3 * o + 2] = s11;
    }
    __syncthreads();
    const int n3 = 3 * min(nb, count - i0);
    for (int j = t; j < n3; j += kBlock) {
        bearing[3 * (size_t)i0 + j] = stage[j];
        sqrt_info[3 * (size_t)i0 + j] = stage[3 * nb + j];
    }
}


// ---------------------------------------------------------------------------------------------
// k_reanchor: the anchor search of a SECOND_NEW slide (SlidingWindowManager.cpp:312-457) for
// landmarks anchored in the redundant slot v, one thread per landmark. The surviving views
// (observations outside v) are tried in the reference's preference order (newer frame first,
// then lower camera index, then observation order); the first view whose ray meets the
// landmark's world point at positive range, inverse range >= min_rho and within 0.03 rad, and
// under which every other surviving observation still evaluates, becomes the new anchor.
// chosen[i]: that observation (landmark order) or -1 (the landmark is retired); rho[i]: 1/range.
// With lms null, i runs over every landmark and only those of class 3 in cls are searched.
// ---------------------------------------------------------------------------------------------
__device__ bool reanchor_view_valid(const WinDev& w, const FrameTable& ft, int f, int frame, int cam, V3<double> bearing,
                                    double rho) {
    if (!(rho > 0) || !isfinite(rho)) return false;
    const size_t F = (size_t)w.F;
    const int fb = w.ob_fc[f] >> 8, cb = w.ob_fc[f] & 255;
    const M3<double> Ra = *reinterpret_cast<const M3<double>*>(ft.Rc[frame][cam]);
    const M3<double> Rb = *reinterpret_cast<const M3<double>*>(ft.Rc[fb][cb]);
    const V3<double> ta = load3(ft.tc[frame][cam]), tb = load3(ft.tc[fb][cb]);
    const V3<double> ba = (1.0 / sqrt(sqnorm(bearing))) * bearing;
    const V3<double> raw = {w.ob[OB_BX * F + f], w.ob[OB_BY * F + f], w.ob[OB_BZ * F + f]};
    const V3<double> bb = (1.0 / sqrt(sqnorm(raw))) * raw;
    const V3<double> pc = mulT(Rb, mul(Ra, (1.0 / rho) * ba) + ta - tb);
    const double range = sqrt(sqnorm(pc));
    if (!isfinite(pc.x) || !isfinite(pc.y) || !isfinite(pc.z) || !isfinite(range) || range < 1e-10) return false;
    const V3<double> n = (1.0 / range) * pc;
    const double c = fmin(1.0, fmax(-1.0, dot(bb, n))), sn = sqrt(sqnorm(cross(bb, n)));
    if (c < -1 + 1e-8) return false;
    const double f_theta = sn < 1e-8 ? 1.0 : atan2(sn, c) / sn;
    const V3<double> lg = f_theta * (n - c * bb);
    V3<double> c0, c1;
    sphere_basis(raw, c0, c1);
    const double s00 = w.ob[OB_S00 * F + f], s10 = w.ob[OB_S10 * F + f], s11 = w.ob[OB_S11 * F + f];
    const double r0 = s00 * dot(c0, lg), r1 = s10 * dot(c0, lg) + s11 * dot(c1, lg);
    return isfinite(r0) && isfinite(r1);
}

__global__ __launch_bounds__(kBlock) void k_reanchor(const WinDev* wins, int v, const int32_t* lms, const uint8_t* cls,
                                                     int count, int32_t* chosen, double* new_rho) {
    __shared__ FrameTable ft;
    const WinDev& w = wins[0];
    build_frame_table(w, w.pose, w.rig, ft);
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int l = lms ? lms[i] : i;
    if (!lms && cls[l] != 3) {
        chosen[i] = -1;
        new_rho[i] = 0;
        return;
    }
    const size_t L = (size_t)w.L, F = (size_t)w.F;
    const int ca = w.lm_fc[l] & 255;
    const M3<double> Ra = *reinterpret_cast<const M3<double>*>(ft.Rc[v][ca]);
    const V3<double> ta = load3(ft.tc[v][ca]);
    const V3<double> raw = {w.lmd[LM_BX * L + l], w.lmd[LM_BY * L + l], w.lmd[LM_BZ * L + l]};
    const V3<double> world = mul(Ra, (1.0 / (w.rho[l] * sqrt(sqnorm(raw)))) * raw) + ta;
    const int f0 = w.lm_off[l], f1 = w.lm_off[l + 1];
    auto key = [&](int f) {  // preference order: newer frame, lower camera, earlier observation
        return ((kMaxFrames - 1 - (w.ob_fc[f] >> 8)) << 24) | ((w.ob_fc[f] & 255) << 16) | (f - f0);
    };
    int result = -1;
    double rho_out = 0;
    int prev = -1;
    for (;;) {
        int cand = -1, best = 0x7fffffff;
        for (int f = f0; f < f1; ++f) {
            if ((w.ob_fc[f] >> 8) == v) continue;
            const int k = key(f);
            if (k > prev && k < best) { best = k; cand = f; }
        }
        if (cand < 0) break;
        prev = best;
        const int fc = w.ob_fc[cand] >> 8, cc = w.ob_fc[cand] & 255;
        const M3<double> Rc = *reinterpret_cast<const M3<double>*>(ft.Rc[fc][cc]);
        const V3<double> delta = mulT(Rc, world - load3(ft.tc[fc][cc]));
        const V3<double> ray = {w.ob[OB_BX * F + cand], w.ob[OB_BY * F + cand], w.ob[OB_BZ * F + cand]};
        const double range = dot(delta, ray);
        if (!(range > 0) || !isfinite(range) || 1 / range < w.min_rho ||
            atan2(sqrt(sqnorm(cross(delta, ray))), range) > 0.03)
            continue;
        bool valid = true;
        for (int f = f0; f < f1 && valid; ++f)
            if ((w.ob_fc[f] >> 8) != v && f != cand) valid = reanchor_view_valid(w, ft, f, fc, cc, ray, 1 / range);
        if (valid) {
            result = cand;
            rho_out = 1 / range;
            break;
        }
    }
    chosen[i] = result;
    new_rho[i] = rho_out;
}


// ---------------------------------------------------------------------------------------------
// k_elim_set: the landmarks a MARGIN_OLD slide of slot `victim` eliminates
// (Marginalizer::compute's landmark set: anchored in the slot or observed from it), one thread
// per landmark; elim[l] = 1 for those.
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_elim_set(const WinDev* wins, int victim, uint8_t* elim) {
    const WinDev& w = wins[0];
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= w.L) return;
    bool e = (w.lm_fc[l] >> 8) == victim;
    for (int f = w.lm_off[l]; f < w.lm_off[l + 1] && !e; ++f) e = (w.ob_fc[f] >> 8) == victim;
    elim[l] = e ? 1 : 0;
}

// ---------------------------------------------------------------------------------------------
// k_preintegrate: IMU preintegration of many intervals at once, one wavefront per interval, the
// arithmetic of ImuError::redoPreintegration (src/factor/ImuError.cpp:254-462) as restated by
// the host vib_preintegrate (midpoint rule, 15x15 covariance propagated per noise source through
// the transition F, whitening W = Lambda^-1/2 U^T of PseudoInverse::symmSqrtU,
// src/factor/PseudoInverse.hpp:116-145, by the same cyclic Jacobi). Every lane carries the
// interval's delta rotation, velocity, position and bias Jacobians (identical arithmetic in all
// lanes); the four 15x15 sensitivity matrices live in LDS, each F P F^T product spread over the
// lanes (14 entries each).
// ---------------------------------------------------------------------------------------------
namespace preint {
struct Mat3 {
    double a[9];
};
__device__ __forceinline__ Mat3 zero3() {
    Mat3 m;
#pragma unroll
    for (int i = 0; i < 9; ++i) m.a[i] = 0;
    return m;
}
__device__ __forceinline__ Mat3 eye3() { Mat3 m = zero3(); m.a[0] = m.a[4] = m.a[8] = 1; return m; }
__device__ __forceinline__ Mat3 mul(const Mat3& x, const Mat3& y) {
    Mat3 o;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) o.a[3 * r + c] = x.a[3 * r] * y.a[c] + x.a[3 * r + 1] * y.a[3 + c] + x.a[3 * r + 2] * y.a[6 + c];
    return o;
}
__device__ __forceinline__ Mat3 add(const Mat3& x, const Mat3& y) { Mat3 o;
#pragma unroll
    for (int i = 0; i < 9; ++i) o.a[i] = x.a[i] + y.a[i]; return o; }
__device__ __forceinline__ Mat3 scl(double s, const Mat3& x) { Mat3 o;
#pragma unroll
    for (int i = 0; i < 9; ++i) o.a[i] = s * x.a[i]; return o; }
__device__ __forceinline__ void mv(const Mat3& m, const double* v, double* o) {
#pragma unroll
    for (int r = 0; r < 3; ++r) o[r] = m.a[3 * r] * v[0] + m.a[3 * r + 1] * v[1] + m.a[3 * r + 2] * v[2];
}
__device__ __forceinline__ Mat3 skew(const double* v) {
    Mat3 m = zero3();
    m.a[1] = -v[2]; m.a[2] = v[1]; m.a[3] = v[2]; m.a[5] = -v[0]; m.a[6] = -v[1]; m.a[7] = v[0];
    return m;
}
struct Q4 {
    double w, x, y, z;
};
__device__ __forceinline__ Q4 qmul(const Q4& a, const Q4& b) {
    return {a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z, a.w * b.x + b.w * a.x + a.y * b.z - a.z * b.y,
            a.w * b.y + b.w * a.y + a.z * b.x - a.x * b.z, a.w * b.z + b.w * a.z + a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ Mat3 qmat(const Q4& q) {
    const double tx = 2 * q.x, ty = 2 * q.y, tz = 2 * q.z;
    const double twx = tx * q.w, twy = ty * q.w, twz = tz * q.w, txx = tx * q.x, txy = ty * q.x, txz = tz * q.x;
    const double tyy = ty * q.y, tyz = tz * q.y, tzz = tz * q.z;
    Mat3 r;
    r.a[0] = 1 - (tyy + tzz); r.a[1] = txy - twz; r.a[2] = txz + twy;
    r.a[3] = txy + twz; r.a[4] = 1 - (txx + tzz); r.a[5] = tyz - twx;
    r.a[6] = txz - twy; r.a[7] = tyz + twx; r.a[8] = 1 - (txx + tyy);
    return r;
}
__device__ __forceinline__ double pick(const Mat3& m, int q) {  // m.a[q] without a stack copy
    double v = m.a[0];
#pragma unroll
    for (int i = 1; i < 9; ++i) v = q == i ? m.a[i] : v;
    return v;
}
__device__ __forceinline__ double sinc(double x) {
    if (fabs(x) > 1e-6) return sin(x) / x;
    const double x2 = x * x;
    return 1.0 - x2 / 6.0 + x2 * x2 / 120.0 - x2 * x2 * x2 / 5040.0;
}
__device__ __forceinline__ Mat3 right_jacobian(const double* phi) {  // Forster et al. RSS 2015 eq. (8)
    const double n = sqrt(phi[0] * phi[0] + phi[1] * phi[1] + phi[2] * phi[2]);
    const Mat3 px = skew(phi), px2 = mul(px, px);
    if (n < 1.0e-4) return add(add(eye3(), scl(-0.5, px)), scl(1.0 / 6.0, px2));
    return add(add(eye3(), scl(-(1.0 - cos(n)) / (n * n), px)), scl((n - sin(n)) / (n * n * n), px2));
}
}  // namespace preint

__global__ __launch_bounds__(64) void k_preintegrate(const double* smp_all, const int* begin, const int* end,
                                                     const double* t0s, const double* t1s, const double* sbl_all,
                                                     vib_imu_params prm, const double* g0_all, const double* g1_all,
                                                     double* out_all, int* status) {
    using namespace preint;
    // 15x15 matrices padded to 16x16 tiles (zero row and column 15) at a pitch of 17 doubles, so
    // the strided operand reads of the MFMA products fall in distinct LDS banks
    constexpr int P = 17, PT = 16 * P;
    __shared__ double dP[4][PT];
    __shared__ double Tm[4][PT];
    __shared__ double Fm[PT];
    __shared__ double Qd[4][15];
    const int e = blockIdx.x, lane = threadIdx.x;
    const int S = VIB_IMU_SAMPLE_STRIDE;
    const double* smp = smp_all + (size_t)begin[e] * S;
    const int count = end[e] - begin[e];
    const double t0 = t0s[e], t1 = t1s[e];
    const double* sbl = sbl_all + 9 * (size_t)e;
    for (int i = lane; i < 4 * PT; i += 64) (&dP[0][0])[i] = 0;
    Q4 Dq{1, 0, 0, 0};
    Mat3 Cint = zero3(), Cdint = zero3(), cr = zero3(), dadbg = zero3(), dvdbg = zero3(), dpdbg = zero3();
    double vint[3] = {0, 0, 0}, pint[3] = {0, 0, 0};
    double time = t0;
    bool started = false;
    int steps = 0;
    wave_sync();
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
#pragma unroll
            for (int k = 0; k < 3; ++k) { w1[k] = (1 - r) * w0[k] + r * w1[k]; a1[k] = (1 - r) * a0[k] + r * a1[k]; }
        }
        if (dt <= 0.0) continue;
        if (!started) {
            started = true;
            const double r = dt / (next - s0[0]);
#pragma unroll
            for (int k = 0; k < 3; ++k) { w0[k] = r * w0[k] + (1 - r) * w1[k]; a0[k] = r * a0[k] + (1 - r) * a1[k]; }
        }
        double gsat = 1, asat = 1;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (fabs(w0[k]) > prm.g_max || fabs(w1[k]) > prm.g_max) gsat = 100;
            if (fabs(a0[k]) > prm.a_max || fabs(a1[k]) > prm.a_max) asat = 100;
        }
        double om[3], ac[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            om[k] = 0.5 * (w0[k] + w1[k]) - sbl[3 + k];
            ac[k] = 0.5 * (a0[k] + a1[k]) - sbl[6 + k];
        }
        const double half = 0.5 * dt * sqrt(om[0] * om[0] + om[1] * om[1] + om[2] * om[2]);
        const double sc = sinc(half) * 0.5 * dt;
        const Q4 dq{cos(half), sc * om[0], sc * om[1], sc * om[2]};
        const Q4 Dq1 = qmul(Dq, dq);
        const Mat3 C0 = qmat(Dq), C1 = qmat(Dq1), CC = add(C0, C1);
        double CCa[3];
        mv(CC, ac, CCa);
        const Mat3 Cint1 = add(Cint, scl(0.5 * dt, CC));
        double vint1[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) vint1[k] = vint[k] + 0.5 * dt * CCa[k];
        const Mat3 Cdint1 = add(add(Cdint, scl(dt, Cint)), scl(0.25 * dt * dt, CC));
        double pint1[3];  // the position update uses the velocity before this step
#pragma unroll
        for (int k = 0; k < 3; ++k) pint1[k] = pint[k] + dt * vint[k] + 0.25 * dt * dt * CCa[k];
        double odt[3] = {om[0] * dt, om[1] * dt, om[2] * dt};
        const Mat3 Jr = right_jacobian(odt);
        const Mat3 dadbg1 = add(dadbg, scl(dt, mul(C1, Jr)));
        const Q4 dqi{dq.w, -dq.x, -dq.y, -dq.z};
        const Mat3 cr1 = add(mul(qmat(dqi), cr), scl(dt, Jr));
        const Mat3 ax = skew(ac);
        const Mat3 mix = add(mul(mul(C0, ax), cr), mul(mul(C1, ax), cr1));
        const Mat3 dvdbg1 = add(dvdbg, scl(0.5 * dt, mix));
        const Mat3 dpdbg1 = add(add(dpdbg, scl(dt, dvdbg)), scl(0.25 * dt * dt, mix));
        // transition F (identity plus nine 3x3 blocks), built from the state before this step
        double tmpA[3], tmpB[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            tmpA[k] = vint[k] * dt + 0.25 * dt * dt * CCa[k];
            tmpB[k] = 0.5 * dt * CCa[k];
        }
        const Mat3 b03 = scl(-1, skew(tmpA)), b09 = add(scl(dt, dvdbg), scl(0.25 * dt * dt, mix));
        const Mat3 b012 = add(scl(-dt, Cint), scl(0.25 * dt * dt, CC)), b39 = scl(-dt, C1);
        const Mat3 b63 = scl(-1, skew(tmpB)), b69 = scl(0.5 * dt, mix), b612 = scl(-0.5 * dt, CC);
        for (int idx = lane; idx < PT; idx += 64) {
            const int r = idx / P, c = idx % P, br = r / 3, bc = c / 3, q = 3 * (r % 3) + c % 3;
            double v = r == c && r < 15 ? 1.0 : 0.0;
            if (r >= 15 || c >= 15) {
                Fm[idx] = 0.0;
                continue;
            }
            if (br == 0) {
                if (bc == 1) v = pick(b03, q);
                else if (bc == 2) v = (r % 3 == c % 3) ? dt : 0.0;
                else if (bc == 3) v = pick(b09, q);
                else if (bc == 4) v = pick(b012, q);
            } else if (br == 1) {
                if (bc == 3) v = pick(b39, q);
            } else if (br == 2) {
                if (bc == 1) v = pick(b63, q);
                else if (bc == 3) v = pick(b69, q);
                else if (bc == 4) v = pick(b612, q);
            }
            Fm[idx] = v;
        }
        if (lane < 15) {
            const int r = lane;
            Qd[0][r] = (r >= 3 && r < 6) ? gsat * dt : 0.0;
            Qd[1][r] = r < 3 ? 0.5 * dt * dt * dt * asat * asat * asat : ((r >= 6 && r < 9) ? asat * dt : 0.0);
            Qd[2][r] = (r >= 9 && r < 12) ? dt : 0.0;
            Qd[3][r] = r >= 12 ? dt : 0.0;
        }
        wave_sync();
        {
            // T_j = F dP_j, then dP_j = T_j F^T + Q_j, on v_mfma_f64_16x16x4f64: lane l gives
            // A[l % 16][k0 + l / 16] and B[k0 + l / 16][l % 16]; accumulator register rr holds
            // D[4 rr + l / 16][l % 16]
            const int li = lane % 16, lk = lane / 16;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int k0 = 0; k0 < 16; k0 += 4)
                    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Fm[li * P + k0 + lk], dP[j][(k0 + lk) * P + li], acc, 0, 0, 0);
#pragma unroll
                for (int rr = 0; rr < 4; ++rr) Tm[j][(4 * rr + lk) * P + li] = acc[rr];
            }
            wave_sync();
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int k0 = 0; k0 < 16; k0 += 4)
                    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Tm[j][li * P + k0 + lk], Fm[li * P + k0 + lk], acc, 0, 0, 0);
#pragma unroll
                for (int rr = 0; rr < 4; ++rr) {
                    const int r = 4 * rr + lk;
                    dP[j][r * P + li] = acc[rr] + (r == li && r < 15 ? Qd[j][r] : 0.0);
                }
            }
            wave_sync();
        }
        Dq = Dq1;
        Cint = Cint1;
        Cdint = Cdint1;
#pragma unroll
        for (int k = 0; k < 3; ++k) { vint[k] = vint1[k]; pint[k] = pint1[k]; }
        cr = cr1;
        dadbg = dadbg1;
        dvdbg = dvdbg1;
        dpdbg = dpdbg1;
        time = next;
        ++steps;
        if (next == t1) break;
    }
    if (steps <= 0) {
        if (lane == 0) status[e] = 1;
        return;
    }
    // P = sum_j sigma_j^2 sym(dP_j), then its eigendecomposition: Jacobi with the host's
    // rotation and stopping rule, in round-robin order (Brent and Luk 1985): each of the 15
    // rounds of a sweep applies 7 rotations of disjoint index pairs at once, so a sweep has 15
    // dependent phases instead of 105. The eigenvectors agree with the host's cyclic order up to
    // sign and rounding; the whitening they give is the same.
    const double s0 = prm.sigma_g_c * prm.sigma_g_c, s1 = prm.sigma_a_c * prm.sigma_a_c;
    const double s2 = prm.sigma_gw_c * prm.sigma_gw_c, s3 = prm.sigma_aw_c * prm.sigma_aw_c;
    double* A = Tm[0];
    double* V = Tm[1];
    for (int idx = lane; idx < 225; idx += 64) {
        const int r = idx / 15, c = idx % 15;
        const double v = (0.5 * dP[0][r * P + c] + 0.5 * dP[0][c * P + r]) * s0 +
                         (0.5 * dP[1][r * P + c] + 0.5 * dP[1][c * P + r]) * s1 +
                         (0.5 * dP[2][r * P + c] + 0.5 * dP[2][c * P + r]) * s2 +
                         (0.5 * dP[3][r * P + c] + 0.5 * dP[3][c * P + r]) * s3;
        A[idx] = v;
        V[idx] = r == c ? 1.0 : 0.0;
    }
    __shared__ double rot_c[8], rot_s[8];
    __shared__ int rot_p[8], rot_q[8];
    wave_sync();
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0, tot = 0;
        for (int i = 0; i < 225; ++i) {
            const double v = A[i] * A[i];
            tot += v;
            if (i / 15 != i % 15) off += v;
        }
        if (off == 0 || off <= 1e-32 * tot) break;
        for (int round = 0; round < 15; ++round) {
            if (lane < 7) {  // pair k = lane + 1 of the tournament on 15 indices
                const int x = (round + lane + 1) % 15, y = (round + 14 - lane) % 15;
                const int p = min(x, y), q = max(x, y);
                const double apq = A[p * 15 + q];
                double c = 1, sn = 0;
                if (apq != 0) {
                    const double theta = (A[q * 15 + q] - A[p * 15 + p]) / (2 * apq);
                    const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1));
                    c = 1 / sqrt(t * t + 1);
                    sn = t * c;
                }
                rot_c[lane] = c; rot_s[lane] = sn; rot_p[lane] = p; rot_q[lane] = q;
            }
            wave_sync();
            for (int idx = lane; idx < 105; idx += 64) {  // columns p, q of every row
                const int k = idx / 15, r = idx % 15, p = rot_p[k], q = rot_q[k];
                const double c = rot_c[k], sn = rot_s[k];
                const double x = A[r * 15 + p], y = A[r * 15 + q];
                A[r * 15 + p] = c * x - sn * y;
                A[r * 15 + q] = sn * x + c * y;
            }
            wave_sync();
            for (int idx = lane; idx < 105; idx += 64) {  // rows p, q of every column, and V
                const int k = idx / 15, r = idx % 15, p = rot_p[k], q = rot_q[k];
                const double c = rot_c[k], sn = rot_s[k];
                const double x = A[p * 15 + r], y = A[q * 15 + r];
                A[p * 15 + r] = c * x - sn * y;
                A[q * 15 + r] = sn * x + c * y;
                const double vx = V[r * 15 + p], vy = V[r * 15 + q];
                V[r * 15 + p] = c * vx - sn * vy;
                V[r * 15 + q] = sn * vx + c * vy;
            }
            wave_sync();
        }
    }
    // eigenvalues ascending (selection sort as on the host, by lane 0 in LDS), then W rows
    // s_i u_i^T
    __shared__ double ev[15];
    __shared__ int order[15];
    if (lane == 0) {
        for (int i = 0; i < 15; ++i) { ev[i] = A[i * 15 + i]; order[i] = i; }
        for (int i = 0; i < 15; ++i) {
            int m = i;
            for (int j = i + 1; j < 15; ++j) if (ev[j] < ev[m]) m = j;
            if (m != i) {
                const double tv = ev[i]; ev[i] = ev[m]; ev[m] = tv;
                const int to = order[i]; order[i] = order[m]; order[m] = to;
            }
        }
    }
    wave_sync();
    const double eps = 2.220446049250313e-16;
    const double tol = fmax(eps, eps * 15 * ev[14]);
    double* out = out_all + (size_t)e * VIB_IMU_STRIDE;
    for (int idx = lane; idx < 225; idx += 64) {
        const int i = idx / 15, k = idx % 15;
        const double sv = sqrt(ev[i] > tol ? 1.0 / ev[i] : 1.0 / tol);
        out[76 + idx] = sv * V[k * 15 + order[i]];
    }
    if (lane == 0) {
        const double* g0 = g0_all + 3 * (size_t)e;
        const double* g1 = g1_all + 3 * (size_t)e;
        out[0] = t1 - t0;
        out[1] = prm.g;
#pragma unroll
        for (int k = 0; k < 3; ++k) out[2 + k] = prm.p_BS[k];
        for (int k = 0; k < 4; ++k) out[5 + k] = prm.q_BS[k];
#pragma unroll
        for (int k = 0; k < 3; ++k) { out[9 + k] = g0[k]; out[12 + k] = g1[k]; }
        out[15] = Dq.x; out[16] = Dq.y; out[17] = Dq.z; out[18] = Dq.w;
#pragma unroll
        for (int k = 0; k < 3; ++k) { out[19 + k] = vint[k]; out[22 + k] = pint[k]; }
        for (int k = 0; k < 6; ++k) out[25 + k] = sbl[3 + k];
        for (int k = 0; k < 9; ++k) {
            out[31 + k] = dadbg.a[k]; out[40 + k] = dvdbg.a[k]; out[49 + k] = dpdbg.a[k];
            out[58 + k] = Cint.a[k]; out[67 + k] = Cdint.a[k];
        }
        status[e] = 0;
    }
}

}  // namespace vib

// ---------------------------------------------------------------------------------------------
// Launch wrappers used by the host driver.
// ---------------------------------------------------------------------------------------------
extern "C" {


// One launch per kernel id (the LM iteration is the sequence 0..11; id 0 is the visual
// normal equations plus the per-landmark sums).
const char* vibk_name(int k) {
    static const char* names[] = {"k_lin_visual", "k_retained_lin", "k_reduce_lin", "k_schur", "k_reduce_schur",
                                  "k_assemble", "k_cholesky", "k_backsub", "k_trial_cost", "k_retained_trial",
                                  "k_lm_update", "k_accept", "k_schur_apply", "k_refine"};
    return k >= 0 && k < 14 ? names[k] : "";
}

hipError_t vibk_launch(int k, vib::WinDev* wins, int nwin, const vib::Chunk* chunks, int nchunks, const vib::Chunk* lin,
                       int nlin, double** red, int max_len, int max_edges, int max_landmarks, int refinement, int mode,
                       int mixed, int fuse_reduce, hipStream_t s) {
    using namespace vib;
    const int slices = (max_len + kRedEntries - 1) / kRedEntries;
    const int rper = max_edges + 1, ritems = fuse_reduce ? nwin * rper : 0;  // retained items inside kernel 0
    switch (k) {
        case 0: {  // (small batches) the retained items, then the visual factors, in one launch
            const int rb = ritems ? ((ritems + kWaves - 1) / kWaves + kLinXcds - 1) / kLinXcds * kLinXcds : 0;
            const int grid = rb + nlin;
            if (grid == 0) break;
            if (mixed) hipLaunchKernelGGL(k_lin_visual<true>, dim3(grid), dim3(kBlock), 0, s, wins, lin, rb, ritems, rper);
            else hipLaunchKernelGGL(k_lin_visual<false>, dim3(grid), dim3(kBlock), 0, s, wins, lin, rb, ritems, rper);
        }
            if (nchunks) {
                if (mixed) hipLaunchKernelGGL(k_lm_sums<true>, dim3(nchunks), dim3(kBlock), 0, s, wins, chunks);
                else hipLaunchKernelGGL(k_lm_sums<false>, dim3(nchunks), dim3(kBlock), 0, s, wins, chunks);
            }
            break;
        case 1: hipLaunchKernelGGL(k_retained, dim3(nwin, max_edges + 1), dim3(64), 0, s, wins, 0); break;
        case 2:
            hipLaunchKernelGGL(k_reduce_lin, dim3(nwin, (max_len + kRedEntriesFused - 1) / kRedEntriesFused), dim3(kBlock), 0,
                               s, wins, red);
            break;
        case 3: {  // Schur products, then the visual-block partial sums in the same launch
            const int rslices = fuse_reduce ? (max_len + kRedEntriesFused - 1) / kRedEntriesFused : 0;
            const int grid = nchunks + nwin * rslices;
            if (grid == 0) break;
            if (mixed)
                hipLaunchKernelGGL(k_schur_mixed, dim3(grid), dim3(schur_block<true>()), 0, s, wins, chunks, nchunks, rslices, red);
            else hipLaunchKernelGGL(k_schur, dim3(grid), dim3(schur_block<false>()), 0, s, wins, chunks, nchunks, rslices, red);
            break;
        }
        case 4: hipLaunchKernelGGL(k_reduce, dim3(nwin, slices), dim3(kBlock), 0, s, wins, 1, red); break;
        case 5:
            hipLaunchKernelGGL(k_assemble, dim3(nwin, kAssembleSplit * ((kMaxRetained + kTile - 1) / kTile)), dim3(kBlock), 0,
                               s, wins, red);
            break;
        case 6:
            hipLaunchKernelGGL(k_cholesky, dim3(nwin), dim3(kCholThreads), 0, s, wins, refinement,
                               mixed && refinement > 0 ? 1 : 0);
            break;
        case 7:
            if (nchunks) {
                if (mixed) hipLaunchKernelGGL(k_backsub<true>, dim3(nchunks), dim3(kBlock), 0, s, wins, chunks);
                else hipLaunchKernelGGL(k_backsub<false>, dim3(nchunks), dim3(kBlock), 0, s, wins, chunks);
            }
            break;
        case 8: {  // (small batches) the retained trial items, then the visual trial cost
            const int rb = ritems ? ((ritems + kWaves - 1) / kWaves + kLinXcds - 1) / kLinXcds * kLinXcds : 0;
            if (rb) hipLaunchKernelGGL(k_trial_cost<true>, dim3(rb + nlin), dim3(kBlock), 0, s, wins, lin, rb, ritems, rper);
            else if (nlin) hipLaunchKernelGGL(k_trial_cost<false>, dim3(nlin), dim3(kBlock), 0, s, wins, lin, 0, 0, 1);
            break;
        }
        case 9: hipLaunchKernelGGL(k_retained, dim3(nwin, max_edges + 1), dim3(64), 0, s, wins, 1); break;
        case 10: hipLaunchKernelGGL(k_lm_update, dim3(nwin), dim3(kBlock), 0, s, wins, mode); break;
        case 11:
            hipLaunchKernelGGL(k_accept, dim3(nwin, (max_landmarks + kAcceptSlice - 1) / kAcceptSlice + 1), dim3(kBlock), 0, s,
                               wins);
            break;
        case 12: {  // exact-operator landmark sum and state residual of one refinement pass
            const int hslices = (kMaxRetained + kResidRows - 1) / kResidRows;
            const int grid = nchunks + nwin * hslices;
            if (mixed) hipLaunchKernelGGL(k_schur_apply<true>, dim3(grid), dim3(kBlock), 0, s, wins, chunks, nchunks, hslices);
            else hipLaunchKernelGGL(k_schur_apply<false>, dim3(grid), dim3(kBlock), 0, s, wins, chunks, nchunks, hslices);
            break;
        }
        case 13: hipLaunchKernelGGL(k_refine, dim3(nwin), dim3(kCholThreads), 0, s, wins, mode); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Plain streaming copy of n 16-byte words (traffic calibration of the HBM counters).
__global__ __launch_bounds__(256) void k_copy_probe(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

hipError_t vibk_nis_gate(const vib::WinDev* wins, const vib::Chunk* chunks, int nchunks, double threshold,
                         double* energy, uint8_t* decision, int F, double huber, double* stats, hipStream_t s) {
    if (nchunks > 0)
        hipLaunchKernelGGL(vib::k_nis_gate, dim3(nchunks), dim3(vib::kBlock), 0, s, wins, chunks, threshold, energy,
                           decision);
    hipLaunchKernelGGL(vib::k_nis_stats, dim3(5), dim3(vib::kStatThreads), 0, s, energy, decision, F, huber, stats);
    return hipGetLastError();
}

hipError_t vibk_nis_apply(const vib::WinDev* wins, const uint8_t* decision, int F, hipStream_t s) {
    const int blocks = std::max(1, std::min(1024, (F + vib::kBlock - 1) / vib::kBlock));
    hipLaunchKernelGGL(vib::k_nis_apply, dim3(blocks), dim3(vib::kBlock), 0, s, wins, decision);
    return hipGetLastError();
}

hipError_t vibk_whiten(int count, const double* px, const int32_t* cam, const double* intr, int ncam,
                       const double* sigma, double sigma_all, double* bearing, double* sqrt_info, int32_t* status,
                       hipStream_t s) {
    hipLaunchKernelGGL(vib::k_whiten, dim3((count + vib::kWhitenPer * vib::kBlock - 1) / (vib::kWhitenPer * vib::kBlock)),
                       dim3(vib::kBlock), 0, s, count, px,
                       cam, intr, ncam, sigma, sigma_all, bearing, sqrt_info, status);
    return hipGetLastError();
}

hipError_t vibk_reanchor(const vib::WinDev* wins, int v, const int32_t* lms, const uint8_t* cls, int count,
                         int32_t* chosen, double* new_rho, hipStream_t s) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(vib::k_reanchor, dim3((count + vib::kBlock - 1) / vib::kBlock), dim3(vib::kBlock), 0, s, wins, v,
                       lms, cls, count, chosen, new_rho);
    return hipGetLastError();
}

hipError_t vibk_elim_set(const vib::WinDev* wins, int L, int victim, uint8_t* elim, hipStream_t s) {
    hipLaunchKernelGGL(vib::k_elim_set, dim3((L + vib::kBlock - 1) / vib::kBlock), dim3(vib::kBlock), 0, s, wins, victim,
                       elim);
    return hipGetLastError();
}

hipError_t vibk_preintegrate(int count, const double* smp, const int* begin, const int* end, const double* t0,
                             const double* t1, const double* sbl, const vib_imu_params* prm, const double* g0,
                             const double* g1, double* out, int* status, hipStream_t s) {
    hipLaunchKernelGGL(vib::k_preintegrate, dim3(count), dim3(64), 0, s, smp, begin, end, t0, t1, sbl, *prm, g0, g1, out,
                       status);
    return hipGetLastError();
}

hipError_t vibk_copy_probe(const void* src, void* dst, size_t bytes, hipStream_t s) {
    const size_t n = bytes / 16;
    hipLaunchKernelGGL(k_copy_probe, dim3(4096), dim3(256), 0, s, (const uint4*)src, (uint4*)dst, n);
    return hipGetLastError();
}

size_t vibk_sym_eigen_log(int n) { return vib::eigen_log_doubles(n); }

int vibk_sym_eigen_max(void) { return vib::kEigMax; }

// count jobs (device array), every n <= max_n <= kEigMax (checked by the caller on the host).
hipError_t vibk_sym_eigen_batch(const vib::EigenJob* jobs, int count, int max_n, hipStream_t s) {
    if (count < 1 || max_n < 1 || max_n > vib::kEigMax) return hipErrorInvalidValue;
    if (max_n <= vib::kEigLds)
        hipLaunchKernelGGL(vib::k_sym_eigen_lds, dim3(count), dim3(vib::kEigThreads), 0, s, jobs);
    else hipLaunchKernelGGL(vib::k_sym_eigen, dim3(count), dim3(vib::kEigThreads), 0, s, jobs);
    return hipGetLastError();
}

hipError_t vibk_prior_gram(vib::WinDev* wins, int nwin, double* const* G, int max_cols, hipStream_t s) {
    if (nwin < 1 || max_cols < 1) return hipErrorInvalidValue;
    const int tiles = (max_cols + vib::kGramTile - 1) / vib::kGramTile;
    hipLaunchKernelGGL(vib::k_prior_gram, dim3(tiles * tiles, nwin), dim3(vib::kBlock), 0, s, wins, G, tiles);
    return hipGetLastError();
}

}  // extern "C"
