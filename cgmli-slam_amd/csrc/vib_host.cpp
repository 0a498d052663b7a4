// Host driver of the MI355X Schur backend: window slots in HBM, the batched LM loop and the
// C-ABI of include/vib.h.
//
// Counterparts in the reference:
//   window upload     VisualCudaBatch::linearizeSchur(newTopology) + configureRetained +
//                     configureResident (src/backend/VisualCudaBatch.cu:433-511,606-624,673-691)
//   LM loop           solveVisualSchur (src/backend/BackendOptimizer.cpp:27-331), here with the
//                     accept/reject decision evaluated on the device for every window at once
//   piecewise API     linearizeResident / solveOnline / acceptResident / readResident
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/vib.h"
#define VIB_HOST_VIEW 1  // host code only: the descriptor's pointers stay plain in both passes
#include "vib_device.hpp"
#include "vib_host.hpp"
#include "vib_math.hpp"

extern "C" {
const char* vibk_name(int k);
hipError_t vibk_launch(int k, vib::WinDev* wins, int nwin, const vib::Chunk* chunks, int nchunks, const vib::Chunk* lin,
                       int nlin, double** red, int max_len, int max_edges, int max_landmarks, int refinement, int mode,
                       int mixed, int fuse_reduce, hipStream_t s);
hipError_t vibk_prior_gram(vib::WinDev* wins, int nwin, double* const* G, int max_cols, hipStream_t s);
hipError_t vibk_copy_probe(const void* src, void* dst, size_t bytes, hipStream_t s);
hipError_t vibk_elim_set(const vib::WinDev* wins, int L, int victim, uint8_t* elim, hipStream_t s);
hipError_t vibk_reanchor(const vib::WinDev* wins, int v, const int32_t* lms, const uint8_t* cls, int count,
                         int32_t* chosen, double* new_rho, hipStream_t s);
int vibk_edit_scan_tiles(int n);
int vibk_edit_sort_blocks(int n);
int vibk_edit_pair_keys(void);
hipError_t vibk_edit_scan(const void* in, int n, void* tile_sum, void* out, int wide, hipStream_t s);
hipError_t vibk_edit_mark(const vib::WinDev* src, int L, const void* spec, unsigned long long* lmv, uint32_t* fv,
                          uint32_t* cnt, int32_t* lm_of, int* bad, hipStream_t s);
hipError_t vibk_edit_gather(const vib::WinDev* src, const vib::WinDev* dst, int Ls, int Fs, const void* spec,
                            const unsigned long long* lscan, const uint32_t* fscan, const uint32_t* cscan,
                            const int32_t* lm_of, int32_t* lmn, hipStream_t s);
hipError_t vibk_insert_resolve(const vib::WinDev* src, const int64_t* obs_uid, int nobs, const int64_t* new_uid, int nnew,
                               int32_t* obs_lm, int32_t* add_cnt, int* unmatched, hipStream_t s);
hipError_t vibk_edit_pair_hist(const vib::WinDev* dst, int F, const int32_t* lmn, uint32_t* counts, hipStream_t s);
hipError_t vibk_edit_pair_scatter(const vib::WinDev* dst, int F, const int32_t* lmn, const uint32_t* offs, hipStream_t s);
hipError_t vibk_strip_select(const vib::WinDev* src, int L, uint8_t* keep_lm, uint8_t* keep_f, int* out, hipStream_t s);
hipError_t vibk_second_new_classify(const vib::WinDev* src, int L, int v, uint8_t* cls, hipStream_t s);
hipError_t vibk_second_new_flags(const vib::WinDev* src, int L, int v, const uint8_t* cls, const int32_t* chosen,
                                 const double* chosen_rho, uint8_t* keep_lm, uint8_t* keep_f, int32_t* over_f,
                                 double* over_rho, uint8_t* marg_lm, int* counts, hipStream_t s);
hipError_t vibk_touch_frames(const vib::WinDev* src, int L, const uint8_t* sel, int* out, hipStream_t s);
hipError_t vibk_two_frame_select(const vib::WinDev* src, int L, int prev, int last, uint8_t* keep_lm, uint8_t* keep_f,
                                 int* out, hipStream_t s);
hipError_t vibk_whiten(int count, const double* px, const int32_t* cam, const double* intr, int ncam,
                       const double* sigma, double sigma_all, double* bearing, double* sqrt_info, int32_t* status,
                       hipStream_t s);
hipError_t vibk_nis_gate(const vib::WinDev* wins, const vib::Chunk* chunks, int nchunks, double threshold,
                         double* energy, uint8_t* decision, int F, double huber, double* stats, hipStream_t s);
hipError_t vibk_nis_apply(const vib::WinDev* wins, const uint8_t* decision, int F, hipStream_t s);
hipError_t vibk_preintegrate(int count, const double* smp, const int* begin, const int* end, const double* t0,
                             const double* t1, const double* sbl, const vib_imu_params* prm, const double* g0,
                             const double* g1, double* out, int* status, hipStream_t s);
hipError_t vibk_copy_segments(const vib::CopySeg* segs, int n, size_t max_bytes, hipStream_t s);
hipError_t vibk_sym_eigen_batch(const vib::EigenJob* jobs, int count, int max_n, hipStream_t s);
int vibk_sym_eigen_max(void);
size_t vibk_sym_eigen_log(int n);
}

namespace vib {

namespace {

int tangent_size(int kind) { return kind == VIB_KEY_POSE || kind == VIB_KEY_RIG ? 6 : (kind == VIB_KEY_SPEED_BIAS ? 9 : 1); }
int ambient_size(int kind) { return kind == VIB_KEY_POSE || kind == VIB_KEY_RIG ? 7 : (kind == VIB_KEY_SPEED_BIAS ? 9 : 1); }

// Bump allocator over one hipMalloc arena per slot.
struct Arena {
    char* base = nullptr;
    size_t size = 0, used = 0;
    template <class T>
    T* take(size_t n) {
        used = (used + 255) & ~size_t(255);
        T* p = reinterpret_cast<T*>(base + used);
        used += sizeof(T) * std::max<size_t>(n, 1);
        return p;
    }
};

}  // namespace

// Scratch of the no-allocation gate (vib_nis_gate_static): the window descriptor, its landmark
// chunks, the 18 statistics, the per-factor energies and decisions of one call, in one
// file-scope device array. Every GPU the library runs on holds its own copy until the process
// ends; host code reaches it only through the per-device address that vib_create looks up
// (the array's name is not a device address in host code).
constexpr size_t kStaticGateFactors = size_t(1) << 20;  // factors of a window
constexpr size_t kStaticGateChunks = size_t(1) << 17;   // landmark chunks of a window
constexpr size_t kScratchChunkOffset = 4096;            // after the descriptor (sizeof(WinDev) < 4 KB)
constexpr size_t kScratchStatsOffset = kScratchChunkOffset + kStaticGateChunks * sizeof(Chunk);
constexpr size_t kScratchEnergyOffset = kScratchStatsOffset + 256;
constexpr size_t kScratchDecisionOffset = kScratchEnergyOffset + kStaticGateFactors * sizeof(double);
constexpr size_t kScratchBytes = kScratchDecisionOffset + kStaticGateFactors;
static_assert(sizeof(WinDev) <= kScratchChunkOffset, "descriptor region");
static_assert(18 * sizeof(double) <= 256, "statistics region");
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device (`device`), looked up once per device.
int scratch_address(int device, void** out) {
    if (device < 0 || device >= kMaxDevices) return VIB_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            (void)hipGetLastError();
            return VIB_ERR_DEVICE;
        }
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return VIB_OK;
}

// 64 landmarks per chunk workgroup: a 20000-landmark window spreads over 313 workgroups (more
// than one per CU), against twice the chunk partials to reduce (measured against 32..192).
int HostWindow::chunk_size = 64;

// ------------------------------------------------------------------------------------------
// Host-side window copy with derived layout.
// ------------------------------------------------------------------------------------------
int HostWindow::build(const vib_window& w, std::string& err) {
    const int BAD = VIB_ERR_INVALID_ARGUMENT;
    *this = HostWindow();
    K = w.num_frames; C = w.num_cameras; L = w.num_landmarks; F = w.num_observations; E = w.num_imu;
    flags = w.flags; huber = w.huber_delta; min_rho = w.min_inverse_depth;
    td_lo = w.td_lower; td_hi = w.td_upper;
    if (K < 1 || C < 1) { err = "window must hold at least one frame and one camera"; return BAD; }
    // limits of the device layout (the visual block and the reduced system are sized for them)
    if (K > kMaxFrames) { err = "unsupported Schur dimensions: more than 20 frames"; return VIB_ERR_UNSUPPORTED_DIMENSIONS; }
    if (C > kMaxCameras) { err = "unsupported Schur dimensions: more than 4 cameras"; return VIB_ERR_UNSUPPORTED_DIMENSIONS; }
    if (L < 0 || F < 0 || E < 0) { err = "negative sizes"; return BAD; }
    if (!w.pose || !w.speed_bias || !w.rig || !w.td || !w.cam_rotation || !w.cam_translation) { err = "missing state"; return BAD; }
    if (L && (!w.lm_anchor_frame || !w.lm_anchor_cam || !w.lm_bearing || !w.lm_velocity || !w.lm_tref || !w.lm_rho)) { err = "missing landmarks"; return BAD; }
    if (F && (!w.obs_landmark || !w.obs_frame || !w.obs_cam || !w.obs_bearing || !w.obs_velocity || !w.obs_tref || !w.obs_sqrt_info)) { err = "missing observations"; return BAD; }
    if (E && (!w.imu_from || !w.imu_to || !w.imu_data)) { err = "missing IMU"; return BAD; }
    if (!(huber >= 0) || !std::isfinite(huber)) { err = "unsupported visual loss (Huber scale must be finite and >= 0)"; return VIB_ERR_UNSUPPORTED_LOSS; }
    if (!(min_rho > 0)) { err = "invalid inverse depth bound"; return BAD; }
    frame_id.assign(K, 0);
    for (int k = 0; k < K; ++k) frame_id[k] = w.frame_id ? w.frame_id[k] : k;
    pose.assign(w.pose, w.pose + 7 * K);
    sb.assign(w.speed_bias, w.speed_bias + 9 * K);
    pose_free.assign(K, 1);
    sb_width.assign(K, 9);
    for (int k = 0; k < K; ++k) {
        if (w.pose_free) pose_free[k] = w.pose_free[k] ? 1 : 0;
        if (w.sb_width) sb_width[k] = w.sb_width[k];
        if (sb_width[k] != 0 && sb_width[k] != 3 && sb_width[k] != 9) { err = "sb_width must be 0, 3 or 9"; return BAD; }
    }
    std::copy_n(w.rig, 7, rig);
    td = *w.td;
    cam.assign(12 * C, 0);
    for (int c = 0; c < C; ++c) {
        for (int i = 0; i < 9; ++i) cam[12 * c + i] = w.cam_rotation[9 * c + i];
        for (int i = 0; i < 3; ++i) cam[12 * c + 9 + i] = w.cam_translation[3 * c + i];
    }
    lm_frame.assign(L, 0); lm_cam.assign(L, 0); lm_free.assign(L, 1);
    lmd.assign((size_t)LM_PLANES * L, 0);
    rho.assign(L, 0);
    lm_uid.assign(L, 0);
    for (int l = 0; l < L; ++l) {
        lm_uid[l] = w.lm_id ? w.lm_id[l] : l;
        if (l && lm_uid[l] <= lm_uid[l - 1]) { err = "landmark ids must be strictly increasing"; return BAD; }
    }
    uid_next = L ? lm_uid[L - 1] + 1 : 0;
    for (int l = 0; l < L; ++l) {
        lm_frame[l] = w.lm_anchor_frame[l]; lm_cam[l] = w.lm_anchor_cam[l];
        if (lm_frame[l] < 0 || lm_frame[l] >= K || lm_cam[l] < 0 || lm_cam[l] >= C) { err = "landmark anchor out of range"; return BAD; }
        for (int i = 0; i < 3; ++i) {
            lmd[(size_t)(LM_BX + i) * L + l] = w.lm_bearing[3 * l + i];
            lmd[(size_t)(LM_VX + i) * L + l] = w.lm_velocity[3 * l + i];
        }
        lmd[(size_t)LM_T * L + l] = w.lm_tref[l];
        rho[l] = w.lm_rho[l];
        if (w.lm_free) lm_free[l] = w.lm_free[l] ? 1 : 0;
    }
    // observations sorted by landmark (stable counting sort)
    lm_off.assign(L + 1, 0);
    for (int f = 0; f < F; ++f) {
        const int l = w.obs_landmark[f];
        if (l < 0 || l >= L) { err = "observation landmark out of range"; return BAD; }
        if (w.obs_frame[f] < 0 || w.obs_frame[f] >= K || w.obs_cam[f] < 0 || w.obs_cam[f] >= C) { err = "observation frame/camera out of range"; return BAD; }
        lm_off[l + 1]++;
    }
    for (int l = 0; l < L; ++l) lm_off[l + 1] += lm_off[l];
    std::vector<int> cursor(lm_off.begin(), lm_off.end() - 1);
    ob_fc.assign(F, 0); ob_lm.assign(F, 0);
    ob.assign((size_t)OB_PLANES * F, 0);
    for (int f = 0; f < F; ++f) {
        const int l = w.obs_landmark[f];
        const int d = cursor[l]++;
        ob_fc[d] = (w.obs_frame[f] << 8) | w.obs_cam[f];
        ob_lm[d] = l;
        for (int i = 0; i < 3; ++i) {
            ob[(size_t)(OB_BX + i) * F + d] = w.obs_bearing[3 * f + i];
            ob[(size_t)(OB_VX + i) * F + d] = w.obs_velocity[3 * f + i];
            ob[(size_t)(OB_S00 + i) * F + d] = w.obs_sqrt_info[3 * f + i];
        }
        ob[(size_t)OB_T * F + d] = w.obs_tref[f];
    }
    imu_from.assign(w.imu_from, w.imu_from + E);
    imu_to.assign(w.imu_to, w.imu_to + E);
    imu.assign(w.imu_data, w.imu_data + (size_t)VIB_IMU_STRIDE * E);
    for (int e = 0; e < E; ++e)
        if (imu_from[e] < 0 || imu_from[e] >= K || imu_to[e] < 0 || imu_to[e] >= K || imu_from[e] == imu_to[e]) {
            err = "IMU edge out of range"; return BAD;
        }
    // prior
    prior_rows = w.prior_rows > 0 ? w.prior_rows : 0;
    prior_cols = 0;
    if (prior_rows) {
        const int B = w.prior_blocks;
        if (B <= 0 || !w.prior_key_kind || !w.prior_ref || !w.prior_A || !w.prior_r0) { err = "incomplete prior"; return BAD; }
        int refs = 0;
        for (int b = 0; b < B; ++b) {
            const int k = w.prior_key_kind[b];
            if (k < 0 || k > 3) { err = "invalid prior key kind"; return BAD; }
            prior_kind.push_back(k);
            prior_id.push_back(w.prior_key_id ? w.prior_key_id[b] : 0);
            refs += ambient_size(k);
            prior_cols += tangent_size(k);
        }
        prior_ref.assign(w.prior_ref, w.prior_ref + refs);
        prior_A.assign(w.prior_A, w.prior_A + (size_t)prior_rows * prior_cols);
        prior_r0.assign(w.prior_r0, w.prior_r0 + prior_rows);
    }
    return layout(err) ? VIB_OK : BAD;
}

bool HostWindow::layout(std::string& err) {
    if (!layout_states(err)) return false;
    lm_col.assign(L, -1);
    ne = 0;
    for (int l = 0; l < L; ++l) if (lm_free[l]) lm_col[l] = nr + ne++;
    make_chunks();
    // factors in (anchor frame, observer frame) pair order (stable counting sort)
    std::vector<int> start(kMaxFrames * kMaxFrames + 1, 0), order(F);
    auto pair_key = [&](int f) { return lm_frame[ob_lm[f]] * kMaxFrames + (ob_fc[f] >> 8); };
    for (int f = 0; f < F; ++f) start[pair_key(f) + 1]++;
    for (size_t k = 1; k < start.size(); ++k) start[k] += start[k - 1];
    for (int f = 0; f < F; ++f) order[start[pair_key(f)]++] = f;
    pf_ob.assign((size_t)OB_PLANES * F, 0);
    pf_an.assign((size_t)LM_PLANES * F, 0);
    pf_ids.assign(F, 0); pf_lm.assign(F, 0); pf_pos.assign(F, 0);
    for (int p = 0; p < F; ++p) {
        const int f = order[p], l = ob_lm[f];
        for (int q = 0; q < OB_PLANES; ++q) pf_ob[(size_t)q * F + p] = ob[(size_t)q * F + f];
        for (int q = 0; q < LM_PLANES; ++q) pf_an[(size_t)q * F + p] = lmd[(size_t)q * L + l];
        pf_ids[p] = (lm_frame[l] << 24) | (lm_cam[l] << 16) | ob_fc[f];
        pf_lm[p] = l;
        pf_pos[f] = p;
    }
    bulk = true;
    return true;
}

void HostWindow::make_chunks() {
    chunks.clear();
    const int cs = std::max(1, chunk_size);
    for (int l0 = 0, i = 0; l0 < L; l0 += cs, ++i) chunks.push_back(Chunk{0, i, l0, std::min(L, l0 + cs)});
    lin_chunks.clear();
    for (int p0 = 0, i = 0; p0 < F; p0 += kLinFactors, ++i) lin_chunks.push_back(Chunk{0, i, p0, std::min(F, p0 + kLinFactors)});
}

bool HostWindow::layout_states(std::string& err) {
    int c = 0;
    pose_col.assign(K, -1); sb_col.assign(K, -1);
    for (int k = 0; k < K; ++k) if (pose_free[k]) { pose_col[k] = c; c += 6; }
    rig_col = (flags & VIB_RIG_FREE) ? c : -1;
    if (rig_col >= 0) c += 6;
    td_col = (flags & VIB_TD_FREE) ? c : -1;
    if (td_col >= 0) c += 1;
    nv = c;
    for (int k = 0; k < K; ++k) if (sb_width[k]) { sb_col[k] = c; c += sb_width[k]; }
    nr = c;
    if (nr <= 0) { err = "no retained state to solve"; return false; }
    // prior blocks
    pblk.clear(); pfree.clear();
    int chart = 0, ref = 0;
    for (size_t b = 0; b < prior_kind.size(); ++b) {
        PriorBlockDev d{};
        d.kind = prior_kind[b];
        d.slot = -1;
        if (d.kind == VIB_KEY_POSE || d.kind == VIB_KEY_SPEED_BIAS) {
            for (int k = 0; k < K; ++k) if (frame_id[k] == prior_id[b]) d.slot = k;
            if (d.slot < 0) { err = "prior key refers to a frame outside the window"; return false; }
        }
        d.tsize = tangent_size(d.kind);
        d.chart = chart; d.ref = ref;
        chart += d.tsize; ref += ambient_size(d.kind);
        if (d.kind == VIB_KEY_POSE) { d.col = pose_col[d.slot]; d.width = d.col >= 0 ? 6 : 0; }
        else if (d.kind == VIB_KEY_SPEED_BIAS) { d.col = sb_col[d.slot]; d.width = d.col >= 0 ? sb_width[d.slot] : 0; }
        else if (d.kind == VIB_KEY_RIG) { d.col = rig_col; d.width = rig_col >= 0 ? 6 : 0; }
        else { d.col = td_col; d.width = td_col >= 0 ? 1 : 0; }
        for (int i = 0; i < d.width; ++i) pfree.push_back(((int)b << 8) | i);
        pblk.push_back(d);
    }
    // static column map of the retained factors: IMU (edge << 8 | local column) x2, prior
    colmap.assign(4 * (size_t)nr, -1);
    for (int e = 0; e < E; ++e) {
        const int s0 = imu_from[e], s1 = imu_to[e];
        const int cols[4] = {pose_col[s0], sb_col[s0], pose_col[s1], sb_col[s1]};
        const int width[4] = {cols[0] >= 0 ? 6 : 0, cols[1] >= 0 ? sb_width[s0] : 0, cols[2] >= 0 ? 6 : 0,
                              cols[3] >= 0 ? sb_width[s1] : 0};
        const int off[4] = {0, 6, 15, 21};
        for (int b = 0; b < 4; ++b)
            for (int i = 0; i < width[b]; ++i) {
                int* slot = &colmap[4 * (size_t)(cols[b] + i)];
                if (slot[0] < 0) slot[0] = (e << 8) | (off[b] + i);
                else if (slot[1] < 0) slot[1] = (e << 8) | (off[b] + i);
                else { err = "a state block is shared by more than two IMU edges"; return false; }
            }
    }
    for (const int32_t f : pfree) {
        const PriorBlockDev& d = pblk[f >> 8];
        colmap[4 * (size_t)(d.col + (f & 255)) + 2] = f;
    }
    // structural tiles of the factor in the speed/bias-first order: the visual block is dense
    // (landmark Schur complement), an IMU edge couples its four blocks, the prior all of its
    // columns; then the fill of a tile Cholesky (tile (i, j) once (i, k) and (j, k) are).
    {
        const int nsb = nr - nv, T = (nr + kTile - 1) / kTile;
        auto tile_of = [&](int col) { return (col < nv ? col + nsb : col - nv) / kTile; };
        std::vector<uint8_t> S((size_t)T * T, 0);
        auto mark = [&](const std::vector<int>& cols) {
            for (int a : cols)
                for (int b : cols) {
                    const int ta = tile_of(a), tb = tile_of(b);
                    S[(size_t)std::max(ta, tb) * T + std::min(ta, tb)] = 1;
                }
        };
        std::vector<int> cols;
        for (int c = 0; c < nv; ++c) cols.push_back(c);
        mark(cols);
        for (int e = 0; e < E; ++e) {
            cols.clear();
            for (int s : {imu_from[e], imu_to[e]}) {
                if (pose_col[s] >= 0) for (int i = 0; i < 6; ++i) cols.push_back(pose_col[s] + i);
                if (sb_col[s] >= 0) for (int i = 0; i < sb_width[s]; ++i) cols.push_back(sb_col[s] + i);
            }
            mark(cols);
        }
        cols.clear();
        for (const int32_t f : pfree) cols.push_back(pblk[f >> 8].col + (f & 255));
        mark(cols);
        for (int i = 0; i < T; ++i) S[(size_t)i * T + i] = 1;
        for (int k = 0; k < T; ++k)
            for (int i = k + 1; i < T; ++i)
                if (S[(size_t)i * T + k])
                    for (int j = k + 1; j <= i; ++j)
                        if (S[(size_t)j * T + k]) S[(size_t)i * T + j] = 1;
        chol_mask.assign(kMaxTiles, 0);
        for (int k = 0; k < T; ++k)
            for (int i = k; i < T; ++i)
                if (S[(size_t)i * T + k]) chol_mask[k] |= 1u << i;
    }
    return true;
}

// ------------------------------------------------------------------------------------------
// Slot: device arena holding one window.
// ------------------------------------------------------------------------------------------
bool PinnedBuffer::reserve(size_t bytes) {
    if (bytes <= cap && p) return true;
    const size_t want = std::max({bytes, cap + cap / 2, (size_t)1 << 16});
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    if (hipHostMalloc(reinterpret_cast<void**>(&p), want, hipHostMallocDefault) != hipSuccess) {
        p = nullptr;
        (void)hipGetLastError();
        return false;
    }
    cap = want;
    return true;
}

PinnedBuffer::~PinnedBuffer() {
    if (p) (void)hipHostFree(p);
}

Staging::~Staging() {
    if (buf) (void)hipFree(buf);
    for (hipEvent_t e : done)
        if (e) (void)hipEventDestroy(e);
}

void CopyBatch::up(void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return;
    const size_t off = (pay_.size() + 7) & ~size_t(7);
    pay_.resize(off + bytes);
    std::memcpy(pay_.data() + off, src, bytes);
    ups_.push_back(Up{dst, off, bytes});
}

void CopyBatch::zero(void* dst, size_t bytes) {
    if (bytes) ups_.push_back(Up{dst, SIZE_MAX, bytes});
}

void CopyBatch::down(void* host, const void* dev, size_t bytes) {
    if (bytes) downs_.push_back(Down{host, dev, bytes});
}

bool CopyBatch::flush(hipStream_t s, Staging& st) {
    if (ups_.empty() && downs_.empty()) return true;
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t nu = ups_.size(), nd = downs_.size(), n = nu + nd;
    std::vector<size_t> doff(nd);
    size_t dbytes = 0, max_up = 0, max_down = 0;
    for (size_t j = 0; j < nd; ++j) {
        doff[j] = dbytes;
        dbytes += (downs_[j].bytes + 7) & ~size_t(7);
        max_down = std::max(max_down, downs_[j].bytes);
    }
    for (const Up& u : ups_) max_up = std::max(max_up, u.bytes);
    // device staging: upload payload | segment table | download area
    const size_t seg_off = al(pay_.size()), down_off = al(seg_off + n * sizeof(CopySeg)), total = down_off + dbytes;
    auto drop = [&]() {
        ups_.clear();
        downs_.clear();
        pay_.clear();
        (void)hipGetLastError();
        return false;
    };
    if (total > st.cap) {  // grow: the old buffer may still be read by issued work
        if (st.buf && (hipStreamSynchronize(s) != hipSuccess || hipFree(st.buf) != hipSuccess)) return drop();
        st.buf = nullptr;
        st.cap = 0;
        const size_t want = std::max(total, size_t(1) << 20);
        if (hipMalloc(&st.buf, want) != hipSuccess) {
            st.buf = nullptr;
            return drop();
        }
        st.cap = want;
    }
    const int k = st.next;
    st.next ^= 1;
    if (!st.done[k] && hipEventCreateWithFlags(&st.done[k], hipEventDisableTiming) != hipSuccess) {
        st.done[k] = nullptr;
        return drop();
    }
    if (st.recorded[k] && hipEventSynchronize(st.done[k]) != hipSuccess) return drop();  // the transfer that last read host[k]
    char* dev = static_cast<char*>(st.buf);
    const size_t hbytes = seg_off + n * sizeof(CopySeg);
    if (!st.host[k].reserve(hbytes)) return drop();
    char* h = st.host[k].p;
    if (!pay_.empty()) std::memcpy(h, pay_.data(), pay_.size());
    CopySeg* seg = reinterpret_cast<CopySeg*>(h + seg_off);
    for (size_t i = 0; i < nu; ++i)
        seg[i] = CopySeg{(uint64_t)(uintptr_t)ups_[i].dst,
                         ups_[i].off == SIZE_MAX ? 0 : (uint64_t)(uintptr_t)(dev + ups_[i].off), ups_[i].bytes};
    for (size_t j = 0; j < nd; ++j)
        seg[nu + j] = CopySeg{(uint64_t)(uintptr_t)(dev + down_off + doff[j]), (uint64_t)(uintptr_t)downs_[j].dev,
                              downs_[j].bytes};
    const CopySeg* dseg = reinterpret_cast<const CopySeg*>(dev + seg_off);
    bool ok = hipMemcpyAsync(dev, h, hbytes, hipMemcpyHostToDevice, s) == hipSuccess &&
              hipEventRecord(st.done[k], s) == hipSuccess && (st.recorded[k] = true) &&
              vibk_copy_segments(dseg, (int)nu, max_up, s) == hipSuccess &&
              vibk_copy_segments(dseg + nu, (int)nd, max_down, s) == hipSuccess;
    if (ok && nd) {
        ok = st.back.reserve(dbytes) &&
             hipMemcpyAsync(st.back.p, dev + down_off, dbytes, hipMemcpyDeviceToHost, s) == hipSuccess &&
             hipStreamSynchronize(s) == hipSuccess;
        for (size_t j = 0; ok && j < nd; ++j) std::memcpy(downs_[j].host, st.back.p + doff[j], downs_[j].bytes);
    }
    if (!ok) return drop();
    ups_.clear();
    downs_.clear();
    pay_.clear();
    return true;
}

Slot::~Slot() { release(); }

void Slot::release() {
    if (arena) (void)hipFree(arena);
    if (host_terms) (void)hipFree(host_terms);
    host_terms = nullptr;
    host_cap = 0;
    host_set = false;
    arena = nullptr;
    bytes = 0;
    capacity = 0;
    used = false;
}

size_t Slot::footprint(const HostWindow& h) {
    const size_t nr = h.nr, nv = h.nv, tri = nv * (nv + 1) / 2, nch = h.chunks.size();
    const size_t B = h.pblk.size(), m = h.prior_rows, n = h.prior_cols;
    size_t d = 0;
    d += 2 * (16 * h.K + 8 + h.L);                     // state + candidate
    d += 12 * h.C + LM_PLANES * h.L + OB_PLANES * h.F;  // streams
    d += 6 * h.F + LO_PLANES * h.L + h.L;                // linearization outputs
    d += h.L;                                            // landmark ids
    d += (size_t)VIB_IMU_STRIDE * h.E;
    d += m * n + m + n * n + h.prior_ref.size();
    d += 2 * nr * nr + factor_doubles((int)nr + kTile) + 8 * nr + 64;
    d += 465 * (size_t)h.E + 36 * B + n + 4 * (size_t)h.E + 4;
    d += (sizeof(WinDev) + 256) / 8 + 1;  // the descriptor and pointer of the prior Gram launch
    const size_t nlin = h.lin_chunks.size();
    d += 2 * nch * (tri + nv) + nch * CS_PLANES + 2 * (tri + nv);
    d += nlin * (tri + nv + CS_PLANES);
    d += (size_t)(OB_PLANES + LM_PLANES + FL_PLANES) * h.F;
    size_t i = 3 * h.L + 1 + 4 * h.F + 2 * h.E + h.pfree.size() + 4 * nr;
    return 8 * d + 4 * i + sizeof(PriorBlockDev) * B + sizeof(LmState) + 64 * 256;
}

bool Slot::carve(const HostWindow& h, char* base, size_t bytes, WinDev& d, SlotCarve& x) {
    Arena a{base, bytes, 0};
    d = WinDev{};
    d.K = h.K; d.C = h.C; d.L = h.L; d.F = h.F; d.E = h.E;
    d.nv = h.nv; d.nr = h.nr; d.ne = h.ne; d.rig_col = h.rig_col; d.td_col = h.td_col; d.flags = h.flags;
    d.tri = h.nv * (h.nv + 1) / 2;
    d.huber = h.huber; d.min_rho = h.min_rho; d.td_lo = h.td_lo; d.td_hi = h.td_hi;
    for (int k = 0; k < kMaxTiles; ++k) d.chol_mask[k] = k < (int)h.chol_mask.size() ? h.chol_mask[k] : 0;
    for (int k = 0; k < kMaxFrames; ++k) {
        d.pose_col[k] = k < h.K ? h.pose_col[k] : -1;
        d.sb_col[k] = k < h.K ? h.sb_col[k] : -1;
        d.sb_w[k] = k < h.K ? h.sb_width[k] : 0;
    }
    d.pose = a.take<double>(7 * h.K); d.sb = a.take<double>(9 * h.K); d.rig = a.take<double>(7);
    d.td = a.take<double>(1); d.rho = a.take<double>(h.L);
    d.cpose = a.take<double>(7 * h.K); d.csb = a.take<double>(9 * h.K); d.crig = a.take<double>(7);
    d.ctd = a.take<double>(1); d.crho = a.take<double>(h.L);
    d.cam = a.take<double>(12 * h.C);
    d.lm_fc = a.take<int32_t>(h.L);
    d.lm_off = a.take<int32_t>(h.L + 1);
    d.lm_col = a.take<int32_t>(h.L);
    d.lmd = a.take<double>((size_t)LM_PLANES * h.L);
    d.ob_fc = a.take<int32_t>(h.F);
    d.ob = a.take<double>((size_t)OB_PLANES * h.F);
    d.fac_u = a.take<double>((size_t)6 * h.F);
    d.fac_lm = a.take<double>((size_t)FL_PLANES * h.F);
    d.pf_ob = a.take<double>((size_t)OB_PLANES * h.F);
    d.pf_an = a.take<double>((size_t)LM_PLANES * h.F);
    d.pf_ids = a.take<int32_t>(h.F);
    d.pf_lm = a.take<int32_t>(h.F);
    d.pf_pos = a.take<int32_t>(h.F);
    d.lm_out = a.take<double>((size_t)LO_PLANES * h.L);
    d.lm_delta = a.take<double>(h.L);
    d.lm_uid = a.take<int64_t>(h.L);
    d.imu_from = a.take<int32_t>(h.E);
    d.imu_to = a.take<int32_t>(h.E);
    d.imu = a.take<double>((size_t)VIB_IMU_STRIDE * h.E);
    d.prior_rows = h.prior_rows; d.prior_cols = h.prior_cols; d.prior_blocks = (int)h.pblk.size();
    d.pblk = a.take<PriorBlockDev>(h.pblk.size());
    d.pA = a.take<double>((size_t)h.prior_rows * h.prior_cols);
    d.pr0 = a.take<double>(h.prior_rows);
    d.pG = a.take<double>((size_t)h.prior_cols * h.prior_cols);
    d.pref = a.take<double>(h.prior_ref.size());
    d.pfree = a.take<int32_t>(h.pfree.size());
    d.prior_free = (int)h.pfree.size();
    const size_t nr = h.nr;
    d.imu_J = a.take<double>((size_t)450 * h.E); d.imu_r = a.take<double>((size_t)15 * h.E);
    d.pD = a.take<double>(36 * h.pblk.size()); d.patr = a.take<double>(h.prior_cols);
    d.ret_sc = a.take<double>(4 * (size_t)h.E + 4);
    d.colmap = a.take<int32_t>(4 * nr);
    d.Hfull = a.take<double>(nr * nr); d.gfull = a.take<double>(nr);
    d.fac = a.take<double>(factor_doubles((int)nr + kTile)); d.Scopy = a.take<double>(nr * nr);
    d.rhs = a.take<double>(nr); d.dstate = a.take<double>(nr); d.scratch = a.take<double>(4 * nr + 64);
    const size_t len = d.tri + h.nv, nch = h.chunks.size();
    const size_t nlin = h.lin_chunks.size();
    d.lin_part = a.take<double>(nlin * len); d.schur_part = a.take<double>(nch * len);
    d.lin_sc = a.take<double>(nlin * CS_PLANES);
    d.chunk_sc = a.take<double>(nch * CS_PLANES);
    d.nchunks = (int)nch;
    d.nlin = (int)nlin;
    x.reduced_lin = a.take<double>(len); x.reduced_schur = a.take<double>(len);
    d.lm = a.take<LmState>(1);
    x.gram_win = a.take<WinDev>(1);
    x.gram_out = a.take<double*>(1);
    return a.used <= bytes;
}

bool Slot::fill(const HostWindow& h, const WinDev& d, const SlotCarve& x, char* base, size_t bytes, bool streams,
                hipStream_t stream, Staging& st, CopyBatch& cb) {
    // every copy into the arena stays inside the carved footprint (host-side bounds check)
    const char* lo = base;
    bool ok = true;
    auto inside = [&](const void* dst, const void* src, size_t n) {
        const char* p = reinterpret_cast<const char*>(dst);
        return src && p >= lo && p + n <= lo + bytes;
    };
    auto up = [&](const void* dst, const void* src, size_t n) {  // small: one batched transfer
        if (n == 0) return;
        if (!inside(dst, src, n)) ok = false;
        else cb.up(const_cast<void*>(dst), src, n);
    };
    auto big = [&](const void* dst, const void* src, size_t n) {  // streams: transfers of their own
        if (n == 0) return true;
        return inside(dst, src, n) &&
               hipMemcpyAsync(const_cast<void*>(dst), src, n, hipMemcpyHostToDevice, stream) == hipSuccess;
    };
    up(d.pose, h.pose.data(), 8 * h.pose.size());
    up(d.sb, h.sb.data(), 8 * h.sb.size());
    up(d.rig, h.rig, 56);
    up(d.td, &h.td, 8);
    up(d.cam, h.cam.data(), 8 * h.cam.size());
    up(d.imu_from, h.imu_from.data(), 4 * h.E);
    up(d.imu_to, h.imu_to.data(), 4 * h.E);
    up(d.imu, h.imu.data(), 8 * h.imu.size());
    up(d.pblk, h.pblk.data(), sizeof(PriorBlockDev) * h.pblk.size());
    up(d.pA, h.prior_A.data(), 8 * h.prior_A.size());
    up(d.pr0, h.prior_r0.data(), 8 * h.prior_r0.size());
    up(d.pref, h.prior_ref.data(), 8 * h.prior_ref.size());
    up(d.pfree, h.pfree.data(), 4 * h.pfree.size());
    up(d.colmap, h.colmap.data(), 4 * h.colmap.size());
    // candidate starts equal to the state
    up(d.cpose, h.pose.data(), 8 * h.pose.size());
    up(d.csb, h.sb.data(), 8 * h.sb.size());
    up(d.crig, h.rig, 56);
    up(d.ctd, &h.td, 8);
    LmState lm{};
    lm.lambda = 1e-4; lm.increase = 2; lm.new_point = 1; lm.status = 0; lm.max_iterations = 1;
    up(d.lm, &lm, sizeof(lm));
    if (h.prior_rows) {  // descriptor and output of the prior Gram launch
        up(x.gram_win, &d, sizeof(WinDev));
        up(x.gram_out, &d.pG, sizeof(double*));
    }
    std::vector<int32_t> fc;  // staged landmark keys: kept until the stream has completed
    if (streams) {
        fc.resize(h.L);
        for (int l = 0; l < h.L; ++l) fc[l] = (h.lm_frame[l] << 8) | h.lm_cam[l];
        ok = ok && big(d.rho, h.rho.data(), 8 * h.rho.size()) && big(d.lm_fc, fc.data(), 4 * fc.size()) &&
             big(d.lm_off, h.lm_off.data(), 4 * h.lm_off.size()) && big(d.lm_col, h.lm_col.data(), 4 * h.lm_col.size()) &&
             big(d.lmd, h.lmd.data(), 8 * h.lmd.size()) && big(d.ob_fc, h.ob_fc.data(), 4 * h.ob_fc.size()) &&
             big(d.ob, h.ob.data(), 8 * h.ob.size()) && big(d.pf_ob, h.pf_ob.data(), 8 * h.pf_ob.size()) &&
             big(d.pf_an, h.pf_an.data(), 8 * h.pf_an.size()) && big(d.pf_ids, h.pf_ids.data(), 4 * h.pf_ids.size()) &&
             big(d.pf_lm, h.pf_lm.data(), 4 * h.pf_lm.size()) && big(d.pf_pos, h.pf_pos.data(), 4 * h.pf_pos.size()) &&
             big(d.crho, h.rho.data(), 8 * h.rho.size()) && big(d.lm_uid, h.lm_uid.data(), 8 * h.lm_uid.size());
        // the layouts of the copies match the carved footprint
        ok = ok && (int)h.rho.size() == h.L && (int)h.lm_uid.size() == h.L && (int)h.lm_off.size() == h.L + 1 && h.lmd.size() == (size_t)LM_PLANES * h.L &&
             (int)h.ob_fc.size() == h.F && h.ob.size() == (size_t)OB_PLANES * h.F && h.pf_ob.size() == (size_t)OB_PLANES * h.F;
    }
    ok = ok && cb.flush(stream, st);
    // G = A^T A once per prior (one-window launch)
    if (ok && h.prior_rows) ok = vibk_prior_gram(x.gram_win, 1, x.gram_out, h.prior_cols, stream) == hipSuccess;
    return ok && hipStreamSynchronize(stream) == hipSuccess;
}

int Slot::upload(const HostWindow& h, hipStream_t stream, Staging& st, std::string& err) {
    const size_t need = footprint(h);
    // The slot is live only once every part of the window is on the device: a failure on any
    // path below leaves it unused (find_slot no longer returns it) with its arena kept for reuse.
    used = false;
    if (!arena || need > capacity) {  // a re-upload that fits (a slide) keeps the arena
        // a quarter of headroom: the window's next appends are edited in place
        const size_t want = need + need / 4;
        release();
        if (hipMalloc(&arena, want) != hipSuccess) {
            arena = nullptr;
            (void)hipGetLastError();
            err = "device allocation failed";
            return VIB_ERR_OUT_OF_MEMORY;
        }
        capacity = want;
    }
    host = h;
    bytes = need;
    host_set = false;  // host-evaluated terms belong to the previous window
    SlotCarve x;
    if (!carve(h, static_cast<char*>(arena), bytes, dev, x)) { err = "arena overflow"; return VIB_ERR_DEVICE; }
    reduced_lin = x.reduced_lin;
    reduced_schur = x.reduced_schur;
    CopyBatch cb;
    if (!fill(h, dev, x, static_cast<char*>(arena), bytes, true, stream, st, cb)) {
        (void)hipGetLastError();
        err = "upload failed";
        return VIB_ERR_DEVICE;
    }
    used = true;
    return VIB_OK;
}

// ------------------------------------------------------------------------------------------
// Batch: device copies of the WinDev descriptors and the concatenated chunk list.
// ------------------------------------------------------------------------------------------
Batch::~Batch() {
    if (d_wins) (void)hipFree(d_wins);
    if (d_chunks) (void)hipFree(d_chunks);
    if (d_lin) (void)hipFree(d_lin);
    if (d_red) (void)hipFree(d_red);
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
}

bool Batch::build(std::vector<Slot*>& slots, hipStream_t stream, Staging& xs, std::string& err) {
    nwin = (int)slots.size();
    std::vector<WinDev> wins(nwin);
    std::vector<Chunk> all;
    std::vector<double*> red(2 * nwin);
    max_len = 1;
    max_edges = 0;
    max_landmarks = 0;
    std::vector<Chunk> lin;
    // Landmark chunks are merged `merge` at a time when the batch alone fills the device many
    // times over: fewer per-chunk Schur partials to write and reduce. The per-slot arrays are
    // sized for the unmerged count, so any merge fits.
    size_t fine = 0, fine_lin = 0;
    for (Slot* s : slots) {
        fine += s->host.chunks.size();
        fine_lin += s->host.lin_chunks.size();
    }
    int merge = chunk_merge > 0 ? chunk_merge : (fine >= kMergeChunksAbove ? kChunkMerge : 1);
    if (const char* m = std::getenv("VIB_CHUNK_MERGE")) merge = std::max(1, std::atoi(m));
    // the same for the visual factor ranges: a merged range is one longer pair-ordered range,
    // whose workgroup zeroes, and writes back, its visual-block accumulator once
    int lin_merge = this->lin_merge > 0 ? this->lin_merge : (fine_lin >= kMergeLinAbove ? kLinMerge : 1);
    if (const char* m = std::getenv("VIB_LIN_MERGE")) lin_merge = std::max(1, std::atoi(m));
    for (int i = 0; i < nwin; ++i) {
        Slot& s = *slots[i];
        wins[i] = s.dev;
        wins[i].hostH = s.host_set ? s.host_terms : nullptr;
        wins[i].hostG = s.host_set ? s.host_terms + (size_t)s.host.nr * s.host.nr : nullptr;
        const std::vector<Chunk>& hc = s.host.chunks;
        for (size_t c0 = 0; c0 < hc.size(); c0 += merge) {
            const size_t c1 = std::min(hc.size(), c0 + merge);
            all.push_back(Chunk{i, (int)(c0 / merge), hc[c0].lm_begin, hc[c1 - 1].lm_end});
        }
        wins[i].nchunks = (int)((hc.size() + merge - 1) / merge);
        const std::vector<Chunk>& hl = s.host.lin_chunks;
        for (size_t c0 = 0; c0 < hl.size(); c0 += lin_merge) {
            const size_t c1 = std::min(hl.size(), c0 + lin_merge);
            lin.push_back(Chunk{i, (int)(c0 / lin_merge), hl[c0].lm_begin, hl[c1 - 1].lm_end});
        }
        wins[i].nlin = (int)((hl.size() + lin_merge - 1) / lin_merge);
        red[2 * i] = s.reduced_lin;
        red[2 * i + 1] = s.reduced_schur;
        max_len = std::max(max_len, wins[i].tri + wins[i].nv);
        max_edges = std::max(max_edges, s.host.E);
        max_landmarks = std::max(max_landmarks, s.host.L);
    }
    // XCD-aware order of the visual work list: workgroup i runs on XCD i % 8, so list position
    // 8 s + x holds chunk x * n8 + s, and each XCD takes one contiguous range of the pair-ordered
    // chunks. The per-factor coupling stores of one anchor frame then meet in one L2 instead of
    // leaving eight partial lines. Positions without a chunk are empty (local = -1).
    {
        const int n = (int)lin.size(), n8 = (n + kXcds - 1) / kXcds;
        std::vector<Chunk> order((size_t)n8 * kXcds, Chunk{0, -1, 0, 0});
        for (int c = 0; c < n; ++c) order[(size_t)(c % n8) * kXcds + c / n8] = lin[c];
        lin.swap(order);
    }
    nchunks = (int)all.size();
    nlin = (int)lin.size();
    // host-side bounds check of the work lists against the windows they index
    for (const Chunk& c : all)
        if (c.win < 0 || c.win >= nwin || c.lm_begin < 0 || c.lm_begin > c.lm_end || c.lm_end > wins[c.win].L ||
            c.local >= wins[c.win].nchunks) { err = "landmark chunk outside its window"; return false; }
    for (const Chunk& c : lin)
        if (c.local >= 0 && (c.win < 0 || c.win >= nwin || c.lm_begin < 0 || c.lm_begin > c.lm_end ||
                             c.lm_end > wins[c.win].F || c.local >= wins[c.win].nlin)) {
            err = "visual factor chunk outside its window"; return false;
        }
    // a rebuilt batch keeps its device buffers when they are large enough; a captured graph of
    // the previous contents is dropped
    if (graph_exec) { (void)hipGraphExecDestroy(graph_exec); graph_exec = nullptr; }
    if (graph) { (void)hipGraphDestroy(graph); graph = nullptr; }
    auto grow = [](auto*& ptr, size_t& cap, size_t need, size_t elem) {
        if (ptr && cap >= need) return true;
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        cap = std::max(need, cap + cap / 2);
        return hipMalloc((void**)&ptr, elem * cap) == hipSuccess;
    };
    if (!grow(d_wins, cap_wins, (size_t)nwin, sizeof(WinDev)) ||
        !grow(d_red, cap_red, (size_t)nwin, 2 * sizeof(double*)) ||
        !grow(d_chunks, cap_chunks, (size_t)std::max(1, nchunks), sizeof(Chunk)) ||
        !grow(d_lin, cap_lin, (size_t)std::max(1, nlin), sizeof(Chunk))) { err = "batch allocation failed"; return false; }
    CopyBatch cb;  // descriptors and work lists in one transfer, in stream order before the launches
    cb.up(d_wins, wins.data(), sizeof(WinDev) * nwin);
    cb.up(d_chunks, all.data(), sizeof(Chunk) * nchunks);
    cb.up(d_lin, lin.data(), sizeof(Chunk) * nlin);
    cb.up(d_red, red.data(), sizeof(double*) * 2 * nwin);
    if (!cb.flush(stream, xs)) { err = "batch upload failed"; return false; }
    return true;
}

bool Batch::launch(int k, hipStream_t s, int refinement, int mode) {
    return vibk_launch(k, d_wins, nwin, d_chunks, nchunks, d_lin, nlin, d_red, max_len, max_edges, max_landmarks, refinement,
                       mode, mixed, fuse_reduce() ? 1 : 0, s) == hipSuccess;
}

// Batches of up to kFuseReduceWindows windows run the retained linearization (kernel 1) as extra
// workgroups of kernel 0, the visual-block partial sums (kernel 2) as extra workgroups of
// kernel 3 and the retained trial cost (kernel 9) as extra workgroups of kernel 8, so an
// iteration skips kernels 1, 2 and 9. Larger batches launch them separately: their many small
// items would otherwise run at the occupancy of the large visual, Schur and trial kernels.
// A linearization alone always sums the partials.
bool Batch::fuse_reduce() const { return nwin <= kFuseReduceWindows; }
bool Batch::linearize(hipStream_t s) {
    return launch(0, s, 0, 0) && (fuse_reduce() || launch(1, s, 0, 0)) && launch(2, s, 0, 0);
}

// A mixed-precision step with refinement runs `refinement` passes of the exact-operator sum
// (kernel 12) and the correction solve (kernel 13, mode 1 on the last pass, which also finishes
// the step) between the factorization and the back-substitution.
std::vector<std::pair<int, int>> Batch::sequence(bool iteration, int refinement, int mode) const {
    std::vector<std::pair<int, int>> q;
    if (iteration) {
        q.push_back({0, 0});
        if (!fuse_reduce()) { q.push_back({1, 0}); q.push_back({2, 0}); }
    }
    for (int k = 3; k <= 6; ++k) q.push_back({k, mode});
    if (mixed && refinement > 0)
        for (int i = 0; i < refinement; ++i) {
            q.push_back({12, 0});
            q.push_back({13, i + 1 == refinement ? 1 : 0});
        }
    for (int k = 7; k < 12; ++k)
        if (k != 9 || !fuse_reduce()) q.push_back({k, mode});
    return q;
}

bool Batch::step(hipStream_t s, int refinement, int mode) {
    for (const auto& km : sequence(false, refinement, mode))
        if (!launch(km.first, s, refinement, km.second)) return false;
    return true;
}

bool Batch::iteration(hipStream_t s, int refinement) {
    for (const auto& km : sequence(true, refinement, 0))
        if (!launch(km.first, s, refinement, km.second)) return false;
    return true;
}

bool Batch::capture(hipStream_t s, int refinement) {
    if (graph_exec) return true;
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) return false;
    const bool ok = iteration(s, refinement);
    if (hipStreamEndCapture(s, &graph) != hipSuccess || !ok) return false;
    return hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0) == hipSuccess;
}

}  // namespace vib

// ============================================================================================
// C-ABI
// ============================================================================================
using namespace vib;

struct vib_context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string error;
    std::map<int, std::unique_ptr<Slot>> slots;
    std::mutex mutex;
    bool use_graph = true;
    vib::Batch requests;  // descriptors of the last vib_request_iterate set (buffers reused)
    size_t hbm_budget = 0;       // bytes of resident windows allowed (0: unlimited)
    void* scratch = nullptr;     // device address of this device's `scratch` array
    void* preint_buf = nullptr;  // grow-only device buffer of the preintegration launches
    size_t preint_cap = 0;
    // Grow-only device resources of the piecewise API and the slides, so that steady-state
    // calls allocate nothing: descriptors of one window and of a solve's window set, the
    // eigensolver's matrices, small per-landmark buffers, and the slots of the marginal system
    // and of the two-frame solve.
    vib::Batch one, solve;
    void* eig_buf = nullptr;
    size_t eig_cap = 0;
    void* aux_buf = nullptr;
    size_t aux_cap = 0;
    std::unique_ptr<vib::Slot> two_frame_slot;
    // the marginal systems and the edit buffers of the slides of one batch (one each per slot)
    std::vector<std::unique_ptr<vib::Slot>> marginal_pool;
    std::vector<std::pair<void*, size_t>> edit_pool;
    // Window edits: the buffers of the scans and selections, and a spare arena that an edit in
    // place (a slide) builds the new window in before it takes the slot's place (the slot's old
    // arena becomes the spare).
    void* edit_buf = nullptr;
    size_t edit_cap = 0;
    void* spare = nullptr;
    size_t spare_cap = 0;
    // work-list shapes chosen by vib_autotune: (precision, log2 of a batch's visual factor
    // ranges) -> (chunk merge, range merge)
    std::map<std::pair<int, int>, std::pair<int, int>> tuned;
    void* tune_buf = nullptr;  // state snapshot of the windows being tuned
    size_t tune_cap = 0;
    vib::Staging xfer;         // batched small transfers (CopyBatch)
    ~vib_context() {
        if (preint_buf) (void)hipFree(preint_buf);
        if (eig_buf) (void)hipFree(eig_buf);
        if (aux_buf) (void)hipFree(aux_buf);
        if (edit_buf) (void)hipFree(edit_buf);
        if (spare) (void)hipFree(spare);
        if (tune_buf) (void)hipFree(tune_buf);
        for (auto& e : edit_pool)
            if (e.first) (void)hipFree(e.first);
    }
};

namespace {

int fail(vib_context* ctx, int code, const std::string& msg) {
    if (ctx) ctx->error = msg;
    return code;
}

// A grow-only device buffer of at least `bytes` (null when the allocation fails). It grows by at
// least half its size: the windows of many sessions pass through one buffer, and a reallocation
// (hipFree waits for the whole device) each time a larger one comes would stall every stream.
void* grow_device(void*& buf, size_t& cap, size_t bytes) {
    if (bytes <= cap && buf) return buf;
    const size_t want = std::max({bytes, cap + cap / 2, (size_t)1 << 16});
    if (buf) (void)hipFree(buf);
    buf = nullptr;
    cap = 0;
    if (hipMalloc(&buf, want) != hipSuccess) {
        buf = nullptr;
        (void)hipGetLastError();
        return nullptr;
    }
    cap = want;
    return buf;
}

// Size class of a batch for the autotuned work-list shapes: log2 of its visual factor ranges.
int tune_class(const std::vector<Slot*>& slots) {
    size_t n = 0;
    for (const Slot* s : slots) n += s->host.lin_chunks.size();
    int c = 0;
    while ((size_t(1) << (c + 1)) <= n) ++c;
    return c;
}

// The work-list shapes vib_autotune chose for batches like this one (the built-in rule when it
// did not tune this precision and size class).
void apply_tuning(vib_context* ctx, Batch& b, const std::vector<Slot*>& slots) {
    auto it = ctx->tuned.find({b.mixed, tune_class(slots)});
    b.chunk_merge = it == ctx->tuned.end() ? 0 : it->second.first;
    b.lin_merge = it == ctx->tuned.end() ? 0 : it->second.second;
}

Slot* find_slot(vib_context* ctx, int slot) {
    auto it = ctx->slots.find(slot);
    return it == ctx->slots.end() || !it->second->used ? nullptr : it->second.get();
}

int status_from_lm(int s) {
    if (s >= 0) return VIB_OK;
    return -s;
}

const char* status_message(int code) {
    switch (code) {
        case VIB_ERR_INVALID_RESIDUAL: return "invalid residual during linearization";
        case VIB_ERR_NOT_POSITIVE_DEFINITE: return "device Schur step failed: damped reduced system not positive definite";
        case VIB_ERR_COST_MISMATCH: return "joint cost disagrees with accepted trial";
        case VIB_ERR_NO_ACCEPTABLE_STEP: return "no acceptable device Schur trial";
        default: return "device solve failed";
    }
}

bool read_lm(vib_context* ctx, Slot& s, LmState& lm) {
    CopyBatch cb;
    cb.down(&lm, s.dev.lm, sizeof(LmState));
    return cb.flush(ctx->stream, ctx->xfer);
}

bool write_lm(vib_context* ctx, Slot& s, const LmState& lm) {
    CopyBatch cb;
    cb.up(s.dev.lm, &lm, sizeof(LmState));
    return cb.flush(ctx->stream, ctx->xfer);
}

LmState fresh_lm(const vib_solve_options& o) {
    LmState lm{};
    lm.lambda = o.initial_lambda > 0 ? o.initial_lambda : 1e-4;
    lm.increase = 2; lm.new_point = 1; lm.status = 0; lm.max_iterations = std::max(1, o.max_iterations);
    lm.continuous = (o.flags & VIB_SOLVE_CONTINUOUS) ? 1 : 0;
    return lm;
}

// Report of a finished (or stopped) LM solve from its device state.
int fill_report(const LmState& s, const Slot& slot, double secs, vib_solve_report& r) {
    int st = status_from_lm(s.status);
    if (s.status == 0 && !s.accepted && !s.converged) st = VIB_ERR_NO_ACCEPTABLE_STEP;
    r.status = st;
    r.iterations = s.iteration;
    r.accepted_steps = s.accepted;
    r.converged = s.converged;
    r.effective_parameters = slot.host.nr + slot.host.ne;
    r.retained_dimension = slot.host.nr;
    r.initial_cost = s.initial_cost;
    r.final_cost = s.cost;
    r.final_lambda = s.lambda;
    r.seconds = secs;
    return st;
}

// Runs the LM loop for a set of slots. All windows iterate in lock-step; finished windows are
// skipped on the device. Returns the per-window reports.
int run_lm(vib_context* ctx, std::vector<Slot*>& slots, const vib_solve_options& o, std::vector<vib_solve_report>& reps) {
    const auto t0 = std::chrono::steady_clock::now();
    if (o.precision != VIB_FP64 && o.precision != VIB_MIXED) return fail(ctx, VIB_ERR_INVALID_ARGUMENT, "unknown precision");
    const int maxit = std::max(1, o.max_iterations);
    for (Slot* s : slots)
        if (s->host_set)
            return fail(ctx, VIB_ERR_UNSUPPORTED,
                        "a window with host-evaluated factors is solved through vib_linearize / vib_solve_step");
    for (Slot* s : slots)
        if (!write_lm(ctx, *s, fresh_lm(o))) return fail(ctx, VIB_ERR_DEVICE, "LM state upload failed");
    Batch& batch = ctx->solve;
    batch.mixed = o.precision == VIB_MIXED;
    apply_tuning(ctx, batch, slots);
    std::string err;
    if (!batch.build(slots, ctx->stream, ctx->xfer, err)) return fail(ctx, VIB_ERR_DEVICE, err);
    const bool graph = ctx->use_graph && batch.capture(ctx->stream, o.refinement_steps);
    std::vector<LmState> states(slots.size());
    const auto t_iter = std::chrono::steady_clock::now();  // the time budget covers the iterations
    for (int it = 0; it < maxit; ++it) {
        bool ok;
        if (graph) ok = hipGraphLaunch(batch.graph_exec, ctx->stream) == hipSuccess;
        else ok = batch.iteration(ctx->stream, o.refinement_steps);
        if (!ok) return fail(ctx, VIB_ERR_DEVICE, std::string("kernel launch failed: ") + hipGetErrorString(hipGetLastError()));
        // all done?
        bool running = false;
        for (size_t i = 0; i < slots.size(); ++i) {
            if (hipMemcpyAsync(&states[i], slots[i]->dev.lm, sizeof(LmState), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
                return fail(ctx, VIB_ERR_DEVICE, "LM state read failed");
        }
        if (hipStreamSynchronize(ctx->stream) != hipSuccess)
            return fail(ctx, VIB_ERR_DEVICE, std::string("device failure: ") + hipGetErrorString(hipGetLastError()));
        for (auto& s : states) running = running || s.status == 0;
        if (!running) break;
        if (o.time_limit_seconds > 0 && it + 1 >= o.min_iterations &&
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t_iter).count() >= o.time_limit_seconds)
            break;
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    reps.assign(slots.size(), vib_solve_report{});
    int first_error = VIB_OK;
    for (size_t i = 0; i < slots.size(); ++i) {
        const int st = fill_report(states[i], *slots[i], secs, reps[i]);
        if (st != VIB_OK && first_error == VIB_OK) first_error = st;
    }
    if (first_error != VIB_OK) ctx->error = status_message(first_error);
    return first_error;
}

// Host copy of a slot with the state it holds on the device now.
// Downloads the resident state (poses, speed/bias, rig, time offset, inverse ranges) into the
// slot's host mirror.
bool refresh_state(vib_context* ctx, Slot& s) {
    HostWindow& h = s.host;
    const WinDev& d = s.dev;
    h.pose.resize(7 * (size_t)h.K);
    h.sb.resize(9 * (size_t)h.K);
    h.rho.resize(h.L);
    CopyBatch cb;
    cb.down(h.pose.data(), d.pose, 56 * (size_t)d.K);
    cb.down(h.sb.data(), d.sb, 72 * (size_t)d.K);
    cb.down(h.rig, d.rig, 56);
    cb.down(&h.td, d.td, 8);
    cb.down(h.rho.data(), d.rho, 8 * (size_t)d.L);
    return cb.flush(ctx->stream, ctx->xfer);
}

// Makes the mirror's landmark and observation vectors current after a device edit (they are
// downloaded from the resident window; the host never rebuilds them).
bool ensure_bulk(vib_context* ctx, Slot& s) {
    HostWindow& h = s.host;
    if (h.bulk) return true;
    const size_t L = h.L, F = h.F;
    const WinDev& d = s.dev;
    std::vector<int32_t> fc(L);
    h.lm_off.assign(L + 1, 0); h.lm_col.assign(L, -1); h.lmd.assign((size_t)LM_PLANES * L, 0); h.rho.assign(L, 0);
    h.lm_uid.assign(L, 0);
    h.ob_fc.assign(F, 0); h.ob.assign((size_t)OB_PLANES * F, 0);
    h.pf_ob.assign((size_t)OB_PLANES * F, 0); h.pf_an.assign((size_t)LM_PLANES * F, 0);
    h.pf_ids.assign(F, 0); h.pf_lm.assign(F, 0); h.pf_pos.assign(F, 0);
    auto dn = [&](void* dst, const void* src, size_t n) {
        return n == 0 || hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, ctx->stream) == hipSuccess;
    };
    const bool ok = dn(fc.data(), d.lm_fc, 4 * L) && dn(h.lm_off.data(), d.lm_off, 4 * (L + 1)) &&
                    dn(h.lm_col.data(), d.lm_col, 4 * L) && dn(h.lmd.data(), d.lmd, 8 * h.lmd.size()) &&
                    dn(h.rho.data(), d.rho, 8 * L) && dn(h.lm_uid.data(), d.lm_uid, 8 * L) &&
                    dn(h.ob_fc.data(), d.ob_fc, 4 * F) &&
                    dn(h.ob.data(), d.ob, 8 * h.ob.size()) && dn(h.pf_ob.data(), d.pf_ob, 8 * h.pf_ob.size()) &&
                    dn(h.pf_an.data(), d.pf_an, 8 * h.pf_an.size()) && dn(h.pf_ids.data(), d.pf_ids, 4 * F) &&
                    dn(h.pf_lm.data(), d.pf_lm, 4 * F) && dn(h.pf_pos.data(), d.pf_pos, 4 * F) &&
                    hipStreamSynchronize(ctx->stream) == hipSuccess;
    if (!ok) return false;
    h.lm_frame.assign(L, 0); h.lm_cam.assign(L, 0); h.lm_free.assign(L, 0); h.ob_lm.assign(F, 0);
    for (size_t l = 0; l < L; ++l) {
        h.lm_frame[l] = fc[l] >> 8;
        h.lm_cam[l] = fc[l] & 255;
        h.lm_free[l] = h.lm_col[l] >= 0;
        for (int f = h.lm_off[l]; f < h.lm_off[l + 1]; ++f) h.ob_lm[f] = (int)l;
    }
    h.bulk = true;
    return true;
}

// The frame-level part of a window (frames, state, cameras, IMU, prior) without its landmark
// and observation vectors: the host part of an edited window.
HostWindow frame_meta(const HostWindow& cur) {
    HostWindow m;
    m.K = cur.K; m.C = cur.C; m.E = cur.E;
    m.flags = cur.flags; m.huber = cur.huber; m.min_rho = cur.min_rho; m.td_lo = cur.td_lo; m.td_hi = cur.td_hi;
    m.frame_id = cur.frame_id; m.pose = cur.pose; m.sb = cur.sb; m.pose_free = cur.pose_free; m.sb_width = cur.sb_width;
    std::copy_n(cur.rig, 7, m.rig);
    m.td = cur.td; m.cam = cur.cam;
    m.imu_from = cur.imu_from; m.imu_to = cur.imu_to; m.imu = cur.imu;
    m.prior_rows = cur.prior_rows; m.prior_cols = cur.prior_cols; m.prior_kind = cur.prior_kind;
    m.prior_id = cur.prior_id; m.prior_ref = cur.prior_ref; m.prior_A = cur.prior_A; m.prior_r0 = cur.prior_r0;
    m.uid_next = cur.uid_next;
    m.bulk = false;
    return m;
}

// Removes frame `drop` from the frame-level vectors (later frames move down by one).
void drop_frame(HostWindow& m, int drop) {
    m.K -= 1;
    m.frame_id.erase(m.frame_id.begin() + drop);
    m.pose.erase(m.pose.begin() + 7 * drop, m.pose.begin() + 7 * drop + 7);
    m.sb.erase(m.sb.begin() + 9 * drop, m.sb.begin() + 9 * drop + 9);
    m.pose_free.erase(m.pose_free.begin() + drop);
    m.sb_width.erase(m.sb_width.begin() + drop);
}

// The IMU edges e with keep[e] that do not touch frame `drop` (-1: none), renumbered as
// drop_frame renumbers the frames.
void keep_edges(HostWindow& m, const std::vector<int>& from, const std::vector<int>& to, const std::vector<double>& rec,
                const std::vector<char>& keep, int drop) {
    auto slot = [&](int k) { return drop >= 0 && k > drop ? k - 1 : k; };
    m.imu_from.clear(); m.imu_to.clear(); m.imu.clear();
    for (size_t e = 0; e < from.size(); ++e) {
        if (!keep[e] || from[e] == drop || to[e] == drop) continue;
        m.imu_from.push_back(slot(from[e]));
        m.imu_to.push_back(slot(to[e]));
        m.imu.insert(m.imu.end(), rec.begin() + (size_t)VIB_IMU_STRIDE * e, rec.begin() + (size_t)VIB_IMU_STRIDE * (e + 1));
    }
    m.E = (int)m.imu_from.size();
}

// Device buffers of the window edits of one slide or append (sized for the source window and
// nl inserted landmarks, nf inserted observations), carved from the context's grow-only edit
// buffer.
struct EditBufs {
    WinDev* src = nullptr;
    WinDev* dst = nullptr;
    int* bad = nullptr;  // kept anchors or observations of removed frames (must stay 0)
    int* sel = nullptr;  // 16 counters of the selection kernels
    unsigned long long *lmv = nullptr, *lscan = nullptr, *ltile = nullptr;
    uint32_t *cnt = nullptr, *cscan = nullptr, *cctile = nullptr;
    uint32_t *fv = nullptr, *fscan = nullptr, *ftile = nullptr;
    int32_t *lm_of = nullptr, *lmn = nullptr;
    uint32_t *counts = nullptr, *offs = nullptr, *ctile = nullptr;
    uint8_t *keep_lm = nullptr, *keep_f = nullptr, *marg_lm = nullptr, *cls = nullptr;
    int32_t *over_f = nullptr, *chosen = nullptr;
    double *over_rho = nullptr, *chosen_rho = nullptr;
    // inserted data
    int32_t* add_cnt = nullptr;
    int64_t* ins_obs_uid = nullptr;
    int32_t *ins_obs_lm = nullptr, *ins_obs_rank = nullptr, *ins_obs_fc = nullptr;
    double* ins_ob = nullptr;
    int32_t* ins_lm_fc = nullptr;
    double *ins_lmd = nullptr, *ins_rho = nullptr;
    uint8_t* ins_free = nullptr;
    int64_t* ins_uid = nullptr;
};

bool edit_buffers(void*& buf, size_t& cap, int L, int F, EditBufs& b, int nl = 0, int nf = 0) {
    size_t off = 0;
    auto take = [&](size_t bytes) {
        off = (off + 255) & ~size_t(255);
        const size_t o = off;
        off += std::max<size_t>(bytes, 1);
        return o;
    };
    const size_t Ls = L, Fs = F, Lx = (size_t)L + nl, Fx = (size_t)F + nf, NL = nl, NF = nf;
    const size_t lt = (size_t)vibk_edit_scan_tiles((int)Lx) + 1, ft = (size_t)vibk_edit_scan_tiles(F) + 1;
    const size_t nkeys = (size_t)vibk_edit_pair_keys() * vibk_edit_sort_blocks((int)Fx);
    const size_t ct = (size_t)vibk_edit_scan_tiles((int)nkeys) + 1;
    const size_t o_src = take(sizeof(WinDev)), o_dst = take(sizeof(WinDev)), o_bad = take(4), o_sel = take(64);
    const size_t o_lmv = take(8 * Lx), o_lscan = take(8 * (Lx + 1)), o_ltile = take(8 * lt);
    const size_t o_cnt = take(4 * Lx), o_cscan = take(4 * (Lx + 1)), o_cctile = take(4 * lt);
    const size_t o_fv = take(4 * Fs), o_fscan = take(4 * (Fs + 1)), o_ftile = take(4 * ft);
    const size_t o_lmof = take(4 * Fs), o_lmn = take(4 * Fx);
    const size_t o_counts = take(4 * (nkeys + 1)), o_offs = take(4 * (nkeys + 1)), o_ctile = take(4 * ct);
    const size_t o_klm = take(Ls), o_kf = take(Fs), o_mlm = take(Ls), o_cls = take(Ls);
    const size_t o_overf = take(4 * Ls), o_chosen = take(4 * Ls), o_overr = take(8 * Ls), o_chr = take(8 * Ls);
    const size_t o_add = take(4 * Lx), o_iou = take(8 * NF), o_iol = take(4 * NF), o_ior = take(4 * NF),
                 o_iof = take(4 * NF), o_iob = take(8 * OB_PLANES * NF), o_ilf = take(4 * NL),
                 o_ild = take(8 * LM_PLANES * NL), o_ilr = take(8 * NL), o_ilfr = take(NL), o_ilu = take(8 * NL);
    char* base = static_cast<char*>(grow_device(buf, cap, off));
    if (!base) return false;
    auto at = [&](size_t o) { return static_cast<void*>(base + o); };
    b.src = static_cast<WinDev*>(at(o_src));
    b.dst = static_cast<WinDev*>(at(o_dst));
    b.bad = static_cast<int*>(at(o_bad));
    b.sel = static_cast<int*>(at(o_sel));
    b.lmv = static_cast<unsigned long long*>(at(o_lmv));
    b.lscan = static_cast<unsigned long long*>(at(o_lscan));
    b.ltile = static_cast<unsigned long long*>(at(o_ltile));
    b.cnt = static_cast<uint32_t*>(at(o_cnt));
    b.cscan = static_cast<uint32_t*>(at(o_cscan));
    b.cctile = static_cast<uint32_t*>(at(o_cctile));
    b.fv = static_cast<uint32_t*>(at(o_fv));
    b.fscan = static_cast<uint32_t*>(at(o_fscan));
    b.ftile = static_cast<uint32_t*>(at(o_ftile));
    b.lm_of = static_cast<int32_t*>(at(o_lmof));
    b.lmn = static_cast<int32_t*>(at(o_lmn));
    b.counts = static_cast<uint32_t*>(at(o_counts));
    b.offs = static_cast<uint32_t*>(at(o_offs));
    b.ctile = static_cast<uint32_t*>(at(o_ctile));
    b.keep_lm = static_cast<uint8_t*>(at(o_klm));
    b.keep_f = static_cast<uint8_t*>(at(o_kf));
    b.marg_lm = static_cast<uint8_t*>(at(o_mlm));
    b.cls = static_cast<uint8_t*>(at(o_cls));
    b.over_f = static_cast<int32_t*>(at(o_overf));
    b.chosen = static_cast<int32_t*>(at(o_chosen));
    b.over_rho = static_cast<double*>(at(o_overr));
    b.chosen_rho = static_cast<double*>(at(o_chr));
    b.add_cnt = static_cast<int32_t*>(at(o_add));
    b.ins_obs_uid = static_cast<int64_t*>(at(o_iou));
    b.ins_obs_lm = static_cast<int32_t*>(at(o_iol));
    b.ins_obs_rank = static_cast<int32_t*>(at(o_ior));
    b.ins_obs_fc = static_cast<int32_t*>(at(o_iof));
    b.ins_ob = static_cast<double*>(at(o_iob));
    b.ins_lm_fc = static_cast<int32_t*>(at(o_ilf));
    b.ins_lmd = static_cast<double*>(at(o_ild));
    b.ins_rho = static_cast<double*>(at(o_ilr));
    b.ins_free = static_cast<uint8_t*>(at(o_ilfr));
    b.ins_uid = static_cast<int64_t*>(at(o_ilu));
    return true;
}

// The edit buffers of the context (one edit at a time) and of job i of a batch of slides.
bool edit_buffers(vib_context* ctx, int L, int F, EditBufs& b, int nl = 0, int nf = 0) {
    return edit_buffers(ctx->edit_buf, ctx->edit_cap, L, F, b, nl, nf);
}
bool job_edit_buffers(vib_context* ctx, size_t i, int L, int F, EditBufs& b) {
    while (ctx->edit_pool.size() <= i) ctx->edit_pool.emplace_back(nullptr, 0);
    return edit_buffers(ctx->edit_pool[i].first, ctx->edit_pool[i].second, L, F, b);
}

// Copies the slot's descriptor to b.src and clears the selection counters.
bool edit_begin(vib_context* ctx, const Slot& s, const EditBufs& b) {
    CopyBatch cb;
    cb.up(b.src, &s.dev, sizeof(WinDev));
    cb.zero(b.sel, 16 * sizeof(int));
    return cb.flush(ctx->stream, ctx->xfer);
}

// The window that `sp` selects from the resident window of `src`, made the resident window of
// `dst`. meta holds its frame-level part (frame_meta, then the caller's changes); its landmark
// and observation counts, free inverse ranges and work lists follow from the device scans. With
// dst == src (a slide in place) the edited window is built in the context's spare arena and
// then takes the slot's place; a failure leaves src as it was. `budgeted`: the result counts
// against the context's HBM budget like an upload to dst.
int edit_window(vib_context* ctx, Slot& src, EditSpec sp, HostWindow&& meta, Slot& dst, const EditBufs& b, bool budgeted) {
    std::string err;
    if (!meta.layout_states(err)) return fail(ctx, VIB_ERR_INVALID_ARGUMENT, err);
    sp.nr = meta.nr;
    const int Ls = src.host.L, Fs = src.host.F, Lx = Ls + sp.ins_lm;
    const hipStream_t st = ctx->stream;
    unsigned long long ltot = 0;
    uint32_t ftot = 0;
    int bad = 0;
    CopyBatch cb;
    cb.up(b.src, &src.dev, sizeof(WinDev));
    cb.zero(b.bad, sizeof(int));
    bool ok = cb.flush(st, ctx->xfer) &&
              vibk_edit_mark(b.src, Ls, &sp, b.lmv, b.fv, b.cnt, b.lm_of, b.bad, st) == hipSuccess &&
              vibk_edit_scan(b.lmv, Lx, b.ltile, b.lscan, 1, st) == hipSuccess &&
              vibk_edit_scan(b.cnt, Lx, b.cctile, b.cscan, 0, st) == hipSuccess &&
              vibk_edit_scan(b.fv, Fs, b.ftile, b.fscan, 0, st) == hipSuccess;
    cb.down(&ltot, b.lscan + Lx, sizeof(ltot));
    cb.down(&ftot, b.cscan + Lx, sizeof(ftot));
    cb.down(&bad, b.bad, sizeof(bad));
    ok = ok && cb.flush(st, ctx->xfer);
    if (!ok) {
        (void)hipGetLastError();
        return fail(ctx, VIB_ERR_DEVICE, "window edit failed");
    }
    if (bad) return fail(ctx, VIB_ERR_INVALID_ARGUMENT, "window edit keeps an anchor or observation of a removed frame");
    meta.L = (int)(uint32_t)ltot;
    meta.ne = (int)(ltot >> 32);
    meta.F = (int)ftot;
    meta.bulk = false;
    meta.lm_frame.clear(); meta.lm_cam.clear(); meta.lm_off.clear(); meta.lm_free.clear(); meta.lmd.clear();
    meta.rho.clear(); meta.lm_uid.clear(); meta.ob_fc.clear(); meta.ob_lm.clear(); meta.ob.clear(); meta.lm_col.clear();
    meta.pf_ob.clear(); meta.pf_an.clear(); meta.pf_ids.clear(); meta.pf_lm.clear(); meta.pf_pos.clear();
    meta.make_chunks();
    const size_t need = Slot::footprint(meta);
    if (budgeted && ctx->hbm_budget > 0) {
        size_t resident = 0;
        for (const auto& kv : ctx->slots)
            if (kv.second && kv.second.get() != &dst && kv.second->used) resident += kv.second->bytes;
        if (resident + need > ctx->hbm_budget)
            return fail(ctx, VIB_ERR_OUT_OF_MEMORY, "edited window does not fit the context's HBM budget");
    }
    const bool in_place = &dst == &src;
    char* base = nullptr;
    if (in_place) {
        base = static_cast<char*>(grow_device(ctx->spare, ctx->spare_cap, need));
        if (!base) return fail(ctx, VIB_ERR_OUT_OF_MEMORY, "window edit allocation failed");
    } else {
        dst.used = false;
        if (!dst.arena || need > dst.capacity) {
            // the marginal systems of many sessions pass through one slot: room to grow
            const size_t want = std::max(need, dst.capacity + dst.capacity / 2);
            dst.release();
            if (hipMalloc(&dst.arena, want) != hipSuccess) {
                dst.arena = nullptr;
                (void)hipGetLastError();
                return fail(ctx, VIB_ERR_OUT_OF_MEMORY, "window edit allocation failed");
            }
            dst.capacity = want;
        }
        base = static_cast<char*>(dst.arena);
    }
    WinDev nd;
    SlotCarve x;
    if (!Slot::carve(meta, base, need, nd, x)) return fail(ctx, VIB_ERR_DEVICE, "arena overflow");
    const int Fd = meta.F, nkeys = vibk_edit_pair_keys() * vibk_edit_sort_blocks(Fd);
    cb.up(b.dst, &nd, sizeof(WinDev));
    ok = Slot::fill(meta, nd, x, base, need, false, st, ctx->xfer, cb) &&
         vibk_edit_gather(b.src, b.dst, Ls, Fs, &sp, b.lscan, b.fscan, b.cscan, b.lm_of, b.lmn, st) == hipSuccess &&
         (Fd == 0 || (vibk_edit_pair_hist(b.dst, Fd, b.lmn, b.counts, st) == hipSuccess &&
                      vibk_edit_scan(b.counts, nkeys, b.ctile, b.offs, 0, st) == hipSuccess &&
                      vibk_edit_pair_scatter(b.dst, Fd, b.lmn, b.offs, st) == hipSuccess)) &&
         hipStreamSynchronize(st) == hipSuccess;
    if (!ok) {
        (void)hipGetLastError();
        return fail(ctx, VIB_ERR_DEVICE, "window edit failed");
    }
    if (in_place) {
        // The edited window is copied into the slot's arena in stream order, at the offsets the
        // spare arena had (the layout depends only on the window), so the spare stays the
        // context's largest. A window that outgrew its arena (an append) first gets a new one
        // with a quarter of headroom, which the next appends fit.
        if (need > dst.capacity) {
            void* grown = nullptr;
            const size_t want = need + need / 4;
            if (hipMalloc(&grown, want) != hipSuccess) {
                (void)hipGetLastError();
                return fail(ctx, VIB_ERR_OUT_OF_MEMORY, "window edit allocation failed");
            }
            (void)hipFree(dst.arena);
            dst.arena = grown;
            dst.capacity = want;
        }
        WinDev nd2;
        SlotCarve x2;
        if (!Slot::carve(meta, static_cast<char*>(dst.arena), dst.capacity, nd2, x2))
            return fail(ctx, VIB_ERR_DEVICE, "arena overflow");
        if (hipMemcpyAsync(dst.arena, base, need, hipMemcpyDeviceToDevice, st) != hipSuccess) {
            (void)hipGetLastError();
            dst.used = false;  // the arena may hold part of the edited window
            return fail(ctx, VIB_ERR_DEVICE, "window edit failed");
        }
        nd = nd2;
        x = x2;
    }
    dst.dev = nd;
    dst.reduced_lin = x.reduced_lin;
    dst.reduced_schur = x.reduced_schur;
    dst.host = std::move(meta);
    dst.bytes = need;
    dst.host_set = false;
    dst.used = true;
    return VIB_OK;
}

// Squared whitened residuals (NIS) of two IMU edges of a slot at its resident state.
bool imu_nis(vib_context* ctx, Slot& s, const int edges[2], double nis[2]) {
    LmState lm{};
    lm.lambda = 1e-4; lm.increase = 2; lm.new_point = 1; lm.status = 0; lm.max_iterations = 1;
    std::vector<Slot*> one{&s};
    Batch& b = ctx->one;
    b.mixed = 0;
    std::string err;
    std::vector<double> r((size_t)15 * s.host.E);
    if (!write_lm(ctx, s, lm) || !b.build(one, ctx->stream, ctx->xfer, err) || !b.launch(1, ctx->stream, 0, 0) ||
        hipMemcpyAsync(r.data(), s.dev.imu_r, r.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        return false;
    for (int i = 0; i < 2; ++i) {
        double v = 0;
        for (int m = 0; m < 15; ++m) v += r[(size_t)15 * edges[i] + m] * r[(size_t)15 * edges[i] + m];
        nis[i] = v;
    }
    return true;
}

// Copy of a resident window in the C-ABI layout (vib_read_window); view() points a vib_window
// into it.
struct WindowCopy {
    int K = 0, C = 0, L = 0, F = 0, E = 0;
    uint32_t flags = 0;
    double huber = 1, min_rho = 1e-8, td_lo = -1e300, td_hi = 1e300;
    std::vector<int64_t> frame_id;
    std::vector<double> pose, sb;
    std::vector<uint8_t> pose_free, sb_width;
    double rig[7]{}, td = 0;
    std::vector<double> cam_R, cam_t;
    std::vector<int32_t> lm_frame, lm_cam;
    std::vector<double> lm_bearing, lm_vel, lm_tref, rho;
    std::vector<uint8_t> lm_free;
    std::vector<int32_t> ob_lm, ob_frame, ob_cam;
    std::vector<double> ob_bearing, ob_vel, ob_tref, ob_sqrt;
    std::vector<int32_t> imu_from, imu_to;
    std::vector<double> imu;
    int prior_rows = 0;
    std::vector<int32_t> prior_kind;
    std::vector<int64_t> prior_id;
    std::vector<double> prior_ref, prior_A, prior_r0;

    vib_window view() {
        vib_window w{};
        w.num_frames = K; w.num_cameras = C; w.num_landmarks = L; w.num_observations = F; w.num_imu = E;
        w.flags = flags; w.huber_delta = huber; w.min_inverse_depth = min_rho; w.td_lower = td_lo; w.td_upper = td_hi;
        w.frame_id = frame_id.data(); w.pose = pose.data(); w.speed_bias = sb.data();
        w.pose_free = pose_free.data(); w.sb_width = sb_width.data();
        w.rig = rig; w.td = &td; w.cam_rotation = cam_R.data(); w.cam_translation = cam_t.data();
        w.lm_anchor_frame = lm_frame.data(); w.lm_anchor_cam = lm_cam.data(); w.lm_bearing = lm_bearing.data();
        w.lm_velocity = lm_vel.data(); w.lm_tref = lm_tref.data(); w.lm_rho = rho.data(); w.lm_free = lm_free.data();
        w.obs_landmark = ob_lm.data(); w.obs_frame = ob_frame.data(); w.obs_cam = ob_cam.data();
        w.obs_bearing = ob_bearing.data(); w.obs_velocity = ob_vel.data(); w.obs_tref = ob_tref.data();
        w.obs_sqrt_info = ob_sqrt.data();
        w.imu_from = imu_from.data(); w.imu_to = imu_to.data(); w.imu_data = imu.data();
        w.prior_rows = prior_rows; w.prior_blocks = (int32_t)prior_kind.size();
        w.prior_key_kind = prior_kind.data(); w.prior_key_id = prior_id.data(); w.prior_ref = prior_ref.data();
        w.prior_A = prior_A.data(); w.prior_r0 = prior_r0.data();
        return w;
    }
};

// The resident window as a copy (observations in landmark order).
WindowCopy copy_of(const HostWindow& h) {
    WindowCopy w;
    w.K = h.K; w.C = h.C; w.L = h.L; w.F = h.F; w.E = h.E;
    w.flags = h.flags; w.huber = h.huber; w.min_rho = h.min_rho; w.td_lo = h.td_lo; w.td_hi = h.td_hi;
    w.frame_id = h.frame_id; w.pose = h.pose; w.sb = h.sb; w.pose_free = h.pose_free; w.sb_width = h.sb_width;
    std::copy_n(h.rig, 7, w.rig);
    w.td = h.td;
    for (int c = 0; c < h.C; ++c) {
        w.cam_R.insert(w.cam_R.end(), h.cam.begin() + 12 * c, h.cam.begin() + 12 * c + 9);
        w.cam_t.insert(w.cam_t.end(), h.cam.begin() + 12 * c + 9, h.cam.begin() + 12 * c + 12);
    }
    w.lm_frame = h.lm_frame; w.lm_cam = h.lm_cam; w.rho = h.rho; w.lm_free = h.lm_free;
    const size_t L = h.L, F = h.F;
    w.lm_bearing.reserve(3 * L); w.lm_vel.reserve(3 * L); w.lm_tref.reserve(L);
    w.ob_lm.reserve(F); w.ob_frame.reserve(F); w.ob_cam.reserve(F); w.ob_tref.reserve(F);
    w.ob_bearing.reserve(3 * F); w.ob_vel.reserve(3 * F); w.ob_sqrt.reserve(3 * F);
    for (size_t l = 0; l < L; ++l) {
        for (int i = 0; i < 3; ++i) {
            w.lm_bearing.push_back(h.lmd[(LM_BX + i) * L + l]);
            w.lm_vel.push_back(h.lmd[(LM_VX + i) * L + l]);
        }
        w.lm_tref.push_back(h.lmd[LM_T * L + l]);
    }
    for (size_t f = 0; f < F; ++f) {
        w.ob_lm.push_back(h.ob_lm[f]); w.ob_frame.push_back(h.ob_fc[f] >> 8); w.ob_cam.push_back(h.ob_fc[f] & 255);
        for (int i = 0; i < 3; ++i) {
            w.ob_bearing.push_back(h.ob[(OB_BX + i) * F + f]);
            w.ob_vel.push_back(h.ob[(OB_VX + i) * F + f]);
            w.ob_sqrt.push_back(h.ob[(OB_S00 + i) * F + f]);
        }
        w.ob_tref.push_back(h.ob[OB_T * F + f]);
    }
    w.imu_from = h.imu_from; w.imu_to = h.imu_to; w.imu = h.imu;
    w.prior_rows = h.prior_rows; w.prior_kind = h.prior_kind; w.prior_id = h.prior_id;
    w.prior_ref = h.prior_ref; w.prior_A = h.prior_A; w.prior_r0 = h.prior_r0;
    return w;
}

// Eigendecompositions of symmetric matrices on the device in one launch (k_sym_eigen, one
// workgroup each): eigenvalues ascending, eigenvectors as the columns of U[j] (row-major).
// status[j]: VIB_OK or VIB_ERR_EIGEN_FAILED (no convergence, or an eigenvalue not finite). The
// return value is VIB_OK or the failure of the launch as a whole (allocation, device).
int device_eigen_batch(vib_context* ctx, const std::vector<const std::vector<double>*>& M, const std::vector<int>& n,
                       std::vector<std::vector<double>>& ev, std::vector<std::vector<double>>& U, std::vector<int>& status) {
    const size_t J = M.size();
    ev.assign(J, {});
    U.assign(J, {});
    status.assign(J, VIB_OK);
    if (J == 0) return VIB_OK;
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    std::vector<size_t> off(J);
    size_t bytes = al(sizeof(EigenJob) * J);
    int max_n = 0;
    for (size_t j = 0; j < J; ++j) {
        if (n[j] < 1 || n[j] > vibk_sym_eigen_max())
            return fail(ctx, VIB_ERR_UNSUPPORTED_DIMENSIONS, "eigenproblem beyond the device eigensolver");
        max_n = std::max(max_n, n[j]);
        off[j] = bytes;
        const size_t nn = (size_t)n[j] * n[j];
        bytes += al(8 * (2 * nn + n[j] + vibk_sym_eigen_log(n[j])) + 8);
    }
    // jobs | per job: A | V | eigenvalues | rotation log | convergence flag, in the context's
    // grow-only eigen buffer
    char* buf = static_cast<char*>(grow_device(ctx->eig_buf, ctx->eig_cap, bytes));
    if (!buf) return fail(ctx, VIB_ERR_OUT_OF_MEMORY, "eigensolver buffer allocation failed");
    std::vector<EigenJob> jobs(J);
    std::vector<std::vector<double>> V(J), e(J);
    std::vector<int32_t> conv(J, 0);
    bool ok = true;
    for (size_t j = 0; j < J; ++j) {
        const size_t nn = (size_t)n[j] * n[j];
        EigenJob& q = jobs[j];
        q.A = reinterpret_cast<double*>(buf + off[j]);
        q.V = q.A + nn;
        q.ev = q.V + nn;
        q.rot = q.ev + n[j];
        q.ok = reinterpret_cast<int32_t*>(q.rot + vibk_sym_eigen_log(n[j]));
        q.n = q.ld = n[j];
        V[j].resize(nn);
        e[j].resize(n[j]);
    }
    CopyBatch cb;  // matrices and job list in, eigenvectors, eigenvalues and flags out: one transfer each way
    for (size_t j = 0; j < J; ++j) cb.up(jobs[j].A, M[j]->data(), 8 * (size_t)n[j] * n[j]);
    cb.up(buf, jobs.data(), sizeof(EigenJob) * J);
    const EigenJob* djobs = reinterpret_cast<const EigenJob*>(buf);
    ok = cb.flush(ctx->stream, ctx->xfer) && vibk_sym_eigen_batch(djobs, (int)J, max_n, ctx->stream) == hipSuccess;
    for (size_t j = 0; j < J; ++j) {
        cb.down(V[j].data(), jobs[j].V, 8 * V[j].size());
        cb.down(e[j].data(), jobs[j].ev, 8 * (size_t)n[j]);
        cb.down(&conv[j], jobs[j].ok, sizeof(int32_t));
    }
    ok = ok && cb.flush(ctx->stream, ctx->xfer);
    if (!ok) return fail(ctx, VIB_ERR_DEVICE, "device eigendecomposition failed");
    for (size_t j = 0; j < J; ++j) {
        const int m = n[j];
        std::vector<int> order(m);
        for (int i = 0; i < m; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return e[j][a] < e[j][b]; });
        ev[j].resize(m);
        U[j].resize((size_t)m * m);
        for (int c = 0; c < m; ++c) {  // k_sym_eigen returns eigenvector c as row c of V
            ev[j][c] = e[j][order[c]];
            for (int r = 0; r < m; ++r) U[j][(size_t)r * m + c] = V[j][(size_t)order[c] * m + r];
        }
        bool finite = conv[j] != 0;
        for (double v : ev[j]) finite = finite && std::isfinite(v);
        if (!finite) status[j] = VIB_ERR_EIGEN_FAILED;
    }
    return VIB_OK;
}

struct SchurPrior {
    int rows = 0, cols = 0, eliminated = 0, landmarks = 0, consumed = 0;
    double cutoff = 0;
    std::vector<int32_t> kind;
    std::vector<int64_t> id;
    std::vector<double> ref, A, r0;
};

// One marginalization of a batch (Marginalizer::compute, Marginalizer.cpp:284-474): slot
// `victim` (pose, speed/bias) of the window of `src` with the landmarks of d_sel (device flags;
// all their factors) and the IMU edges (from, to, record) eliminated into a square-root prior
// on the separator; the previous prior always takes part. frames / observed: the frames the
// selected landmarks touch and whether any of them has an observation (k_touch_frames).
struct PriorJob {
    Slot* src = nullptr;
    const EditBufs* eb = nullptr;
    const uint8_t* d_sel = nullptr;
    int frames = 0;
    bool observed = false;
    int victim = 0;
    std::vector<int> e_from, e_to;
    std::vector<double> e_rec;
    Slot* marg = nullptr;  // the marginal system's slot
    int status = VIB_OK;
    std::string error;
    std::vector<int> victim_cols, kept_cols;
    std::vector<double> S, rhs, Hp, bp;
    SchurPrior out;
};

void job_fail(vib_context* ctx, PriorJob& j, int code, const std::string& msg) {
    j.status = code;
    j.error = msg;
    ctx->error = msg;
}

// The marginalizations of a batch of windows: each marginal system is a device edit of its
// window (consumed factors only, every touched block a variable); they linearize, eliminate
// their landmarks through the diagonal pseudo-inverse and assemble their reduced systems in
// one batch of launches; the eliminated state blocks' eigen pseudo-inverses (cutoff
// max(1e-12, 1e-9 lambda_max), Marginalizer::psdSolve, Marginalizer.cpp:240-283) run as one
// eigen launch, and so do the priors' eigen square roots (A = sqrt(lambda) u^T, r0 =
// u^T b / sqrt(lambda) above the same cutoff; indefinite below -10 cutoff,
// Marginalizer.cpp:430-466). A job's failure is its status; the return value is a failure of
// the batch as a whole.
int schur_priors(vib_context* ctx, std::vector<PriorJob>& jobs) {
    // 1. the marginal systems
    for (PriorJob& j : jobs) {
        const HostWindow& cur = j.src->host;
        const int K = cur.K;
        std::vector<char> tpose(K, 0), tsb(K, 0);
        bool rig_used = j.observed, td_used = j.observed;
        for (int k = 0; k < K; ++k) if ((j.frames >> k) & 1) tpose[k] = 1;
        for (size_t e = 0; e < j.e_from.size(); ++e) {
            tpose[j.e_from[e]] = tpose[j.e_to[e]] = 1;
            tsb[j.e_from[e]] = tsb[j.e_to[e]] = 1;
        }
        for (size_t b = 0; b < cur.prior_kind.size(); ++b) {
            int slot = -1;
            for (int k = 0; k < K; ++k) if (cur.frame_id[k] == cur.prior_id[b]) slot = k;
            switch (cur.prior_kind[b]) {
                case VIB_KEY_POSE: if (slot >= 0) tpose[slot] = 1; break;
                case VIB_KEY_SPEED_BIAS: if (slot >= 0) tsb[slot] = 1; break;
                case VIB_KEY_RIG: rig_used = true; break;
                default: td_used = true;
            }
        }
        HostWindow mm = frame_meta(cur);
        mm.flags = (rig_used ? VIB_RIG_FREE : 0u) | (td_used ? VIB_TD_FREE : 0u);
        mm.td_lo = -1e300; mm.td_hi = 1e300;
        mm.pose_free.assign(tpose.begin(), tpose.end());
        mm.sb_width.assign(K, 0);
        for (int k = 0; k < K; ++k) mm.sb_width[k] = tsb[k] ? 9 : 0;
        mm.imu_from = j.e_from; mm.imu_to = j.e_to; mm.imu = j.e_rec;
        mm.E = (int)j.e_from.size();
        EditSpec sp{};
        for (int k = 0; k < kMaxFrames; ++k) sp.renum[k] = k < K ? k : -1;
        sp.lm_mode = 1;  // every eliminated inverse range is a variable
        sp.keep_lm = j.d_sel;
        if (const int st = edit_window(ctx, *j.src, sp, std::move(mm), *j.marg, *j.eb, false)) {
            job_fail(ctx, j, st, ctx->error);
            continue;
        }
        const HostWindow& m = j.marg->host;
        std::vector<char> is_victim(m.nr, 0);
        if (m.pose_col[j.victim] >= 0) for (int i = 0; i < 6; ++i) j.victim_cols.push_back(m.pose_col[j.victim] + i);
        if (m.sb_col[j.victim] >= 0) for (int i = 0; i < 9; ++i) j.victim_cols.push_back(m.sb_col[j.victim] + i);
        for (int c : j.victim_cols) is_victim[c] = 1;
        for (int c = 0; c < m.nr; ++c) if (!is_victim[c]) j.kept_cols.push_back(c);
        if (j.victim_cols.empty()) job_fail(ctx, j, VIB_ERR_INVALID_ARGUMENT, "no eliminated state block");
    }
    // 2. linearization, landmark elimination and reduced systems, batched
    std::vector<PriorJob*> live;
    for (PriorJob& j : jobs) if (j.status == VIB_OK) live.push_back(&j);
    if (live.empty()) return VIB_OK;
    LmState lm{};
    lm.lambda = 0; lm.increase = 2; lm.new_point = 1; lm.status = 0; lm.max_iterations = 1; lm.evaluate_only = 1;
    lm.marginal = 1;
    std::vector<Slot*> slots;
    for (PriorJob* j : live) {
        if (!write_lm(ctx, *j->marg, lm)) return fail(ctx, VIB_ERR_DEVICE, "marginal LM state upload failed");
        slots.push_back(j->marg);
    }
    Batch& b = ctx->one;
    b.mixed = 0;
    std::string err;
    if (!b.build(slots, ctx->stream, ctx->xfer, err) || !b.linearize(ctx->stream))
        return fail(ctx, VIB_ERR_DEVICE, "marginal linearization failed");
    std::vector<std::vector<double>> dvals(live.size());
    CopyBatch cb;
    for (size_t i = 0; i < live.size(); ++i) {
        const Slot& t = *live[i]->marg;
        dvals[i].resize(t.host.L);
        cb.down(dvals[i].data(), t.dev.lm_out + (size_t)LO_D * t.host.L, 8 * (size_t)t.host.L);
    }
    if (!cb.flush(ctx->stream, ctx->xfer)) return fail(ctx, VIB_ERR_DEVICE, "marginal read-back failed");
    std::vector<LmState> l2(live.size(), lm);
    for (size_t i = 0; i < live.size(); ++i) {
        double dmax = 0;
        for (double d : dvals[i]) dmax = std::max(dmax, d);
        l2[i].marg_cut = std::max(1e-12, 1e-9 * dmax);
        cb.up(live[i]->marg->dev.lm, &l2[i], sizeof(LmState));
    }
    bool ok = cb.flush(ctx->stream, ctx->xfer) && b.launch(3, ctx->stream, 0, 0) && b.launch(4, ctx->stream, 0, 0) &&
              b.launch(5, ctx->stream, 0, 0);
    std::vector<LmState> chk(live.size());
    for (size_t i = 0; i < live.size(); ++i) {
        PriorJob* j = live[i];
        const Slot& t = *j->marg;
        const size_t nr = t.host.nr;
        j->S.resize(nr * nr);
        j->rhs.resize(nr);
        cb.down(j->S.data(), t.dev.Scopy, 8 * nr * nr);
        cb.down(j->rhs.data(), t.dev.rhs, 8 * nr);
        cb.down(&chk[i], t.dev.lm, sizeof(LmState));
    }
    if (!ok || !cb.flush(ctx->stream, ctx->xfer)) return fail(ctx, VIB_ERR_DEVICE, "marginal Schur failed");
    for (size_t i = 0; i < live.size(); ++i)
        if (!chk[i].lin_valid) job_fail(ctx, *live[i], VIB_ERR_INVALID_RESIDUAL, "invalid residual in a consumed factor");
    // 3. the eliminated state blocks' eigen pseudo-inverses; H_prior = S_rr - S_rm S_mm^+ S_mr
    std::vector<PriorJob*> stage;
    std::vector<std::vector<double>> mats;
    for (PriorJob* j : live) {
        if (j->status != VIB_OK) continue;
        const int nv = (int)j->victim_cols.size(), nr = j->marg->host.nr;
        std::vector<double> Smm((size_t)nv * nv);
        for (int a = 0; a < nv; ++a)
            for (int c = 0; c < nv; ++c)
                Smm[(size_t)a * nv + c] = 0.5 * (j->S[(size_t)j->victim_cols[a] * nr + j->victim_cols[c]] +
                                                  j->S[(size_t)j->victim_cols[c] * nr + j->victim_cols[a]]);
        stage.push_back(j);
        mats.push_back(std::move(Smm));
    }
    std::vector<const std::vector<double>*> mp;
    std::vector<int> ns;
    for (size_t i = 0; i < stage.size(); ++i) {
        mp.push_back(&mats[i]);
        ns.push_back((int)stage[i]->victim_cols.size());
    }
    std::vector<std::vector<double>> evs, Us;
    std::vector<int> est;
    if (const int st = device_eigen_batch(ctx, mp, ns, evs, Us, est)) return st;
    for (size_t i = 0; i < stage.size(); ++i) {
        PriorJob& j = *stage[i];
        if (est[i] != VIB_OK) {
            job_fail(ctx, j, est[i], "eigensolver did not converge");
            continue;
        }
        const int nv = ns[i], np = (int)j.kept_cols.size(), nr = j.marg->host.nr;
        const std::vector<double>& ev = evs[i];
        const std::vector<double>& U = Us[i];
        const double cut_v = std::max(1e-12, 1e-9 * std::max(0.0, ev[nv - 1]));
        if (ev[0] < -10.0 * cut_v) {
            job_fail(ctx, j, VIB_ERR_INDEFINITE, "Schur prior is indefinite");
            continue;
        }
        std::vector<double> Z, zb;  // U_k^T S_mr / sqrt(lambda_k), U_k^T b_m / sqrt(lambda_k)
        for (int q = 0; q < nv; ++q) {
            if (!(ev[q] > cut_v)) continue;
            const double is = 1.0 / std::sqrt(ev[q]);
            double bz = 0;
            for (int a = 0; a < nv; ++a) bz += U[(size_t)a * nv + q] * -j.rhs[j.victim_cols[a]];
            zb.push_back(bz * is);
            for (int c = 0; c < np; ++c) {
                double v = 0;
                for (int a = 0; a < nv; ++a) v += U[(size_t)a * nv + q] * j.S[(size_t)j.victim_cols[a] * nr + j.kept_cols[c]];
                Z.push_back(v * is);
            }
        }
        const int nk = (int)zb.size();
        j.Hp.assign((size_t)np * np, 0);
        j.bp.assign(np, 0);
        for (int a = 0; a < np; ++a) {
            double g = -j.rhs[j.kept_cols[a]];
            for (int q = 0; q < nk; ++q) g -= Z[(size_t)q * np + a] * zb[q];
            j.bp[a] = g;
            for (int c = 0; c <= a; ++c) {
                double h = 0.5 * (j.S[(size_t)j.kept_cols[a] * nr + j.kept_cols[c]] +
                                  j.S[(size_t)j.kept_cols[c] * nr + j.kept_cols[a]]);
                for (int q = 0; q < nk; ++q) h -= Z[(size_t)q * np + a] * Z[(size_t)q * np + c];
                j.Hp[(size_t)a * np + c] = j.Hp[(size_t)c * np + a] = h;
            }
        }
    }
    // 4. the priors' eigen square roots
    stage.clear();
    mp.clear();
    ns.clear();
    for (PriorJob* j : live)
        if (j->status == VIB_OK) {
            stage.push_back(j);
            mp.push_back(&j->Hp);
            ns.push_back((int)j->kept_cols.size());
        }
    if (const int st = device_eigen_batch(ctx, mp, ns, evs, Us, est)) return st;
    for (size_t i = 0; i < stage.size(); ++i) {
        PriorJob& j = *stage[i];
        if (est[i] != VIB_OK) {
            job_fail(ctx, j, est[i], "eigensolver did not converge");
            continue;
        }
        const int np = ns[i];
        const std::vector<double>& pev = evs[i];
        const std::vector<double>& PU = Us[i];
        const double cut = std::max(1e-12, 1e-9 * std::max(0.0, pev[np - 1]));
        if (pev[0] < -10.0 * cut) {
            job_fail(ctx, j, VIB_ERR_INDEFINITE, "Schur prior is indefinite");
            continue;
        }
        SchurPrior& out = j.out;
        out.cols = np;
        out.rows = 0;
        for (int q = 0; q < np; ++q) {
            if (!(pev[q] > cut)) continue;
            const double root = std::sqrt(pev[q]);
            double ub = 0;
            for (int k = 0; k < np; ++k) {
                out.A.push_back(root * PU[(size_t)k * np + q]);
                ub += PU[(size_t)k * np + q] * j.bp[k];
            }
            out.r0.push_back(ub / root);
            ++out.rows;
        }
        // prior keys in column order: poses, rig, td, speed/bias (the victim excluded)
        const HostWindow& cur = j.src->host;
        const HostWindow& m = j.marg->host;
        for (int k = 0; k < cur.K; ++k)
            if (k != j.victim && m.pose_col[k] >= 0) {
                out.kind.push_back(VIB_KEY_POSE); out.id.push_back(cur.frame_id[k]);
                out.ref.insert(out.ref.end(), cur.pose.begin() + 7 * k, cur.pose.begin() + 7 * k + 7);
            }
        if (m.rig_col >= 0) {
            out.kind.push_back(VIB_KEY_RIG); out.id.push_back(0);
            out.ref.insert(out.ref.end(), cur.rig, cur.rig + 7);
        }
        if (m.td_col >= 0) {
            out.kind.push_back(VIB_KEY_TD); out.id.push_back(0);
            out.ref.push_back(cur.td);
        }
        for (int k = 0; k < cur.K; ++k)
            if (k != j.victim && m.sb_col[k] >= 0) {
                out.kind.push_back(VIB_KEY_SPEED_BIAS); out.id.push_back(cur.frame_id[k]);
                out.ref.insert(out.ref.end(), cur.sb.begin() + 9 * k, cur.sb.begin() + 9 * k + 9);
            }
        out.eliminated = (int)j.victim_cols.size() + m.L;
        out.landmarks = m.L;
        out.consumed = m.F + m.E + (cur.prior_rows ? 1 : 0);
        out.cutoff = cut;
    }
    return VIB_OK;
}

// The marginal slot of job i (grow-only pool of the context).
Slot& marginal_slot(vib_context* ctx, size_t i) {
    while (ctx->marginal_pool.size() <= i) ctx->marginal_pool.push_back(std::make_unique<Slot>());
    return *ctx->marginal_pool[i];
}

// One marginalization (a batch of one).
int schur_prior(vib_context* ctx, Slot& src, const EditBufs& eb, const uint8_t* d_sel, int frames, bool observed,
                int victim, const std::vector<int>& e_from, const std::vector<int>& e_to, const std::vector<double>& e_rec,
                SchurPrior& out) {
    std::vector<PriorJob> jobs(1);
    PriorJob& j = jobs[0];
    j.src = &src; j.eb = &eb; j.d_sel = d_sel; j.frames = frames; j.observed = observed; j.victim = victim;
    j.e_from = e_from; j.e_to = e_to; j.e_rec = e_rec;
    j.marg = &marginal_slot(ctx, 0);
    if (const int st = schur_priors(ctx, jobs)) return st;
    if (j.status != VIB_OK) return fail(ctx, j.status, j.error);
    out = std::move(j.out);
    return VIB_OK;
}

// The prior of an edited window: the square root of a marginalization (no rows: no prior).
void set_prior(HostWindow& m, const SchurPrior& p) {
    m.prior_rows = p.rows;
    if (p.rows == 0) {
        m.prior_cols = 0;
        m.prior_kind.clear(); m.prior_id.clear(); m.prior_ref.clear(); m.prior_A.clear(); m.prior_r0.clear();
        return;
    }
    m.prior_cols = p.cols;
    m.prior_kind = p.kind; m.prior_id = p.id; m.prior_ref = p.ref; m.prior_A = p.A; m.prior_r0 = p.r0;
}

// The frames of the landmarks selected by d_sel (k_touch_frames): bit mask and whether any of
// them has an observation.
bool touched_frames(vib_context* ctx, const EditBufs& b, int L, const uint8_t* d_sel, int& frames, bool& observed) {
    int out[2] = {0, 0};
    CopyBatch cb;
    const bool ok = vibk_touch_frames(b.src, L, d_sel, b.sel + 8, ctx->stream) == hipSuccess &&
                    (cb.down(out, b.sel + 8, sizeof(out)), cb.flush(ctx->stream, ctx->xfer));
    frames = out[0];
    observed = out[1] != 0;
    return ok;
}

// Linear interpolation of the gyroscope at t (gyroAt, BodyImuFactor.cpp:29-38).
void gyro_at(const double* samples, int count, double t, double* out) {
    int next = 0;
    while (next < count && samples[7 * next] < t) ++next;
    if (next == 0 || next == count) {
        const double* g = &samples[7 * (next == 0 ? 0 : count - 1) + 1];
        for (int i = 0; i < 3; ++i) out[i] = g[i];
        return;
    }
    const double t0 = samples[7 * (next - 1)], a = (t - t0) / (samples[7 * next] - t0);
    for (int i = 0; i < 3; ++i) out[i] = (1 - a) * samples[7 * (next - 1) + 1 + i] + a * samples[7 * next + 1 + i];
}

// SECOND_NEW, the redundant-frame path (SlidingWindowManager.cpp:263-482, 587-680): the two IMU
// edges of slot K-2 become one re-integrated edge (mergeBodyImuRecords, BodyImuFactor.cpp:306-388),
// landmarks anchored in K-2 are re-anchored on a surviving view, demoted (one surviving ray) or
// retired into the prior, the slot's other image measurements are discarded, and the prior is
// reused when nothing else touches the slot.
int preintegrate_locked(vib_context* ctx, int32_t count, const double* samples, int32_t num_samples,
                        const int32_t* sample_begin, const int32_t* sample_end, const double* t0, const double* t1,
                        const double* speed_bias, const vib_imu_params* params, const double* gyro_start,
                        const double* gyro_end, double* out, int32_t* status, double* kernel_ms);

// The keyframe as the window's newest frame: its state and IMU edge join the frame-level part
// on the host; its landmarks and observations are inserted into the resident streams by a device
// edit (observations sorted by landmark id here, resolved to resident landmarks on the device).
int append_keyframe(vib_context* ctx, Slot& s, const vib_keyframe& kf, int32_t* unmatched) {
    const int BAD = VIB_ERR_INVALID_ARGUMENT;
    const HostWindow& cur = s.host;
    if (cur.K >= kMaxFrames) return fail(ctx, VIB_ERR_UNSUPPORTED_DIMENSIONS, "unsupported Schur dimensions: the window holds 20 frames");
    const int nl = kf.num_landmarks, nf = kf.num_observations;
    if (nl < 0 || nf < 0 || !kf.pose || !kf.speed_bias) return fail(ctx, BAD, "incomplete keyframe");
    if (nl && (!kf.lm_id || !kf.lm_anchor_cam || !kf.lm_bearing || !kf.lm_velocity || !kf.lm_tref || !kf.lm_rho))
        return fail(ctx, BAD, "missing keyframe landmarks");
    if (nf && (!kf.obs_landmark_id || !kf.obs_cam || !kf.obs_bearing || !kf.obs_velocity || !kf.obs_tref ||
               !kf.obs_sqrt_info))
        return fail(ctx, BAD, "missing keyframe observations");
    if (kf.sb_width != 0 && kf.sb_width != 3 && kf.sb_width != 9) return fail(ctx, BAD, "sb_width must be 0, 3 or 9");
    for (int j = 0; j < nl; ++j) {
        if (kf.lm_anchor_cam[j] < 0 || kf.lm_anchor_cam[j] >= cur.C) return fail(ctx, BAD, "landmark anchor out of range");
        if (kf.lm_id[j] < (j ? kf.lm_id[j - 1] + 1 : cur.uid_next))
            return fail(ctx, BAD, "new landmark ids must increase above every id the window held");
    }
    for (int i = 0; i < nf; ++i)
        if (kf.obs_cam[i] < 0 || kf.obs_cam[i] >= cur.C) return fail(ctx, BAD, "observation camera out of range");
    if (!refresh_state(ctx, s)) return fail(ctx, VIB_ERR_DEVICE, "state download failed");
    const int k = cur.K;
    HostWindow meta = frame_meta(cur);
    meta.K = k + 1;
    meta.frame_id.push_back(kf.frame_id);
    meta.pose.insert(meta.pose.end(), kf.pose, kf.pose + 7);
    meta.sb.insert(meta.sb.end(), kf.speed_bias, kf.speed_bias + 9);
    meta.pose_free.push_back(kf.pose_free ? 1 : 0);
    meta.sb_width.push_back((uint8_t)kf.sb_width);
    if (kf.imu_data) {
        if (k < 1) return fail(ctx, BAD, "an IMU edge needs a previous frame");
        meta.imu_from.push_back(k - 1);
        meta.imu_to.push_back(k);
        meta.imu.insert(meta.imu.end(), kf.imu_data, kf.imu_data + VIB_IMU_STRIDE);
        meta.E += 1;
    }
    if (nl) meta.uid_next = kf.lm_id[nl - 1] + 1;
    // observations in landmark-id order (stable), each with its rank among those of its landmark
    std::vector<int> order(nf);
    for (int i = 0; i < nf; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return kf.obs_landmark_id[a] < kf.obs_landmark_id[b]; });
    const size_t NF = nf, NL = nl;
    std::vector<int64_t> ouid(NF);
    std::vector<int32_t> orank(NF), ofc(NF), lfc(NL);
    std::vector<double> ob((size_t)OB_PLANES * NF), lmd((size_t)LM_PLANES * NL);
    std::vector<uint8_t> lfree(NL);
    for (size_t p = 0; p < NF; ++p) {
        const int i = order[p];
        ouid[p] = kf.obs_landmark_id[i];
        orank[p] = p && ouid[p] == ouid[p - 1] ? orank[p - 1] + 1 : 0;
        ofc[p] = (k << 8) | kf.obs_cam[i];
        for (int q = 0; q < 3; ++q) {
            ob[(OB_BX + q) * NF + p] = kf.obs_bearing[3 * i + q];
            ob[(OB_VX + q) * NF + p] = kf.obs_velocity[3 * i + q];
            ob[(OB_S00 + q) * NF + p] = kf.obs_sqrt_info[3 * i + q];
        }
        ob[OB_T * NF + p] = kf.obs_tref[i];
    }
    for (size_t j = 0; j < NL; ++j) {
        lfc[j] = (k << 8) | kf.lm_anchor_cam[j];
        for (int q = 0; q < 3; ++q) {
            lmd[(LM_BX + q) * NL + j] = kf.lm_bearing[3 * j + q];
            lmd[(LM_VX + q) * NL + j] = kf.lm_velocity[3 * j + q];
        }
        lmd[LM_T * NL + j] = kf.lm_tref[j];
        lfree[j] = kf.lm_free ? (kf.lm_free[j] ? 1 : 0) : 1;
    }
    EditBufs b;
    if (!edit_buffers(ctx, cur.L, cur.F, b, nl, nf)) return fail(ctx, VIB_ERR_OUT_OF_MEMORY, "append buffer allocation failed");
    int miss = 0;
    CopyBatch cb;  // the source descriptor and the keyframe's records in one transfer
    cb.up(b.src, &s.dev, sizeof(WinDev));
    cb.zero(b.sel, 16 * sizeof(int));
    cb.up(b.ins_obs_uid, ouid.data(), 8 * NF);
    cb.up(b.ins_obs_rank, orank.data(), 4 * NF);
    cb.up(b.ins_obs_fc, ofc.data(), 4 * NF);
    cb.up(b.ins_ob, ob.data(), 8 * ob.size());
    cb.up(b.ins_lm_fc, lfc.data(), 4 * NL);
    cb.up(b.ins_lmd, lmd.data(), 8 * lmd.size());
    cb.up(b.ins_rho, kf.lm_rho, 8 * NL);
    cb.up(b.ins_free, lfree.data(), NL);
    cb.up(b.ins_uid, kf.lm_id, 8 * NL);
    cb.zero(b.add_cnt, 4 * ((size_t)cur.L + NL));
    bool ok = cb.flush(ctx->stream, ctx->xfer) &&
              vibk_insert_resolve(b.src, b.ins_obs_uid, nf, b.ins_uid, nl, b.ins_obs_lm, b.add_cnt, b.sel,
                                  ctx->stream) == hipSuccess;
    cb.down(&miss, b.sel, sizeof(int));
    ok = ok && cb.flush(ctx->stream, ctx->xfer);
    if (!ok) return fail(ctx, VIB_ERR_DEVICE, "keyframe upload failed");
    EditSpec sp{};
    for (int f = 0; f < kMaxFrames; ++f) sp.renum[f] = f < k ? f : -1;
    sp.ins_lm = nl;
    sp.ins_obs = nf;
    sp.add_cnt = b.add_cnt;
    sp.ins_lm_fc = b.ins_lm_fc;
    sp.ins_lmd = b.ins_lmd;
    sp.ins_rho = b.ins_rho;
    sp.ins_free = b.ins_free;
    sp.ins_uid = b.ins_uid;
    sp.ins_obs_lm = b.ins_obs_lm;
    sp.ins_obs_rank = b.ins_obs_rank;
    sp.ins_obs_fc = b.ins_obs_fc;
    sp.ins_ob = b.ins_ob;
    const int st = edit_window(ctx, s, sp, std::move(meta), s, b, true);
    if (st == VIB_OK && unmatched) *unmatched = miss;
    return st;
}

int marginalize_second_new(vib_context* ctx, Slot& s, const double* samples, int count, double t0, double t1,
                           const vib_imu_params* params, vib_marginal_report* rep) {
    if (!refresh_state(ctx, s)) return fail(ctx, VIB_ERR_DEVICE, "state download failed");
    const HostWindow& cur = s.host;
    if (cur.K < 3) return fail(ctx, VIB_ERR_INVALID_ARGUMENT, "SECOND_NEW needs at least three frames");
    const int K = cur.K, v = K - 2, left = K - 3, right = K - 1;
    // IMU edges of the slot (SlidingWindowManager.cpp:275-311)
    std::vector<int> incident;
    for (int e = 0; e < cur.E; ++e) if (cur.imu_from[e] == v || cur.imu_to[e] == v) incident.push_back(e);
    bool sb_in_prior = false, prior_touches = false;
    for (size_t b = 0; b < cur.prior_kind.size(); ++b)
        if ((cur.prior_kind[b] == VIB_KEY_POSE || cur.prior_kind[b] == VIB_KEY_SPEED_BIAS) &&
            cur.prior_id[b] == cur.frame_id[v]) {
            prior_touches = true;
            if (cur.prior_kind[b] == VIB_KEY_SPEED_BIAS) sb_in_prior = true;
        }
    int imu_mode = incident.empty() ? VIB_IMU_MERGE_NONE : VIB_IMU_MERGE_BRIDGE;
    std::vector<char> keep_e(cur.E, 1);
    std::vector<double> merged;
    if (incident.size() == 2 && !sb_in_prior && samples && count > 0 && params) {
        int a = incident[0], b = incident[1];
        if (cur.imu_from[a] != left) std::swap(a, b);
        if (cur.imu_from[a] != left || cur.imu_to[a] != v || cur.imu_from[b] != v || cur.imu_to[b] != right)
            return fail(ctx, VIB_ERR_INVALID_ARGUMENT, "IMU merge endpoint/layout mismatch");
        const double* da = &cur.imu[(size_t)VIB_IMU_STRIDE * a];
        const double* db = &cur.imu[(size_t)VIB_IMU_STRIDE * b];
        if (!(t1 > t0) || std::fabs((t1 - t0) - (da[0] + db[0])) > 1e-9)
            return fail(ctx, VIB_ERR_INVALID_ARGUMENT, "IMU merge interval does not span the two edges");
        if (count > 10000) return fail(ctx, VIB_ERR_INVALID_ARGUMENT, "merged IMU interval exceeds sample budget");
        for (int i = 0; i < count; ++i) {
            for (int c = 0; c < 7; ++c)
                if (!std::isfinite(samples[7 * i + c])) return fail(ctx, VIB_ERR_INVALID_ARGUMENT, "non-finite IMU sample");
            if (i && !(samples[7 * i] > samples[7 * (i - 1)]))
                return fail(ctx, VIB_ERR_INVALID_ARGUMENT, "non-increasing IMU samples");
        }
        keep_e[a] = keep_e[b] = 0;
        if (t1 - t0 > 10.0) {
            imu_mode = VIB_IMU_MERGE_SUPPRESSED;  // maxImuIntervalSeconds: kept out of the solve
        } else {
            // the left edge's reference bias is kept while the estimate is within 0.0003 rad/s
            // (gyro) and 0.02 m/s^2 (accel) of it (BodyImuFactor.cpp:146-170)
            double lin[9];
            std::copy_n(&cur.sb[9 * left], 9, lin);
            double dg = 0, dacc = 0;
            for (int i = 0; i < 3; ++i) {
                dg += (lin[3 + i] - da[25 + i]) * (lin[3 + i] - da[25 + i]);
                dacc += (lin[6 + i] - da[28 + i]) * (lin[6 + i] - da[28 + i]);
            }
            const bool incremental = std::sqrt(dg) <= 3e-4 && std::sqrt(dacc) <= 0.02;
            if (incremental) for (int i = 0; i < 6; ++i) lin[3 + i] = da[25 + i];
            double g0[3], g1[3];
            gyro_at(samples, count, t0, g0);
            gyro_at(samples, count, t1, g1);
            // the re-integration runs on the device (k_preintegrate), as for any batch of edges
            merged.assign(VIB_IMU_STRIDE, 0);
            const int32_t b0 = 0, e0 = count;
            int32_t est = VIB_OK;
            const int st = preintegrate_locked(ctx, 1, samples, count, &b0, &e0, &t0, &t1, lin, params, g0, g1,
                                               merged.data(), &est, nullptr);
            if (st == VIB_ERR_DEVICE || st == VIB_ERR_OUT_OF_MEMORY) return st;
            if (st != VIB_OK || est != VIB_OK) return fail(ctx, VIB_ERR_INVALID_ARGUMENT, "IMU re-integration failed");
            imu_mode = incremental ? VIB_IMU_MERGE_INCREMENTAL : VIB_IMU_MERGE_REINTEGRATED;
        }
    }
    // the IMU edges after the merge: the kept ones, then the merged edge
    std::vector<int> w_from, w_to;
    std::vector<double> w_rec;
    for (int e = 0; e < cur.E; ++e) {
        if (!keep_e[e]) continue;
        w_from.push_back(cur.imu_from[e]);
        w_to.push_back(cur.imu_to[e]);
        w_rec.insert(w_rec.end(), cur.imu.begin() + (size_t)VIB_IMU_STRIDE * e, cur.imu.begin() + (size_t)VIB_IMU_STRIDE * (e + 1));
    }
    if (!merged.empty()) {
        w_from.push_back(left);
        w_to.push_back(right);
        w_rec.insert(w_rec.end(), merged.begin(), merged.end());
    }
    // landmarks anchored in the slot (SlidingWindowManager.cpp:312-457) on the device: classes,
    // the anchor search (preferred anchors: newer frames first, then lower camera index; every
    // retained frame counts as a keyframe), keep flags of the next window and of the marginal
    // system; the slot's image measurements of landmarks anchored elsewhere are discarded
    // (:458-470)
    EditBufs b;
    if (!edit_buffers(ctx, cur.L, cur.F, b)) return fail(ctx, VIB_ERR_OUT_OF_MEMORY, "slide buffer allocation failed");
    int cnt[4] = {0, 0, 0, 0};
    int frames = 0;
    bool observed = false;
    bool ok = edit_begin(ctx, s, b);
    if (ok && cur.L) {
        ok = vibk_second_new_classify(b.src, cur.L, v, b.cls, ctx->stream) == hipSuccess &&
             vibk_reanchor(b.src, v, nullptr, b.cls, cur.L, b.chosen, b.chosen_rho, ctx->stream) == hipSuccess &&
             vibk_second_new_flags(b.src, cur.L, v, b.cls, b.chosen, b.chosen_rho, b.keep_lm, b.keep_f, b.over_f,
                                   b.over_rho, b.marg_lm, b.sel, ctx->stream) == hipSuccess &&
             hipMemcpyAsync(cnt, b.sel, sizeof(cnt), hipMemcpyDeviceToHost, ctx->stream) == hipSuccess &&
             touched_frames(ctx, b, cur.L, b.marg_lm, frames, observed);
    }
    if (!ok) return fail(ctx, VIB_ERR_DEVICE, "slide topology launch failed");
    // IMU edges incident to the slot after the merge are eliminated with it
    std::vector<int> m_from, m_to;
    std::vector<double> m_rec;
    for (size_t e = 0; e < w_from.size(); ++e)
        if (w_from[e] == v || w_to[e] == v) {
            m_from.push_back(w_from[e]);
            m_to.push_back(w_to[e]);
            m_rec.insert(m_rec.end(), w_rec.begin() + (size_t)VIB_IMU_STRIDE * e, w_rec.begin() + (size_t)VIB_IMU_STRIDE * (e + 1));
        }
    const bool consumed = !m_from.empty() || observed;
    *rep = vib_marginal_report{};
    rep->imu_mode = imu_mode;
    rep->reset_landmarks = cnt[0];
    rep->demoted_landmarks = cnt[1];
    rep->dropped_factors = cnt[2];
    rep->eliminated_landmarks = cnt[3];
    HostWindow next = frame_meta(cur);
    drop_frame(next, v);
    keep_edges(next, w_from, w_to, w_rec, std::vector<char>(w_from.size(), 1), v);
    EditSpec sp{};
    for (int k = 0; k < kMaxFrames; ++k) sp.renum[k] = k < K ? (k == v ? -1 : (k > v ? k - 1 : k)) : -1;
    sp.keep_lm = b.keep_lm;
    sp.keep_f = b.keep_f;
    sp.over_f = b.over_f;
    sp.over_rho = b.over_rho;
    if (!consumed && !prior_touches) {
        // Marginalizer.cpp:299-310: nothing incident is left, the frozen prior is kept as it is
        const int up = edit_window(ctx, s, sp, std::move(next), s, b, true);
        if (up != VIB_OK) return up;
        const HostWindow& h = s.host;
        rep->status = VIB_OK;
        rep->prior_reused = 1;
        rep->retained_dimension = h.prior_rows ? (int)(h.prior_A.size() / h.prior_rows) : 0;
        rep->prior_rank = h.prior_rows;
        return VIB_OK;
    }
    SchurPrior p;
    int st = schur_prior(ctx, s, b, b.marg_lm, frames, observed, v, m_from, m_to, m_rec, p);
    if (st != VIB_OK) return st;
    set_prior(next, p);
    st = edit_window(ctx, s, sp, std::move(next), s, b, true);
    if (st != VIB_OK) return st;
    rep->status = VIB_OK;
    rep->eliminated_dimension = p.eliminated;
    rep->retained_dimension = p.cols;
    rep->prior_rank = p.rows;
    rep->consumed_factors = p.consumed;
    rep->eigen_cutoff = p.cutoff;
    return VIB_OK;
}

}  // namespace

extern "C" {

int vib_abi_version(void) { return VIB_ABI_VERSION; }

int vib_create(vib_context** out, int device) {
    if (!out) return VIB_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return VIB_ERR_NO_DEVICE;
    if (device < 0 || device >= count) return VIB_ERR_INVALID_ARGUMENT;
    if (hipSetDevice(device) != hipSuccess) return VIB_ERR_DEVICE;
    auto* ctx = new vib_context();
    ctx->device = device;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return VIB_ERR_DEVICE;
    }
    void* sp = nullptr;  // the no-allocation gate's scratch on this device
    if (const int st = scratch_address(device, &sp)) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return st;
    }
    ctx->scratch = sp;
    const char* g = std::getenv("VIB_NO_GRAPH");
    ctx->use_graph = !(g && g[0] == '1');
    if (const char* c = std::getenv("VIB_CHUNK")) HostWindow::chunk_size = std::max(8, std::atoi(c));
    *out = ctx;
    return VIB_OK;
}

int vib_context_device(const vib_context* ctx, int32_t* device) {
    if (!ctx || !device) return VIB_ERR_INVALID_ARGUMENT;
    *device = ctx->device;
    return VIB_OK;
}

int vib_destroy(vib_context* ctx) {
    if (!ctx) return VIB_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->slots.clear();
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return VIB_OK;
}

int vib_device_synchronize(vib_context* ctx) {
    if (!ctx) return VIB_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    return hipDeviceSynchronize() == hipSuccess ? VIB_OK : fail(ctx, VIB_ERR_DEVICE, "device synchronization failed");
}

const char* vib_last_error(const vib_context* ctx) { return ctx ? ctx->error.c_str() : "null context"; }

int64_t vib_window_footprint(const vib_window* w, int32_t precision) {
    if (!w) return -1;
    HostWindow h;
    std::string err;
    if (h.build(*w, err) != VIB_OK) return -1;
    (void)precision;
    return (int64_t)Slot::footprint(h);
}

int vib_window_upload(vib_context* ctx, int32_t slot, const vib_window* w) {
    if (!ctx || !w || slot < 0) return VIB_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mutex);
    (void)hipSetDevice(ctx->device);
    auto h = std::make_unique<HostWindow>();
    std::string err;
    // a failed upload leaves the slot empty (its previous window is released)
    auto drop = [&](int code) {
        ctx->slots.erase(slot);
        return fail(ctx, code, err);
    };
    if (const int st = h->build(*w, err)) return drop(st);
    const size_t need = Slot::footprint(*h);
    if (ctx->hbm_budget > 0) {
        size_t resident = 0;
        for (const auto& kv : ctx->slots)
            if (kv.first != slot && kv.second && kv.second->used) resident += kv.second->bytes;
        if (resident + need > ctx->hbm_budget) {
            err = "window does not fit the context's HBM budget";
            return drop(VIB_ERR_OUT_OF_MEMORY);
        }
    }
    auto& s = ctx->slots[slot];
    if (!s) s = std::make_unique<Slot>();
    if (const int st = s->upload(*h, ctx->stream, ctx->xfer, err)) return drop(st);
    return VIB_OK;
}

int vib_set_hbm_budget(vib_context* ctx, int64_t bytes) {
    if (!ctx || bytes < 0) return VIB_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mutex);
    ctx->hbm_budget = (size_t)bytes;
    return VIB_OK;
}

int vib_resident_bytes(vib_context* ctx, int64_t* bytes) {
    if (!ctx || !bytes) return VIB_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mutex);
    size_t total = 0;
    for (const auto& kv : ctx->slots)
        if (kv.second && kv.second->used) total += kv.second->bytes;
    *bytes = (int64_t)total;
    return VIB_OK;
}

int vib_slot_bytes(vib_context* ctx, int32_t slot, int64_t* bytes) {
    if (!ctx || !bytes) return VIB_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mutex);
    Slot* s = find_slot(ctx, slot);
    if (!s) return fail(ctx, VIB_ERR_NOT_FOUND, "unknown slot");
    *bytes = (int64_t)s->bytes;
    return VIB_OK;
}

int vib_set_host_retained(vib_context* ctx, int32_t slot, const double* H, const double* g) {
    if (!ctx || (!H) != (!g)) return VIB_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mutex);
    (void)hipSetDevice(ctx->device);
    Slot* s = find_slot(ctx, slot);
    if (!s) return fail(ctx, VIB_ERR_NOT_FOUND, "unknown slot");
    if (!H) {
        s->host_set = false;
        return VIB_OK;
    }
    const size_t nr = s->host.nr, need = nr * nr + nr;
    for (size_t i = 0; i < nr * nr; ++i)
        if (!std::isfinite(H[i])) return fail(ctx, VIB_ERR_INVALID_ARGUMENT, "non-finite host normal matrix");
    for (size_t i = 0; i < nr; ++i)
        if (!std::isfinite(g[i])) return fail(ctx, VIB_ERR_INVALID_ARGUMENT, "non-finite host gradient");
    if (need > s->host_cap) {
        if (s->host_terms) (void)hipFree(s->host_terms);
        s->host_terms = nullptr;
        s->host_cap = 0;
        s->host_set = false;
        if (hipMalloc(&s->host_terms, need * sizeof(double)) != hipSuccess) {
            s->host_terms = nullptr;
            (void)hipGetLastError();
            return fail(ctx, VIB_ERR_OUT_OF_MEMORY, "host retained terms allocation failed");
        }
        s->host_cap = need;
    }
    const bool ok = hipMemcpyAsync(s->host_terms, H, nr * nr * sizeof(double), hipMemcpyHostToDevice, ctx->stream) == hipSuccess &&
                    hipMemcpyAsync(s->host_terms + nr * nr, g, nr * sizeof(double), hipMemcpyHostToDevice, ctx->stream) == hipSuccess &&
                    hipStreamSynchronize(ctx->stream) == hipSuccess;
    if (!ok) return fail(ctx, VIB_ERR_DEVICE, "host retained terms upload failed");
    s->host_set = true;
    return VIB_OK;
}

int vib_window_append(vib_context* ctx, int32_t slot, const vib_keyframe* kf, int32_t* unmatched) {
    if (!ctx || !kf) return VIB_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mutex);
    (void)hipSetDevice(ctx->device);
    Slot* s = find_slot(ctx, slot);
    if (!s) return fail(ctx, VIB_ERR_NOT_FOUND, "unknown slot");
    return append_keyframe(ctx, *s, *kf, unmatched);
}

int vib_window_release(vib_context* ctx, int32_t slot) {
    if (!ctx) return VIB_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mutex);
    auto it = ctx->slots.find(slot);
    if (it == ctx->slots.end()) return fail(ctx, VIB_ERR_NOT_FOUND, "unknown slot");
    ctx->slots.erase(it);
    return VIB_OK;
}

int vib_read_state(vib_context* ctx, int32_t slot, vib_window* w) {
    if (!ctx || !w) return VIB_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mutex);
    Slot* s = find_slot(ctx, slot);
    if (!s) return fail(ctx, VIB_ERR_NOT_FOUND, "unknown slot");
    const WinDev& d = s->dev;
    if (w->num_frames != d.K || w->num_landmarks != d.L) return fail(ctx, VIB_ERR_INVALID_ARGUMENT, "window shape mismatch");
    bool ok = hipMemcpyAsync(w->pose, d.pose, 56 * d.K, hipMemcpyDeviceToHost, ctx->stream) == hipSuccess &&
              hipMemcpyAsync(w->speed_bias, d.sb, 72 * d.K, hipMemcpyDeviceToHost, ctx->stream) == hipSuccess &&
              hipMemcpyAsync(w->rig, d.rig, 56, hipMemcpyDeviceToHost, ctx->stream) == hipSuccess &&
              hipMemcpyAsync(w->td, d.td, 8, hipMemcpyDeviceToHost, ctx->stream) == hipSuccess &&
              (d.L == 0 || hipMemcpyAsync(w->lm_rho, d.rho, 8 * d.L, hipMemcpyDeviceToHost, ctx->stream) == hipSuccess) &&
              hipStreamSynchronize(ctx->stream) == hipSuccess;
    return ok ? VIB_OK : fail(ctx, VIB_ERR_DEVICE, "state download failed");
}

int vib_window_shape(vib_context* ctx, int32_t slot, int32_t* frames, int32_t* landmarks, int32_t* observations,
                     int32_t* imu, int32_t* prior_rows, int32_t* prior_cols) {
    if (!ctx) return VIB_ERR_INVALID_ARGUMENT;
    Slot* s = find_slot(ctx, slot);
    if (!s) return fail(ctx, VIB_ERR_NOT_FOUND, "unknown slot");
    if (frames) *frames = s->host.K;
    if (landmarks) *landmarks = s->host.L;
    if (observations) *observations = s->host.F;
    if (imu) *imu = s->host.E;
    if (prior_rows) *prior_rows = s->host.prior_rows;
    if (prior_cols) *prior_cols = s->host.prior_cols;
    return VIB_OK;
}

int vib_retained_dimension(vib_context* ctx, int32_t slot, int32_t* nr, int32_t* nv) {
    if (!ctx) return VIB_ERR_INVALID_ARGUMENT;
    Slot* s = find_slot(ctx, slot);
    if (!s) return fail(ctx, VIB_ERR_NOT_FOUND, "unknown slot");
    if (nr) *nr = s->host.nr;
    if (nv) *nv = s->host.nv;
    return VIB_OK;
}

int vib_linearize(vib_context* ctx, int32_t slot, int32_t precision) {
    if (!ctx) return VIB_ERR_INVALID_ARGUMENT;
    if (precision != VIB_FP64 && precision != VIB_MIXED) return fail(ctx, VIB_ERR_INVALID_ARGUMENT, "unknown precision");
    std::lock_guard<std::mutex> lock(ctx->mutex);
    Slot* s = find_slot(ctx, slot);
    if (!s) return fail(ctx, VIB_ERR_NOT_FOUND, "unknown slot");
    LmState lm;
    if (!read_lm(ctx, *s, lm)) return fail(ctx, VIB_ERR_DEVICE, "LM state read failed");
    lm.new_point = 1; lm.status = 0; lm.iteration = 0;
    std::vector<Slot*> one{s};
    Batch& b = ctx->one;
    b.mixed = precision == VIB_MIXED;
    std::string err;
    if (!write_lm(ctx, *s, lm) || !b.build(one, ctx->stream, ctx->xfer, err) || !b.linearize(ctx->stream) ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        return fail(ctx, VIB_ERR_DEVICE, "linearization launch failed");
    s->lin_precision = precision;
    return VIB_OK;
}

int vib_solve_step(vib_context* ctx, int32_t slot, double lambda, int32_t precision, int32_t refinement,
                   vib_step_report* rep) {
    if (!ctx || !rep || !(lambda > 0) || !std::isfinite(lambda)) return VIB_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mutex);
    *rep = vib_step_report{};
    Slot* s = find_slot(ctx, slot);
    if (!s) return fail(ctx, VIB_ERR_NOT_FOUND, "unknown slot");
    if (precision != s->lin_precision)
        return fail(ctx, VIB_ERR_INVALID_ARGUMENT, "step precision differs from the linearization's (vib_linearize)");
    LmState lm;
    if (!read_lm(ctx, *s, lm)) return fail(ctx, VIB_ERR_DEVICE, "LM state read failed");
    lm.lambda = lambda; lm.status = 0; lm.iteration = 0;
    std::vector<Slot*> one{s};
    Batch& b = ctx->one;
    b.mixed = precision == VIB_MIXED;
    std::string err;
    if (!write_lm(ctx, *s, lm) || !b.build(one, ctx->stream, ctx->xfer, err) || !b.step(ctx->stream, refinement, 1) ||
        !read_lm(ctx, *s, lm))
        return fail(ctx, VIB_ERR_DEVICE, "step launch failed");
    rep->cost = lm.cost;
    rep->predicted_reduction = lm.predicted;
    rep->trial_cost = lm.trial_cost;
    rep->max_increment = lm.max_increment;
    rep->trial_valid = lm.trial_valid;
    rep->status = status_from_lm(lm.status);
    if (rep->status != VIB_OK) return fail(ctx, rep->status, status_message(rep->status));
    // the increment goes back to the caller: a non-finite one is the reference's
    // "nonfinite CUDA Schur step" (BackendOptimizer.cpp:253-254)
    if (lm.step_nonfinite) {
        rep->status = VIB_ERR_NONFINITE_STEP;
        return fail(ctx, VIB_ERR_NONFINITE_STEP, "nonfinite Schur step");
    }
    return VIB_OK;
}

int vib_accept(vib_context* ctx, int32_t slot) {
    if (!ctx) return VIB_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mutex);
    Slot* s = find_slot(ctx, slot);
    if (!s) return fail(ctx, VIB_ERR_NOT_FOUND, "unknown slot");
    std::vector<Slot*> one{s};
    Batch& b = ctx->one;
    b.mixed = 0;
    std::string err;
    if (!b.build(one, ctx->stream, ctx->xfer, err) || !b.launch(10, ctx->stream, 0, 2) || !b.launch(11, ctx->stream, 0, 0) ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        return fail(ctx, VIB_ERR_DEVICE, "accept failed");
    return VIB_OK;
}

int vib_read_step(vib_context* ctx, int32_t slot, double* delta, int32_t capacity) {
    if (!ctx || !delta) return VIB_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mutex);
    Slot* s = find_slot(ctx, slot);
    if (!s) return fail(ctx, VIB_ERR_NOT_FOUND, "unknown slot");
    if (!ensure_bulk(ctx, *s)) return fail(ctx, VIB_ERR_DEVICE, "window download failed");
    const HostWindow& h = s->host;
    if (capacity < h.nr + h.ne) return fail(ctx, VIB_ERR_INVALID_ARGUMENT, "capacity too small");
    std::vector<double> dl(h.L);
    bool ok = hipMemcpyAsync(delta, s->dev.dstate, 8 * h.nr, hipMemcpyDeviceToHost, ctx->stream) == hipSuccess &&
              (h.L == 0 || hipMemcpyAsync(dl.data(), s->dev.lm_delta, 8 * h.L, hipMemcpyDeviceToHost, ctx->stream) == hipSuccess) &&
              hipStreamSynchronize(ctx->stream) == hipSuccess;
    if (!ok) return fail(ctx, VIB_ERR_DEVICE, "step download failed");
    for (int l = 0; l < h.L; ++l) if (h.lm_col[l] >= 0) delta[h.lm_col[l]] = dl[l];
    return VIB_OK;
}

int vib_read_reduced(vib_context* ctx, int32_t slot, double* S, double* rhs, int32_t capacity) {
    if (!ctx) return VIB_ERR_INVALID_ARGUMENT;
    Slot* s = find_slot(ctx, slot);
    if (!s) return fail(ctx, VIB_ERR_NOT_FOUND, "unknown slot");
    const int nr = s->host.nr;
    if (capacity < nr) return fail(ctx, VIB_ERR_INVALID_ARGUMENT, "capacity too small");
    bool ok = (!S || hipMemcpyAsync(S, s->dev.Scopy, 8 * (size_t)nr * nr, hipMemcpyDeviceToHost, ctx->stream) == hipSuccess) &&
              (!rhs || hipMemcpyAsync(rhs, s->dev.rhs, 8 * nr, hipMemcpyDeviceToHost, ctx->stream) == hipSuccess) &&
              hipStreamSynchronize(ctx->stream) == hipSuccess;
    return ok ? VIB_OK : fail(ctx, VIB_ERR_DEVICE, "download failed");
}

int vib_read_normal(vib_context* ctx, int32_t slot, double* H, double* g, double* cost, int32_t capacity) {
    if (!ctx) return VIB_ERR_INVALID_ARGUMENT;
    Slot* s = find_slot(ctx, slot);
    if (!s) return fail(ctx, VIB_ERR_NOT_FOUND, "unknown slot");
    const int nr = s->host.nr;
    if (capacity < nr) return fail(ctx, VIB_ERR_INVALID_ARGUMENT, "capacity too small");
    LmState lm;
    bool ok = (!H || hipMemcpyAsync(H, s->dev.Hfull, 8 * (size_t)nr * nr, hipMemcpyDeviceToHost, ctx->stream) == hipSuccess) &&
              (!g || hipMemcpyAsync(g, s->dev.gfull, 8 * nr, hipMemcpyDeviceToHost, ctx->stream) == hipSuccess) &&
              read_lm(ctx, *s, lm);
    if (cost) *cost = lm.lin_cost;
    return ok ? VIB_OK : fail(ctx, VIB_ERR_DEVICE, "download failed");
}

int vib_solve(vib_context* ctx, int32_t slot, const vib_solve_options* o, vib_solve_report* rep) {
    if (!ctx || !o || !rep) return VIB_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mutex);
    Slot* s = find_slot(ctx, slot);
    if (!s) return fail(ctx, VIB_ERR_NOT_FOUND, "unknown slot");
    std::vector<Slot*> one{s};
    std::vector<vib_solve_report> reps;
    const int st = run_lm(ctx, one, *o, reps);
    if (!reps.empty()) *rep = reps[0];
    else rep->status = st;
    return st;
}

int vib_solve_batch(vib_context* ctx, int32_t count, const int32_t* ids, const vib_solve_options* o,
                    vib_solve_report* reports) {
    if (!ctx || !o || !ids || !reports || count <= 0) return VIB_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mutex);
    std::vector<Slot*> slots;
    for (int i = 0; i < count; ++i) {
        Slot* s = find_slot(ctx, ids[i]);
        if (!s) return fail(ctx, VIB_ERR_NOT_FOUND, "unknown slot");
        slots.push_back(s);
    }
    std::vector<vib_solve_report> reps;
    const int st = run_lm(ctx, slots, *o, reps);
    for (size_t i = 0; i < reps.size(); ++i) reports[i] = reps[i];
    return st;
}

int vib_request_begin(vib_context* ctx, int32_t slot, const vib_solve_options* o) {
    if (!ctx || !o) return VIB_ERR_INVALID_ARGUMENT;
    if (o->precision != VIB_FP64 && o->precision != VIB_MIXED) return fail(ctx, VIB_ERR_INVALID_ARGUMENT, "unknown precision");
    std::lock_guard<std::mutex> lock(ctx->mutex);
    (void)hipSetDevice(ctx->device);
    Slot* s = find_slot(ctx, slot);
    if (!s) return fail(ctx, VIB_ERR_NOT_FOUND, "unknown slot");
    if (!write_lm(ctx, *s, fresh_lm(*o))) return fail(ctx, VIB_ERR_DEVICE, "LM state upload failed");
    return VIB_OK;
}

int vib_request_iterate(vib_context* ctx, int32_t count, const int32_t* ids, int32_t precision, int32_t refinement,
                        int32_t* done, vib_solve_report* reps) {
    if (!ctx || !ids || !done || count <= 0) return VIB_ERR_INVALID_ARGUMENT;
    if (precision != VIB_FP64 && precision != VIB_MIXED) return fail(ctx, VIB_ERR_INVALID_ARGUMENT, "unknown precision");
    std::lock_guard<std::mutex> lock(ctx->mutex);
    (void)hipSetDevice(ctx->device);
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<Slot*> slots;
    for (int i = 0; i < count; ++i) {
        Slot* s = find_slot(ctx, ids[i]);
        if (!s) return fail(ctx, VIB_ERR_NOT_FOUND, "unknown slot");
        for (Slot* o : slots)
            if (o == s) return fail(ctx, VIB_ERR_INVALID_ARGUMENT, "a slot is listed twice");
        if (s->host_set)
            return fail(ctx, VIB_ERR_UNSUPPORTED,
                        "a window with host-evaluated factors is solved through vib_linearize / vib_solve_step");
        slots.push_back(s);
    }
    Batch& b = ctx->requests;
    b.mixed = precision == VIB_MIXED;
    apply_tuning(ctx, b, slots);
    std::string err;
    if (!b.build(slots, ctx->stream, ctx->xfer, err)) return fail(ctx, VIB_ERR_DEVICE, err);
    if (!b.iteration(ctx->stream, refinement))
        return fail(ctx, VIB_ERR_DEVICE, std::string("kernel launch failed: ") + hipGetErrorString(hipGetLastError()));
    std::vector<LmState> st(slots.size());
    CopyBatch cb;
    for (size_t i = 0; i < slots.size(); ++i) cb.down(&st[i], slots[i]->dev.lm, sizeof(LmState));
    if (!cb.flush(ctx->stream, ctx->xfer))
        return fail(ctx, VIB_ERR_DEVICE, std::string("device failure: ") + hipGetErrorString(hipGetLastError()));
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (size_t i = 0; i < slots.size(); ++i) {
        done[i] = st[i].status != 0;
        if (reps) fill_report(st[i], *slots[i], secs, reps[i]);
    }
    return VIB_OK;
}

int vib_copy_probe(vib_context* ctx, int64_t bytes) {
    if (!ctx || bytes <= 0 || bytes % 16) return VIB_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mutex);
    (void)hipSetDevice(ctx->device);
    void *src = nullptr, *dst = nullptr;
    if (hipMalloc(&src, bytes) != hipSuccess || hipMalloc(&dst, bytes) != hipSuccess) {
        if (src) (void)hipFree(src);
        return fail(ctx, VIB_ERR_OUT_OF_MEMORY, "copy probe allocation failed");
    }
    bool ok = hipMemsetAsync(src, 0, bytes, ctx->stream) == hipSuccess &&
              vibk_copy_probe(src, dst, (size_t)bytes, ctx->stream) == hipSuccess &&
              hipStreamSynchronize(ctx->stream) == hipSuccess;
    (void)hipFree(src);
    (void)hipFree(dst);
    return ok ? VIB_OK : fail(ctx, VIB_ERR_DEVICE, "copy probe failed");
}

}  // extern "C"

namespace {

// Device preintegration of `count` intervals of one sample array (one k_preintegrate launch),
// with the context's lock held. Inputs are staged in the context's grow-only device buffer.
int preintegrate_locked(vib_context* ctx, int32_t count, const double* samples, int32_t num_samples,
                        const int32_t* sample_begin, const int32_t* sample_end, const double* t0, const double* t1,
                        const double* speed_bias, const vib_imu_params* params, const double* gyro_start,
                        const double* gyro_end, double* out, int32_t* status, double* kernel_ms) {
    const int S = VIB_IMU_SAMPLE_STRIDE;
    // the checks of the single-interval routine, per interval; invalid intervals are not launched
    std::vector<int> live;
    live.reserve(count);
    for (int e = 0; e < count; ++e) {
        const int b = sample_begin[e], n = sample_end[e];
        const bool ok = b >= 0 && n <= num_samples && n - b >= 2 && t1[e] > t0[e] && samples[(size_t)b * S] <= t0[e] &&
                        samples[(size_t)(n - 1) * S] >= t1[e];
        status[e] = ok ? VIB_OK : VIB_ERR_INVALID_ARGUMENT;
        std::memset(out + (size_t)e * VIB_IMU_STRIDE, 0, sizeof(double) * VIB_IMU_STRIDE);
        if (ok) live.push_back(e);
    }
    const int m = (int)live.size();
    if (m == 0) return VIB_ERR_INVALID_ARGUMENT;
    std::vector<int> hb(m), he(m);
    std::vector<double> ht0(m), ht1(m), hsb((size_t)m * 9), hg0((size_t)m * 3), hg1((size_t)m * 3);
    for (int i = 0; i < m; ++i) {
        const int e = live[i];
        hb[i] = sample_begin[e]; he[i] = sample_end[e]; ht0[i] = t0[e]; ht1[i] = t1[e];
        std::memcpy(&hsb[(size_t)i * 9], speed_bias + (size_t)e * 9, 9 * sizeof(double));
        std::memcpy(&hg0[(size_t)i * 3], gyro_start + (size_t)e * 3, 3 * sizeof(double));
        std::memcpy(&hg1[(size_t)i * 3], gyro_end + (size_t)e * 3, 3 * sizeof(double));
    }
    // one buffer: samples | t0 | t1 | speed/bias | gyro | records | begin | end | status
    const size_t nd = (size_t)num_samples * S + 2 * (size_t)m + 15 * (size_t)m + (size_t)m * VIB_IMU_STRIDE;
    const size_t bytes = nd * sizeof(double) + 3 * (size_t)m * sizeof(int);
    if (bytes > ctx->preint_cap) {
        if (ctx->preint_buf) (void)hipFree(ctx->preint_buf);
        ctx->preint_buf = nullptr;
        ctx->preint_cap = 0;
        const size_t cap = std::max(bytes, (size_t)1 << 20);
        if (hipMalloc(&ctx->preint_buf, cap) != hipSuccess) {
            ctx->preint_buf = nullptr;
            return fail(ctx, VIB_ERR_OUT_OF_MEMORY, "preintegration allocation failed");
        }
        ctx->preint_cap = cap;
    }
    double* d_smp = reinterpret_cast<double*>(ctx->preint_buf);
    double* d_t0 = d_smp + (size_t)num_samples * S;
    double* d_t1 = d_t0 + m;
    double* d_sb = d_t1 + m;
    double* d_g0 = d_sb + (size_t)m * 9;
    double* d_g1 = d_g0 + (size_t)m * 3;
    double* d_out = d_g1 + (size_t)m * 3;
    int* d_b = reinterpret_cast<int*>(d_out + (size_t)m * VIB_IMU_STRIDE);
    int* d_e = d_b + m;
    int* d_st = d_e + m;
    hipStream_t st = ctx->stream;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool ok = hipMemcpyAsync(d_smp, samples, (size_t)num_samples * S * sizeof(double), hipMemcpyHostToDevice, st) == hipSuccess &&
              hipMemcpyAsync(d_t0, ht0.data(), m * sizeof(double), hipMemcpyHostToDevice, st) == hipSuccess &&
              hipMemcpyAsync(d_t1, ht1.data(), m * sizeof(double), hipMemcpyHostToDevice, st) == hipSuccess &&
              hipMemcpyAsync(d_sb, hsb.data(), hsb.size() * sizeof(double), hipMemcpyHostToDevice, st) == hipSuccess &&
              hipMemcpyAsync(d_g0, hg0.data(), hg0.size() * sizeof(double), hipMemcpyHostToDevice, st) == hipSuccess &&
              hipMemcpyAsync(d_g1, hg1.data(), hg1.size() * sizeof(double), hipMemcpyHostToDevice, st) == hipSuccess &&
              hipMemcpyAsync(d_b, hb.data(), m * sizeof(int), hipMemcpyHostToDevice, st) == hipSuccess &&
              hipMemcpyAsync(d_e, he.data(), m * sizeof(int), hipMemcpyHostToDevice, st) == hipSuccess;
    if (ok && kernel_ms) ok = hipEventCreate(&ev0) == hipSuccess && hipEventCreate(&ev1) == hipSuccess &&
                             hipEventRecord(ev0, st) == hipSuccess;
    ok = ok && vibk_preintegrate(m, d_smp, d_b, d_e, d_t0, d_t1, d_sb, params, d_g0, d_g1, d_out, d_st, st) == hipSuccess;
    if (ok && kernel_ms) ok = hipEventRecord(ev1, st) == hipSuccess;
    std::vector<double> hout((size_t)m * VIB_IMU_STRIDE);
    std::vector<int> hst(m);
    ok = ok && hipMemcpyAsync(hout.data(), d_out, hout.size() * sizeof(double), hipMemcpyDeviceToHost, st) == hipSuccess &&
         hipMemcpyAsync(hst.data(), d_st, m * sizeof(int), hipMemcpyDeviceToHost, st) == hipSuccess &&
         hipStreamSynchronize(st) == hipSuccess;
    if (ok && kernel_ms) {
        float ms = 0;
        ok = hipEventElapsedTime(&ms, ev0, ev1) == hipSuccess;
        *kernel_ms = ms;
    }
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (!ok) return fail(ctx, VIB_ERR_DEVICE, "preintegration launch failed");
    int rc = m == count ? VIB_OK : VIB_ERR_INVALID_ARGUMENT;
    for (int i = 0; i < m; ++i) {
        const int e = live[i];
        if (hst[i] != 0) {  // no sample step inside [t0, t1]
            status[e] = VIB_ERR_INVALID_ARGUMENT;
            rc = VIB_ERR_INVALID_ARGUMENT;
            continue;
        }
        std::memcpy(out + (size_t)e * VIB_IMU_STRIDE, &hout[(size_t)i * VIB_IMU_STRIDE], sizeof(double) * VIB_IMU_STRIDE);
    }
    return rc;
}

}  // namespace

extern "C" {

int vib_preintegrate_batch(vib_context* ctx, int32_t count, const double* samples, int32_t num_samples,
                           const int32_t* sample_begin, const int32_t* sample_end, const double* t0, const double* t1,
                           const double* speed_bias, const vib_imu_params* params, const double* gyro_start,
                           const double* gyro_end, double* out, int32_t* status, double* kernel_ms) {
    if (!ctx || count <= 0 || !samples || num_samples < 2 || !sample_begin || !sample_end || !t0 || !t1 || !speed_bias ||
        !params || !gyro_start || !gyro_end || !out || !status)
        return VIB_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mutex);
    (void)hipSetDevice(ctx->device);
    return preintegrate_locked(ctx, count, samples, num_samples, sample_begin, sample_end, t0, t1, speed_bias, params,
                               gyro_start, gyro_end, out, status, kernel_ms);
}

namespace {

// The statistical gate of one resident window (the lock is held by the caller).
// Device buffers of one gate call: the window descriptor and its landmark chunks, the per-factor
// energies and decisions, and the 18 statistics.
struct GateBuffers {
    const WinDev* wins;
    const Chunk* chunks;
    int nchunks;
    double* energy;
    uint8_t* decision;
    double* stats;
};
constexpr int kGateStats = 18;

static int nis_gate_core(vib_context* ctx, Slot* s, double threshold, int apply, vib_gate_report* rep, uint8_t* dec_out,
                         const GateBuffers& g) {
    HostWindow& h = s->host;
    const int F = h.F;
    std::vector<uint8_t> dec(F);
    double st[kGateStats];
    bool ok = vibk_nis_gate(g.wins, g.chunks, g.nchunks, threshold, g.energy, g.decision, F, h.huber, g.stats,
                            ctx->stream) == hipSuccess &&
              hipMemcpyAsync(dec.data(), g.decision, F, hipMemcpyDeviceToHost, ctx->stream) == hipSuccess &&
              hipMemcpyAsync(st, g.stats, sizeof(st), hipMemcpyDeviceToHost, ctx->stream) == hipSuccess &&
              hipStreamSynchronize(ctx->stream) == hipSuccess;
    int over = 0, invalid = 0, skipped = 0;
    for (uint8_t d : dec) {
        over += d == 1;
        invalid += d == 2;
        skipped += d == 3;
    }
    const bool reject = ok && apply && invalid == 0 && over > 0;
    if (reject)
        ok = vibk_nis_apply(g.wins, g.decision, F, ctx->stream) == hipSuccess && hipStreamSynchronize(ctx->stream) == hipSuccess;
    if (!ok) return fail(ctx, VIB_ERR_DEVICE, "gate launch failed");
    auto fill = [](vib_residual_stats& o, const double* v) {
        o.factors = v[0]; o.components = v[1]; o.squared_norm = v[2]; o.over_gate = v[3]; o.max_nis = v[4];
        o.worst_factor = (int32_t)v[5]; o.robust_cost = v[6]; o.median_nis = v[7]; o.p95_nis = v[8];
    };
    fill(rep->evaluated, st);
    fill(rep->admitted, st + 9);
    rep->skipped = skipped;
    rep->invalid = invalid;
    if (reject && h.bulk) {  // a downloaded host mirror follows the device
        const size_t Fs = F;
        for (int f = 0; f < F; ++f) {
            if (dec[f] != 1) continue;
            const size_t p = h.pf_pos[f];
            for (int q = OB_S00; q <= OB_S11; ++q) {
                h.ob[q * Fs + f] = 0.0;
                h.pf_ob[q * Fs + p] = 0.0;
            }
        }
    }
    if (reject) rep->rejected = over;
    if (dec_out) std::memcpy(dec_out, dec.data(), F);
    rep->status = invalid ? VIB_ERR_INVALID_RESIDUAL : VIB_OK;
    return invalid ? fail(ctx, VIB_ERR_INVALID_RESIDUAL, "measurement evaluation failed/nonfinite before robust loss")
                   : VIB_OK;
}

static int nis_gate_locked(vib_context* ctx, Slot* s, double threshold, int apply, vib_gate_report* rep, uint8_t* dec_out) {
    *rep = vib_gate_report{};
    rep->evaluated.worst_factor = rep->admitted.worst_factor = -1;
    const int F = s->host.F;
    if (F == 0) return VIB_OK;
    (void)hipSetDevice(ctx->device);
    std::vector<Slot*> one{s};
    Batch& b = ctx->one;
    b.mixed = 0;
    std::string err;
    if (!b.build(one, ctx->stream, ctx->xfer, err)) return fail(ctx, VIB_ERR_DEVICE, err);
    char* buf = static_cast<char*>(grow_device(ctx->aux_buf, ctx->aux_cap, (size_t)F * 9 + kGateStats * sizeof(double)));
    if (!buf) return fail(ctx, VIB_ERR_OUT_OF_MEMORY, "gate allocation failed");
    double* d_energy = reinterpret_cast<double*>(buf);
    double* d_stats = d_energy + F;
    uint8_t* d_dec = reinterpret_cast<uint8_t*>(d_stats + kGateStats);
    const int st = nis_gate_core(ctx, s, threshold, apply, rep, dec_out,
                                 GateBuffers{b.d_wins, b.d_chunks, b.nchunks, d_energy, d_dec, d_stats});
    return st;
}

}  // namespace

int vib_nis_gate(vib_context* ctx, int32_t slot, double threshold, int32_t apply, vib_gate_report* report,
                 uint8_t* decisions) {
    if (!ctx || !report || !(threshold > 0)) return VIB_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mutex);
    Slot* s = find_slot(ctx, slot);
    if (!s) return fail(ctx, VIB_ERR_NOT_FOUND, "unknown slot");
    return nis_gate_locked(ctx, s, threshold, apply, report, decisions);
}

int vib_nis_gate_static(vib_context* ctx, int32_t slot, double threshold, int32_t apply, vib_gate_report* report,
                        uint8_t* decisions) {
    if (!ctx || !report || !(threshold > 0)) return VIB_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mutex);
    Slot* s = find_slot(ctx, slot);
    if (!s) return fail(ctx, VIB_ERR_NOT_FOUND, "unknown slot");
    *report = vib_gate_report{};
    report->evaluated.worst_factor = report->admitted.worst_factor = -1;
    const HostWindow& h = s->host;
    if (h.F == 0) return VIB_OK;
    if ((size_t)h.F > kStaticGateFactors || h.chunks.size() > kStaticGateChunks)
        return fail(ctx, VIB_ERR_UNSUPPORTED_DIMENSIONS,
                    "window beyond the no-allocation gate's scratch (2^20 factors, 2^17 landmark chunks)");
    (void)hipSetDevice(ctx->device);
    char* base = reinterpret_cast<char*>(ctx->scratch);  // device address of `scratch`, never the name
    WinDev wd = s->dev;
    wd.nchunks = (int)h.chunks.size();
    wd.hostH = wd.hostG = nullptr;
    const int nch = (int)h.chunks.size();
    if (hipMemcpyAsync(base, &wd, sizeof(WinDev), hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
        (nch && hipMemcpyAsync(base + kScratchChunkOffset, h.chunks.data(), nch * sizeof(Chunk), hipMemcpyHostToDevice,
                               ctx->stream) != hipSuccess))
        return fail(ctx, VIB_ERR_DEVICE, "gate descriptor copy failed");
    const GateBuffers g{reinterpret_cast<const WinDev*>(base), reinterpret_cast<const Chunk*>(base + kScratchChunkOffset),
                        nch, reinterpret_cast<double*>(base + kScratchEnergyOffset),
                        reinterpret_cast<uint8_t*>(base + kScratchDecisionOffset),
                        reinterpret_cast<double*>(base + kScratchStatsOffset)};
    return nis_gate_core(ctx, s, threshold, apply, report, decisions, g);
}

int vib_solve_refit(vib_context* ctx, int32_t slot, const vib_solve_options* o, const vib_refit_options* r,
                    vib_refit_report* rep) {
    if (!ctx || !o || !r || !rep || !(r->nis_threshold > 0) || r->max_outlier_refits < 0) return VIB_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mutex);
    *rep = vib_refit_report{};
    rep->pre.evaluated.worst_factor = rep->pre.admitted.worst_factor = -1;
    rep->post.evaluated.worst_factor = rep->post.admitted.worst_factor = -1;
    Slot* s = find_slot(ctx, slot);
    if (!s) return fail(ctx, VIB_ERR_NOT_FOUND, "unknown slot");
    auto failed = [&](int st) {
        rep->solve.status = st;
        rep->termination = VIB_TERM_FAILURE;
        return st;
    };
    if (!r->skip_pre_nis) {  // BackendOptimizer.cpp:681-690
        const int st = nis_gate_locked(ctx, s, r->nis_threshold, 1, &rep->pre, nullptr);
        if (st != VIB_OK) return failed(st);
        rep->pre_rejected = rep->pre.rejected;
    }
    // passes share the iteration allowance (BackendOptimizer.cpp:720-737, 1036-1079)
    int remaining = std::max(1, o->max_iterations), used_total = 0;
    double seconds = 0, initial = 0;
    for (int pass = 0; pass <= r->max_outlier_refits; ++pass) {
        const int minimum_remaining = std::max(0, o->min_iterations - used_total);
        if (remaining <= 0) {
            rep->termination = VIB_TERM_ITERATION_BUDGET;
            break;
        }
        const int reserve = std::min(r->max_outlier_refits - pass, std::max(0, remaining - std::max(1, minimum_remaining)));
        vib_solve_options po = *o;
        po.max_iterations = remaining - std::max(0, reserve);
        std::vector<Slot*> one{s};
        std::vector<vib_solve_report> reps;
        const int st = run_lm(ctx, one, po, reps);
        if (st != VIB_OK || reps.empty()) return failed(st != VIB_OK ? st : VIB_ERR_DEVICE);
        const vib_solve_report& sr = reps[0];
        if (sr.status != VIB_OK) {
            rep->solve = sr;
            return failed(sr.status);
        }
        ++rep->passes;
        if (pass == 0) initial = sr.initial_cost;
        seconds += sr.seconds;
        used_total += sr.iterations;
        remaining -= std::max(1, sr.iterations);  // a pass is charged at least one iteration
        rep->solve = sr;
        const int gst = nis_gate_locked(ctx, s, r->nis_threshold, 1, &rep->post, nullptr);
        if (gst != VIB_OK) return failed(gst);
        const int rejected = rep->post.rejected;
        rep->post_rejected += rejected;
        rep->unrefitted_outliers = rejected;
        const bool iterations_ended = remaining <= 0, refits_ended = pass == r->max_outlier_refits;
        if (!rejected || iterations_ended || refits_ended) {
            rep->termination = iterations_ended ? VIB_TERM_ITERATION_BUDGET
                               : rejected && refits_ended ? VIB_TERM_REFIT_BUDGET
                                                          : VIB_TERM_SOLVER_FINISHED;
            break;
        }
    }
    rep->solve.iterations = used_total;
    rep->solve.initial_cost = initial;
    rep->solve.seconds = seconds;
    return VIB_OK;
}

int vib_whiten_observations(vib_context* ctx, int32_t count, const double* pixels, const int32_t* camera,
                            const double* intrinsics, int32_t num_cameras, const double* sigma, double sigma_all,
                            double* bearings, double* sqrt_info, int32_t* status, double* kernel_ms) {
    if (!ctx || count < 0 || !intrinsics || num_cameras <= 0 || (count > 0 && (!pixels || !bearings || !sqrt_info || !status)))
        return VIB_ERR_INVALID_ARGUMENT;
    if (!camera && num_cameras != 1) return fail(ctx, VIB_ERR_INVALID_ARGUMENT, "camera indices required for several cameras");
    if (count == 0) return VIB_OK;
    std::lock_guard<std::mutex> lock(ctx->mutex);
    (void)hipSetDevice(ctx->device);
    const size_t n = count;
    // one allocation: pixels | intrinsics | sigma | bearings | sqrt_info | camera | status
    const size_t nd = 2 * n + 6 * (size_t)num_cameras + (sigma ? n : 0) + 6 * n;
    const size_t bytes = nd * sizeof(double) + 2 * n * sizeof(int32_t);
    char* buf = static_cast<char*>(grow_device(ctx->aux_buf, ctx->aux_cap, bytes));
    if (!buf) return fail(ctx, VIB_ERR_OUT_OF_MEMORY, "whitening allocation failed");
    double* d_px = reinterpret_cast<double*>(buf);
    double* d_in = d_px + 2 * n;
    double* d_sg = d_in + 6 * (size_t)num_cameras;
    double* d_b = d_sg + (sigma ? n : 0);
    double* d_s = d_b + 3 * n;
    int32_t* d_c = reinterpret_cast<int32_t*>(d_s + 3 * n);
    int32_t* d_st = d_c + n;
    hipStream_t st = ctx->stream;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool ok = hipMemcpyAsync(d_px, pixels, 2 * n * sizeof(double), hipMemcpyHostToDevice, st) == hipSuccess &&
              hipMemcpyAsync(d_in, intrinsics, 6 * (size_t)num_cameras * sizeof(double), hipMemcpyHostToDevice, st) == hipSuccess &&
              (!sigma || hipMemcpyAsync(d_sg, sigma, n * sizeof(double), hipMemcpyHostToDevice, st) == hipSuccess) &&
              (!camera || hipMemcpyAsync(d_c, camera, n * sizeof(int32_t), hipMemcpyHostToDevice, st) == hipSuccess);
    if (ok && kernel_ms) ok = hipEventCreate(&ev0) == hipSuccess && hipEventCreate(&ev1) == hipSuccess &&
                             hipEventRecord(ev0, st) == hipSuccess;
    ok = ok && vibk_whiten(count, d_px, camera ? d_c : nullptr, d_in, num_cameras, sigma ? d_sg : nullptr, sigma_all, d_b,
                           d_s, d_st, st) == hipSuccess;
    if (ok && kernel_ms) ok = hipEventRecord(ev1, st) == hipSuccess;
    ok = ok && hipMemcpyAsync(bearings, d_b, 3 * n * sizeof(double), hipMemcpyDeviceToHost, st) == hipSuccess &&
         hipMemcpyAsync(sqrt_info, d_s, 3 * n * sizeof(double), hipMemcpyDeviceToHost, st) == hipSuccess &&
         hipMemcpyAsync(status, d_st, n * sizeof(int32_t), hipMemcpyDeviceToHost, st) == hipSuccess &&
         hipStreamSynchronize(st) == hipSuccess;
    if (ok && kernel_ms) {
        float ms = 0;
        ok = hipEventElapsedTime(&ms, ev0, ev1) == hipSuccess;
        *kernel_ms = ms;
    }
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (!ok) return fail(ctx, VIB_ERR_DEVICE, "whitening launch failed");
    for (size_t i = 0; i < n; ++i)
        if (status[i] != 0) return VIB_ERR_INVALID_ARGUMENT;
    return VIB_OK;
}

int vib_kernel_count(void) { return kKernelCount; }

const char* vib_kernel_name(int32_t k) { return vibk_name(k); }

int vib_bench(vib_context* ctx, int32_t count, const int32_t* ids, const vib_solve_options* o, int32_t iterations,
              double* iteration_ms, double* kernel_ms) {
    if (!ctx || !o || !ids || count <= 0 || iterations <= 0) return VIB_ERR_INVALID_ARGUMENT;
    if (o->precision != VIB_FP64 && o->precision != VIB_MIXED) return fail(ctx, VIB_ERR_INVALID_ARGUMENT, "unknown precision");
    std::lock_guard<std::mutex> lock(ctx->mutex);
    (void)hipSetDevice(ctx->device);
    std::vector<Slot*> slots;
    for (int i = 0; i < count; ++i) {
        Slot* s = find_slot(ctx, ids[i]);
        if (!s) return fail(ctx, VIB_ERR_NOT_FOUND, "unknown slot");
        slots.push_back(s);
    }
    for (Slot* s : slots) {
        LmState lm{};
        lm.lambda = o->initial_lambda > 0 ? o->initial_lambda : 1e-4;
        lm.increase = 2; lm.new_point = 1; lm.status = 0; lm.max_iterations = 1 << 30; lm.continuous = 1;
        if (!write_lm(ctx, *s, lm)) return fail(ctx, VIB_ERR_DEVICE, "LM state upload failed");
    }
    Batch& batch = ctx->solve;
    batch.mixed = o->precision == VIB_MIXED;
    apply_tuning(ctx, batch, slots);
    std::string err;
    if (!batch.build(slots, ctx->stream, ctx->xfer, err)) return fail(ctx, VIB_ERR_DEVICE, err);
    if (!batch.capture(ctx->stream, o->refinement_steps)) return fail(ctx, VIB_ERR_DEVICE, "graph capture failed");
    std::vector<hipEvent_t> ev(iterations + 1);
    for (auto& e : ev)
        if (hipEventCreate(&e) != hipSuccess) return fail(ctx, VIB_ERR_DEVICE, "event creation failed");
    bool ok = hipEventRecord(ev[0], ctx->stream) == hipSuccess;
    for (int it = 0; ok && it < iterations; ++it)
        ok = hipGraphLaunch(batch.graph_exec, ctx->stream) == hipSuccess && hipEventRecord(ev[it + 1], ctx->stream) == hipSuccess;
    ok = ok && hipStreamSynchronize(ctx->stream) == hipSuccess;
    for (int it = 0; ok && it < iterations; ++it) {
        float ms = 0;
        ok = hipEventElapsedTime(&ms, ev[it], ev[it + 1]) == hipSuccess;
        if (iteration_ms) iteration_ms[it] = ms;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    if (!ok) return fail(ctx, VIB_ERR_DEVICE, std::string("timed iterations failed: ") + hipGetErrorString(hipGetLastError()));
    if (kernel_ms) {
        // the iteration's launches one by one, each between two events; a kernel launched more
        // than once per iteration (the refinement passes) reports the sum of its launches
        const auto seq = batch.sequence(true, o->refinement_steps, 0);
        const size_t per = seq.size();
        std::vector<hipEvent_t> kev(2 * per * iterations);
        for (auto& e : kev)
            if (hipEventCreate(&e) != hipSuccess) return fail(ctx, VIB_ERR_DEVICE, "event creation failed");
        for (int it = 0; ok && it < iterations; ++it)
            for (size_t q = 0; ok && q < per; ++q) {
                hipEvent_t* e = &kev[2 * (it * per + q)];
                ok = hipEventRecord(e[0], ctx->stream) == hipSuccess &&
                     batch.launch(seq[q].first, ctx->stream, o->refinement_steps, seq[q].second) &&
                     hipEventRecord(e[1], ctx->stream) == hipSuccess;
            }
        ok = ok && hipStreamSynchronize(ctx->stream) == hipSuccess;
        for (int k = 0; k < kKernelCount; ++k) kernel_ms[k] = 0;
        for (int it = 0; ok && it < iterations; ++it)
            for (size_t q = 0; ok && q < per; ++q) {
                float ms = 0;
                const hipEvent_t* e = &kev[2 * (it * per + q)];
                ok = hipEventElapsedTime(&ms, e[0], e[1]) == hipSuccess;
                kernel_ms[seq[q].first] += ms / iterations;
            }
        for (auto& e : kev) (void)hipEventDestroy(e);
        if (!ok) return fail(ctx, VIB_ERR_DEVICE, "kernel profiling failed");
    }
    return VIB_OK;
}

int vib_autotune(vib_context* ctx, int32_t count, const int32_t* ids, const vib_solve_options* o, int32_t iterations,
                 vib_tune_report* rep) {
    if (!ctx || !o || !ids || !rep || count <= 0 || iterations <= 0) return VIB_ERR_INVALID_ARGUMENT;
    if (o->precision != VIB_FP64 && o->precision != VIB_MIXED) return fail(ctx, VIB_ERR_INVALID_ARGUMENT, "unknown precision");
    std::lock_guard<std::mutex> lock(ctx->mutex);
    (void)hipSetDevice(ctx->device);
    *rep = vib_tune_report{};
    std::vector<Slot*> slots;
    for (int i = 0; i < count; ++i) {
        Slot* s = find_slot(ctx, ids[i]);
        if (!s) return fail(ctx, VIB_ERR_NOT_FOUND, "unknown slot");
        if (s->host_set) return fail(ctx, VIB_ERR_UNSUPPORTED, "a window with host-evaluated factors is not tuned");
        slots.push_back(s);
    }
    // the windows' states before the timed iterations, restored afterwards
    size_t snap = 0;
    for (const Slot* s : slots) snap += 16 * (size_t)s->host.K + 8 + s->host.L;
    double* keep = static_cast<double*>(grow_device(ctx->tune_buf, ctx->tune_cap, 8 * snap));
    if (!keep) return fail(ctx, VIB_ERR_OUT_OF_MEMORY, "tuning buffer allocation failed");
    auto copy_state = [&](bool save) {
        double* p = keep;
        bool ok = true;
        for (const Slot* s : slots) {
            const WinDev& d = s->dev;
            const size_t part[5] = {7 * (size_t)d.K, 9 * (size_t)d.K, 7, 1, (size_t)d.L};
            double* const arr[5] = {d.pose, d.sb, d.rig, d.td, d.rho};
            for (int q = 0; q < 5 && ok; ++q) {
                if (part[q])
                    ok = hipMemcpyAsync(save ? p : arr[q], save ? arr[q] : p, 8 * part[q], hipMemcpyDeviceToDevice,
                                        ctx->stream) == hipSuccess;
                p += part[q];
            }
        }
        return ok;
    };
    if (!copy_state(true)) return fail(ctx, VIB_ERR_DEVICE, "state snapshot failed");
    const int merges[4] = {1, 2, 4, 8};
    Batch& batch = ctx->solve;
    batch.mixed = o->precision == VIB_MIXED;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        return fail(ctx, VIB_ERR_DEVICE, "event creation failed");
    std::string err;
    bool ok = true;
    double best = 1e300;
    for (int a = 0; a < 4 && ok; ++a)
        for (int b = 0; b < 4 && ok; ++b) {
            for (Slot* s : slots) {
                LmState lm{};
                lm.lambda = o->initial_lambda > 0 ? o->initial_lambda : 1e-4;
                lm.increase = 2; lm.new_point = 1; lm.status = 0; lm.max_iterations = 1 << 30; lm.continuous = 1;
                ok = ok && write_lm(ctx, *s, lm);
            }
            batch.chunk_merge = merges[a];
            batch.lin_merge = merges[b];
            ok = ok && batch.build(slots, ctx->stream, ctx->xfer, err) && batch.capture(ctx->stream, o->refinement_steps) &&
                 hipGraphLaunch(batch.graph_exec, ctx->stream) == hipSuccess &&  // warm-up
                 hipEventRecord(e0, ctx->stream) == hipSuccess;
            for (int it = 0; ok && it < iterations; ++it) ok = hipGraphLaunch(batch.graph_exec, ctx->stream) == hipSuccess;
            float ms = 0;
            ok = ok && hipEventRecord(e1, ctx->stream) == hipSuccess && hipStreamSynchronize(ctx->stream) == hipSuccess &&
                 hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
            const int k = rep->candidates++;
            rep->chunk_merge[k] = merges[a];
            rep->range_merge[k] = merges[b];
            rep->ms_per_iteration[k] = ms / iterations;
            if (ok && ms / iterations < best) {
                best = ms / iterations;
                rep->best = k;
            }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ok = ok && copy_state(false) && hipStreamSynchronize(ctx->stream) == hipSuccess;
    if (!ok) return fail(ctx, VIB_ERR_DEVICE, std::string("tuning launches failed: ") + err);
    rep->landmarks_per_chunk = HostWindow::chunk_size;
    rep->factors_per_range = kLinFactors;
    rep->size_class = tune_class(slots);
    ctx->tuned[{batch.mixed, rep->size_class}] = {rep->chunk_merge[rep->best], rep->range_merge[rep->best]};
    return VIB_OK;
}

int vib_solve_latest_two(vib_context* ctx, int32_t slot, const vib_solve_options* o, int32_t gate,
                         vib_latest_two_report* rep) {
    if (!ctx || !o || !rep) return VIB_ERR_INVALID_ARGUMENT;
    if (o->precision != VIB_FP64 && o->precision != VIB_MIXED) return fail(ctx, VIB_ERR_INVALID_ARGUMENT, "unknown precision");
    std::lock_guard<std::mutex> lock(ctx->mutex);
    (void)hipSetDevice(ctx->device);
    *rep = vib_latest_two_report{};
    Slot* s = find_slot(ctx, slot);
    if (!s) return fail(ctx, VIB_ERR_NOT_FOUND, "unknown slot");
    // Local graph of the two-frame solve (BackendOptimizer.cpp:476-538): factors touching the pose
    // or speed/bias of slots K-2, K-1; those blocks free with the velocity-only speed/bias
    // tangent, everything else constant. A device edit of the resident window into the
    // context's two-frame slot.
    if (!refresh_state(ctx, *s)) return fail(ctx, VIB_ERR_DEVICE, "state download failed");
    const HostWindow& cur = s->host;
    if (cur.K < 2) return fail(ctx, VIB_ERR_LIGHTWEIGHT_SUPPORT, "missing pose or velocity block in two-frame solve");
    const int prev = cur.K - 2, last = cur.K - 1;
    auto is_free = [&](int k) { return k == prev || k == last; };
    EditBufs eb;
    if (!edit_buffers(ctx, cur.L, cur.F, eb)) return fail(ctx, VIB_ERR_OUT_OF_MEMORY, "two-frame buffer allocation failed");
    int support = 0;
    if (!edit_begin(ctx, *s, eb) ||
        vibk_two_frame_select(eb.src, cur.L, prev, last, eb.keep_lm, eb.keep_f, eb.sel, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(&support, eb.sel, sizeof(int), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        return fail(ctx, VIB_ERR_DEVICE, "two-frame selection failed");
    rep->support_landmarks = support;
    if (support < 6) return fail(ctx, VIB_ERR_LIGHTWEIGHT_SUPPORT, "fewer than six current landmarks in two-frame solve");
    HostWindow w = frame_meta(cur);
    w.flags = cur.flags & ~(VIB_RIG_FREE | VIB_TD_FREE);
    w.pose_free.assign(cur.K, 0);
    w.sb_width.assign(cur.K, 0);
    w.pose_free[prev] = w.pose_free[last] = 1;
    w.sb_width[prev] = w.sb_width[last] = 3;
    // the IMU bridge: exactly the edge ending at K-2 and the edge K-2 -> K-1
    int edges[2] = {-1, -1};
    w.imu_from.clear(); w.imu_to.clear(); w.imu.clear();
    for (int e = 0; e < cur.E; ++e) {
        if (!is_free(cur.imu_from[e]) && !is_free(cur.imu_to[e])) continue;
        const int i = (int)w.imu_from.size();
        w.imu_from.push_back(cur.imu_from[e]);
        w.imu_to.push_back(cur.imu_to[e]);
        w.imu.insert(w.imu.end(), cur.imu.begin() + (size_t)VIB_IMU_STRIDE * e, cur.imu.begin() + (size_t)VIB_IMU_STRIDE * (e + 1));
        if (cur.imu_to[e] == prev) edges[0] = i;
        if (cur.imu_from[e] == prev && cur.imu_to[e] == last) edges[1] = i;
    }
    w.E = (int)w.imu_from.size();
    if (w.E != 2 || edges[0] < 0 || edges[1] < 0) return fail(ctx, VIB_ERR_IMU_BRIDGE, "two-frame IMU bridge missing");
    // the prior takes part when one of its blocks is a free block
    bool touches = false;
    for (size_t b = 0; b < cur.prior_kind.size(); ++b)
        if ((cur.prior_kind[b] == VIB_KEY_POSE || cur.prior_kind[b] == VIB_KEY_SPEED_BIAS) &&
            (cur.prior_id[b] == cur.frame_id[prev] || cur.prior_id[b] == cur.frame_id[last]))
            touches = true;
    if (!touches) {
        w.prior_rows = w.prior_cols = 0;
        w.prior_kind.clear(); w.prior_id.clear(); w.prior_ref.clear(); w.prior_A.clear(); w.prior_r0.clear();
    }
    EditSpec sp{};
    for (int k = 0; k < kMaxFrames; ++k) sp.renum[k] = k < cur.K ? k : -1;
    sp.lm_mode = 2;  // landmarks held constant
    sp.keep_lm = eb.keep_lm;
    sp.keep_f = eb.keep_f;
    if (!ctx->two_frame_slot) ctx->two_frame_slot = std::make_unique<Slot>();
    Slot& local = *ctx->two_frame_slot;  // its arena is kept from call to call
    int st = edit_window(ctx, *s, sp, std::move(w), local, eb, false);
    if (st != VIB_OK) return st;
    double nis[2];
    if (!imu_nis(ctx, local, edges, nis)) return fail(ctx, VIB_ERR_DEVICE, "IMU residual evaluation failed");
    rep->old_imu_nis_before = nis[0];
    rep->new_imu_nis_before = nis[1];
    if (!std::isfinite(nis[0]) || !std::isfinite(nis[1]))
        return fail(ctx, VIB_ERR_IMU_BRIDGE, "two-frame IMU residual invalid before solve");
    vib_solve_options lo = *o;  // budget of the lightweight solve (BackendOptimizer.cpp:549-553)
    lo.max_iterations = std::min(3, std::max(1, o->max_iterations));
    lo.min_iterations = std::min(1, std::max(0, o->min_iterations));
    lo.time_limit_seconds = o->time_limit_seconds > 0 ? std::min(0.008, o->time_limit_seconds) : 0.008;
    lo.flags &= ~VIB_SOLVE_CONTINUOUS;
    std::vector<Slot*> one{&local};
    std::vector<vib_solve_report> reps;
    st = run_lm(ctx, one, lo, reps);
    if (!reps.empty()) rep->solve = reps[0];
    if (st != VIB_OK) return st;
    if (!imu_nis(ctx, local, edges, nis)) return fail(ctx, VIB_ERR_DEVICE, "IMU residual evaluation failed");
    rep->old_imu_nis_after = nis[0];
    rep->new_imu_nis_after = nis[1];
    if (!std::isfinite(nis[0]) || !std::isfinite(nis[1]))
        return fail(ctx, VIB_ERR_IMU_BRIDGE, "two-frame IMU bridge invalid after solve");
    if (gate && nis[1] > 60.0)
        return fail(ctx, VIB_ERR_IMU_GATE, "new IMU edge residual exceeded lightweight health gate");
    // write back poses and velocities of K-2 and K-1 (biases are constant in this solve)
    bool ok = hipMemcpyAsync(s->dev.pose + 7 * prev, local.dev.pose + 7 * prev, 14 * sizeof(double),
                             hipMemcpyDeviceToDevice, ctx->stream) == hipSuccess;
    for (int k = prev; k < prev + 2; ++k)
        ok = ok && hipMemcpyAsync(s->dev.sb + 9 * k, local.dev.sb + 9 * k, 3 * sizeof(double), hipMemcpyDeviceToDevice,
                                  ctx->stream) == hipSuccess;
    ok = ok && hipStreamSynchronize(ctx->stream) == hipSuccess;
    return ok ? VIB_OK : fail(ctx, VIB_ERR_DEVICE, "state write-back failed");
}

namespace {

// A factor with zero square-root information was rejected by the statistical gate. Window
// maintenance archives rejected factors before any marginalization linearization
// (BackendOptimizer.cpp:1053-1054): the window is rebuilt without them, and without the landmarks
// whose every factor was rejected, before the slide.
static int strip_rejected(vib_context* ctx, Slot& s) {
    const HostWindow& h = s.host;
    if (h.L == 0 || h.F == 0) return VIB_OK;
    EditBufs b;
    if (!edit_buffers(ctx, h.L, h.F, b)) return fail(ctx, VIB_ERR_OUT_OF_MEMORY, "slide buffer allocation failed");
    int rejected = 0;
    CopyBatch cb;
    cb.down(&rejected, b.sel, sizeof(int));
    if (!edit_begin(ctx, s, b) || vibk_strip_select(b.src, h.L, b.keep_lm, b.keep_f, b.sel, ctx->stream) != hipSuccess ||
        !cb.flush(ctx->stream, ctx->xfer))
        return fail(ctx, VIB_ERR_DEVICE, "rejected-factor scan failed");
    if (!rejected) return VIB_OK;
    if (!refresh_state(ctx, s)) return fail(ctx, VIB_ERR_DEVICE, "state download failed");
    EditSpec sp{};
    for (int k = 0; k < kMaxFrames; ++k) sp.renum[k] = k < h.K ? k : -1;
    sp.keep_lm = b.keep_lm;
    sp.keep_f = b.keep_f;
    return edit_window(ctx, s, sp, frame_meta(h), s, b, true);
}

// OLD slides of a batch of windows: slot 0 of each with every landmark anchored in or observed
// by it and its IMU edges (Marginalizer::compute with the elimination set of
// SlidingWindowManager.cpp:627-652). Per window: rejected observations stripped, the eliminated
// set and the frames it touches found on the device; then the marginalizations of all windows
// as one batch (schur_priors); then each window slides (a device edit in place). reps[i]
// receives window i's result; the return value is the first failure, if any.
int marginalize_old_batch(vib_context* ctx, const std::vector<Slot*>& slots, vib_marginal_report* reps) {
    const int victim = 0;
    const size_t n = slots.size();
    std::vector<EditBufs> bufs(n);
    std::vector<PriorJob> jobs;
    std::vector<size_t> job_of;  // window of each job
    int first = VIB_OK;
    auto note = [&](size_t i, int st) {
        reps[i].status = st;
        if (first == VIB_OK) first = st;
    };
    for (size_t i = 0; i < n; ++i) {
        reps[i] = vib_marginal_report{};
        Slot& s = *slots[i];
        int st = strip_rejected(ctx, s);
        if (st == VIB_OK && !refresh_state(ctx, s)) st = fail(ctx, VIB_ERR_DEVICE, "state download failed");
        if (st == VIB_OK && s.host.K < 2) st = fail(ctx, VIB_ERR_INVALID_ARGUMENT, "marginalization needs at least two frames");
        if (st == VIB_OK && !job_edit_buffers(ctx, i, s.host.L, s.host.F, bufs[i]))
            st = fail(ctx, VIB_ERR_OUT_OF_MEMORY, "slide buffer allocation failed");
        const HostWindow& cur = s.host;
        PriorJob j;
        if (st == VIB_OK &&
            (!edit_begin(ctx, s, bufs[i]) ||
             (cur.L && vibk_elim_set(bufs[i].src, cur.L, victim, bufs[i].marg_lm, ctx->stream) != hipSuccess) ||
             !touched_frames(ctx, bufs[i], cur.L, bufs[i].marg_lm, j.frames, j.observed)))
            st = fail(ctx, VIB_ERR_DEVICE, "eliminated-set launch failed");
        if (st != VIB_OK) {
            note(i, st);
            continue;
        }
        j.src = &s;
        j.eb = &bufs[i];
        j.d_sel = bufs[i].marg_lm;
        j.victim = victim;
        for (int e = 0; e < cur.E; ++e)
            if (cur.imu_from[e] == victim || cur.imu_to[e] == victim) {
                j.e_from.push_back(cur.imu_from[e]);
                j.e_to.push_back(cur.imu_to[e]);
                j.e_rec.insert(j.e_rec.end(), cur.imu.begin() + (size_t)VIB_IMU_STRIDE * e,
                               cur.imu.begin() + (size_t)VIB_IMU_STRIDE * (e + 1));
            }
        j.marg = &marginal_slot(ctx, jobs.size());
        jobs.push_back(std::move(j));
        job_of.push_back(i);
    }
    if (const int st = schur_priors(ctx, jobs)) {
        for (size_t q = 0; q < jobs.size(); ++q) note(job_of[q], st);
        return first;
    }
    for (size_t q = 0; q < jobs.size(); ++q) {
        PriorJob& j = jobs[q];
        const size_t i = job_of[q];
        if (j.status != VIB_OK) {
            ctx->error = j.error;
            note(i, j.status);
            continue;
        }
        // slide: drop slot 0, its landmarks, factors and IMU edges
        Slot& s = *slots[i];
        const HostWindow& cur = s.host;
        HostWindow next = frame_meta(cur);
        drop_frame(next, victim);
        keep_edges(next, cur.imu_from, cur.imu_to, cur.imu, std::vector<char>(cur.E, 1), victim);
        set_prior(next, j.out);
        EditSpec sp{};
        for (int k = 0; k < kMaxFrames; ++k) sp.renum[k] = k < cur.K ? (k == victim ? -1 : (k > victim ? k - 1 : k)) : -1;
        sp.keep_lm = bufs[i].marg_lm;
        sp.keep_lm_invert = 1;
        const bool bridged = !j.e_from.empty();
        if (const int st = edit_window(ctx, s, sp, std::move(next), s, bufs[i], true)) {
            note(i, st);
            continue;
        }
        vib_marginal_report& rep = reps[i];
        rep.status = VIB_OK;
        rep.prior_reused = 0;
        rep.eliminated_dimension = j.out.eliminated;
        rep.retained_dimension = j.out.cols;
        rep.prior_rank = j.out.rows;
        rep.eliminated_landmarks = j.out.landmarks;
        rep.consumed_factors = j.out.consumed;
        rep.eigen_cutoff = j.out.cutoff;
        rep.imu_mode = bridged ? VIB_IMU_MERGE_BRIDGE : VIB_IMU_MERGE_NONE;
    }
    return first;
}

}  // namespace

int vib_marginalize(vib_context* ctx, int32_t slot, int32_t policy, const double* samples, int32_t count,
                    double t_start, double t_end, const vib_imu_params* params, vib_marginal_report* rep) {
    if (!ctx || !rep) return VIB_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mutex);
    (void)hipSetDevice(ctx->device);
    *rep = vib_marginal_report{};
    Slot* s = find_slot(ctx, slot);
    if (!s) return fail(ctx, VIB_ERR_NOT_FOUND, "unknown slot");
    if (policy == VIB_MARGIN_OLD) return marginalize_old_batch(ctx, {s}, rep);
    if (policy != VIB_MARGIN_SECOND_NEW) return fail(ctx, VIB_ERR_INVALID_ARGUMENT, "unknown marginalization policy");
    int st = strip_rejected(ctx, *s);
    if (st == VIB_OK) st = marginalize_second_new(ctx, *s, samples, count, t_start, t_end, params, rep);
    rep->status = st;
    return st;
}

int vib_marginalize_batch(vib_context* ctx, int32_t count, const int32_t* ids, int32_t policy,
                          vib_marginal_report* reps) {
    if (!ctx || !ids || !reps || count <= 0) return VIB_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mutex);
    (void)hipSetDevice(ctx->device);
    std::vector<Slot*> slots;
    for (int i = 0; i < count; ++i) {
        reps[i] = vib_marginal_report{};
        Slot* s = find_slot(ctx, ids[i]);
        if (!s) return fail(ctx, VIB_ERR_NOT_FOUND, "unknown slot");
        for (Slot* o : slots)
            if (o == s) return fail(ctx, VIB_ERR_INVALID_ARGUMENT, "a slot is listed twice");
        slots.push_back(s);
    }
    if (policy == VIB_MARGIN_OLD) return marginalize_old_batch(ctx, slots, reps);
    if (policy != VIB_MARGIN_SECOND_NEW) return fail(ctx, VIB_ERR_INVALID_ARGUMENT, "unknown marginalization policy");
    int first = VIB_OK;
    for (int i = 0; i < count; ++i) {  // without IMU samples: the slot's edges are bridged
        int st = strip_rejected(ctx, *slots[i]);
        if (st == VIB_OK) st = marginalize_second_new(ctx, *slots[i], nullptr, 0, 0.0, 0.0, nullptr, &reps[i]);
        reps[i].status = st;
        if (first == VIB_OK) first = st;
    }
    return first;
}

int vib_read_window(vib_context* ctx, int32_t slot, vib_window* out) {
    if (!ctx || !out) return VIB_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mutex);
    (void)hipSetDevice(ctx->device);
    Slot* s = find_slot(ctx, slot);
    if (!s) return fail(ctx, VIB_ERR_NOT_FOUND, "unknown slot");
    if (!ensure_bulk(ctx, *s) || !refresh_state(ctx, *s)) return fail(ctx, VIB_ERR_DEVICE, "window download failed");
    const HostWindow& h = s->host;
    if (out->num_frames != h.K || out->num_cameras != h.C || out->num_landmarks != h.L ||
        out->num_observations != h.F || out->num_imu != h.E || out->prior_rows != h.prior_rows)
        return fail(ctx, VIB_ERR_INVALID_ARGUMENT, "window shape mismatch");
    WindowCopy w = copy_of(h);
    auto put = [](const auto& v, const auto* p) {
        using T = std::remove_const_t<std::remove_pointer_t<decltype(p)>>;
        if (p && !v.empty()) std::copy(v.begin(), v.end(), const_cast<T*>(p));
    };
    out->flags = w.flags; out->huber_delta = w.huber; out->min_inverse_depth = w.min_rho;
    out->td_lower = w.td_lo; out->td_upper = w.td_hi;
    put(w.frame_id, out->frame_id); put(w.pose, out->pose); put(w.sb, out->speed_bias);
    put(w.pose_free, out->pose_free); put(w.sb_width, out->sb_width);
    if (out->rig) std::copy_n(w.rig, 7, out->rig);
    if (out->td) *out->td = w.td;
    put(w.cam_R, out->cam_rotation); put(w.cam_t, out->cam_translation);
    put(w.lm_frame, out->lm_anchor_frame); put(w.lm_cam, out->lm_anchor_cam); put(w.lm_bearing, out->lm_bearing);
    put(w.lm_vel, out->lm_velocity); put(w.lm_tref, out->lm_tref); put(w.rho, out->lm_rho); put(w.lm_free, out->lm_free);
    put(w.ob_lm, out->obs_landmark); put(w.ob_frame, out->obs_frame); put(w.ob_cam, out->obs_cam);
    put(w.ob_bearing, out->obs_bearing); put(w.ob_vel, out->obs_velocity); put(w.ob_tref, out->obs_tref);
    put(w.ob_sqrt, out->obs_sqrt_info);
    put(w.imu_from, out->imu_from); put(w.imu_to, out->imu_to); put(w.imu, out->imu_data);
    out->prior_blocks = (int32_t)w.prior_kind.size();
    put(w.prior_kind, out->prior_key_kind); put(w.prior_id, out->prior_key_id); put(w.prior_ref, out->prior_ref);
    put(w.prior_A, out->prior_A); put(w.prior_r0, out->prior_r0);
    put(h.lm_uid, out->lm_id);
    return VIB_OK;
}

int vib_read_prior(vib_context* ctx, int32_t slot, int32_t* rows, int32_t* cols, int32_t* blocks, int32_t* kind,
                   int64_t* id, double* ref, double* A, double* r0, int32_t cap_rows, int32_t cap_cols, int32_t cap_blocks) {
    if (!ctx || !rows || !cols || !blocks) return VIB_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mutex);
    Slot* s = find_slot(ctx, slot);
    if (!s) return fail(ctx, VIB_ERR_NOT_FOUND, "unknown slot");
    const HostWindow& h = s->host;
    *rows = h.prior_rows; *cols = h.prior_cols; *blocks = (int)h.prior_kind.size();
    if (h.prior_rows > cap_rows || h.prior_cols > cap_cols || (int)h.prior_kind.size() > cap_blocks)
        return fail(ctx, VIB_ERR_INVALID_ARGUMENT, "capacity too small");
    for (size_t b = 0; b < h.prior_kind.size(); ++b) {
        if (kind) kind[b] = h.prior_kind[b];
        if (id) id[b] = h.prior_id[b];
    }
    if (ref) std::copy(h.prior_ref.begin(), h.prior_ref.end(), ref);
    if (A) std::copy(h.prior_A.begin(), h.prior_A.end(), A);
    if (r0) std::copy(h.prior_r0.begin(), h.prior_r0.end(), r0);
    return VIB_OK;
}

}  // extern "C"
