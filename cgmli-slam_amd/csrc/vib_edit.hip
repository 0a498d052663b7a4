// Window edits on the device: a slide, the marginal system of a slide, the stripping of
// rejected observations and the two-frame window are all "keep these landmarks and
// observations, renumber these frames" applied to a resident window. The edited window is
// written into another arena by stream compaction (exclusive scans of keep flags), so the host
// never re-sorts or re-uploads the observation streams. The result is the layout
// HostWindow::build + layout would produce for the same window: observations in landmark order,
// then the stable (anchor frame, observer frame) pair order of the visual-normal kernel.
//
// Topology rules follow SlidingWindowManager.cpp:263-470 (SECOND_NEW: landmarks anchored in the
// redundant frame are reset to a surviving view, demoted or retired; the frame's other image
// measurements are discarded) and :627-652 (OLD: landmarks anchored in or observed from the
// oldest frame are eliminated).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "vib_device.hpp"

namespace vib {

constexpr int kEditBlock = 256;
constexpr int kScanThreads = 1024;
constexpr int kScanPer = 4;                               // elements per thread of a scan tile
constexpr int kScanTile = kScanThreads * kScanPer;        // elements per scan workgroup
constexpr int kPairKeys = kMaxFrames * kMaxFrames;        // (anchor frame, observer frame) keys
constexpr int kSortBlock = 1024;                          // new observations per sort workgroup


// ---------------------------------------------------------------------------------------------
// Exclusive scan of n values (uint32 or a pair of uint32 counters packed in a uint64), out[n] =
// total: tile sums, a one-workgroup scan of the tile sums, then each tile rescanned with its
// offset. Tiles are kScanTile long; the tile-sum scan loops, so any n < 2^31 works.
// ---------------------------------------------------------------------------------------------
template <typename T>
__device__ T block_exclusive_scan(T v, T* lds, T& total) {
    const int t = threadIdx.x, lane = t % 64, wave = t / 64, nw = blockDim.x / 64;
    T x = v;
    for (int d = 1; d < 64; d <<= 1) {
        const T y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) lds[wave] = x;
    __syncthreads();
    if (t == 0) {
        T run = 0;
        for (int w = 0; w < nw; ++w) {
            const T s = lds[w];
            lds[w] = run;
            run += s;
        }
        lds[nw] = run;
    }
    __syncthreads();
    total = lds[nw];
    const T r = x - v + lds[wave];
    __syncthreads();
    return r;
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_tiles(const T* in, int n, T* tile_sum) {
    __shared__ T lds[kScanThreads / 64 + 1];
    const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanPer;
    T s = 0;
#pragma unroll
    for (int j = 0; j < kScanPer; ++j)
        if (base + j < (size_t)n) s += in[base + j];
    T total;
    (void)block_exclusive_scan(s, lds, total);
    if (threadIdx.x == 0) tile_sum[blockIdx.x] = total;
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_sums(T* tile_sum, int tiles) {
    __shared__ T lds[kScanThreads / 64 + 1];
    T carry = 0;
    for (int b0 = 0; b0 < tiles; b0 += kScanThreads) {
        const int i = b0 + threadIdx.x;
        const T v = i < tiles ? tile_sum[i] : T(0);
        T total;
        const T e = block_exclusive_scan(v, lds, total);
        if (i < tiles) tile_sum[i] = carry + e;
        carry += total;
    }
    if (threadIdx.x == 0) tile_sum[tiles] = carry;
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const T* in, int n, const T* tile_sum, T* out) {
    __shared__ T lds[kScanThreads / 64 + 1];
    const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanPer;
    T v[kScanPer];
    T s = 0;
#pragma unroll
    for (int j = 0; j < kScanPer; ++j) {
        v[j] = base + j < (size_t)n ? in[base + j] : T(0);
        s += v[j];
    }
    T total;
    T run = block_exclusive_scan(s, lds, total) + tile_sum[blockIdx.x];
#pragma unroll
    for (int j = 0; j < kScanPer; ++j)
        if (base + j < (size_t)n) {
            out[base + j] = run;
            run += v[j];
        }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = tile_sum[gridDim.x];
}

// ---------------------------------------------------------------------------------------------
// k_edit_mark: one thread per source landmark. Landmark flag: 1 (kept) | (kept and free) << 32;
// observation flag: kept landmark and kept observation; cnt[l]: the landmark's observations in
// the edited window (kept ones and inserted ones); lm_of[f] = the observation's landmark.
// bad[0] counts kept landmarks anchored in a removed frame and kept observations from one.
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kEditBlock) void k_edit_mark(const WinDev* srcp, EditSpec sp, unsigned long long* lmv,
                                                          uint32_t* fv, uint32_t* cnt, int32_t* lm_of, int* bad) {
    const WinDev& s = *srcp;
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= s.L) return;
    const bool kl = sp.keep_lm ? (sp.keep_lm[l] != 0) != (sp.keep_lm_invert != 0) : true;
    int anchor = s.lm_fc[l] >> 8;
    if (sp.over_f && sp.over_f[l] >= 0) anchor = s.ob_fc[sp.over_f[l]] >> 8;
    int wrong = kl && sp.renum[anchor] < 0;
    const bool fr = sp.lm_mode == 1 ? true : (sp.lm_mode == 2 ? false : s.lm_col[l] >= 0);
    lmv[l] = kl ? (1ull | ((unsigned long long)(fr ? 1 : 0) << 32)) : 0ull;
    uint32_t kept = 0;
    for (int f = s.lm_off[l]; f < s.lm_off[l + 1]; ++f) {
        const bool kf = kl && (sp.keep_f ? sp.keep_f[f] != 0 : true);
        wrong += kf && sp.renum[s.ob_fc[f] >> 8] < 0;
        fv[f] = kf ? 1u : 0u;
        kept += kf;
        lm_of[f] = l;
    }
    cnt[l] = kl ? kept + (sp.add_cnt ? (uint32_t)sp.add_cnt[l] : 0u) : 0u;
    if (wrong) atomicAdd(bad, wrong);
}

// k_insert_mark: the inserted landmarks' flags and observation counts (after the source's).
__global__ __launch_bounds__(kEditBlock) void k_insert_mark(int L, EditSpec sp, unsigned long long* lmv, uint32_t* cnt) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= sp.ins_lm) return;
    const bool fr = sp.lm_mode == 1 ? true : (sp.lm_mode == 2 ? false : sp.ins_free[j] != 0);
    lmv[L + j] = 1ull | ((unsigned long long)(fr ? 1 : 0) << 32);
    cnt[L + j] = sp.add_cnt ? (uint32_t)sp.add_cnt[L + j] : 0u;
}

// k_edit_landmarks: the kept landmarks at their new index (anchor, planes, inverse range and its
// trial copy, rho column, observation offsets, id).
__global__ __launch_bounds__(kEditBlock) void k_edit_landmarks(const WinDev* srcp, const WinDev* dstp, EditSpec sp,
                                                               const unsigned long long* lscan, const uint32_t* cscan) {
    const WinDev& s = *srcp;
    const WinDev& d = *dstp;
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l == 0) writable(d.lm_off)[d.L] = (int32_t)cscan[s.L + sp.ins_lm];
    if (l >= s.L) return;
    const unsigned long long a = lscan[l], b = lscan[l + 1];
    if ((uint32_t)b == (uint32_t)a) return;  // removed
    const int ln = (int)(uint32_t)a, rank = (int)(a >> 32);
    const bool fr = (b >> 32) != (a >> 32);
    const size_t Ls = (size_t)s.L, Ld = (size_t)d.L, Fs = (size_t)s.F;
    int frame = s.lm_fc[l] >> 8, cam = s.lm_fc[l] & 255;
    double pl[LM_PLANES], rho;
    const int fo = sp.over_f ? sp.over_f[l] : -1;
    if (fo >= 0) {
        frame = s.ob_fc[fo] >> 8;
        cam = s.ob_fc[fo] & 255;
        pl[LM_BX] = s.ob[OB_BX * Fs + fo]; pl[LM_BY] = s.ob[OB_BY * Fs + fo]; pl[LM_BZ] = s.ob[OB_BZ * Fs + fo];
        pl[LM_VX] = s.ob[OB_VX * Fs + fo]; pl[LM_VY] = s.ob[OB_VY * Fs + fo]; pl[LM_VZ] = s.ob[OB_VZ * Fs + fo];
        pl[LM_T] = s.ob[OB_T * Fs + fo];
        rho = sp.over_rho[l];
    } else {
#pragma unroll
        for (int q = 0; q < LM_PLANES; ++q) pl[q] = s.lmd[q * Ls + l];
        rho = s.rho[l];
    }
    writable(d.lm_fc)[ln] = (sp.renum[frame] << 8) | cam;
    writable(d.lm_col)[ln] = fr ? sp.nr + rank : -1;
    writable(d.lm_off)[ln] = (int32_t)cscan[l];
#pragma unroll
    for (int q = 0; q < LM_PLANES; ++q) writable(d.lmd)[q * Ld + ln] = pl[q];
    d.rho[ln] = rho;
    d.crho[ln] = rho;
    d.lm_uid[ln] = s.lm_uid[l];
}

// k_insert_landmarks: the inserted landmarks after the kept ones.
__global__ __launch_bounds__(kEditBlock) void k_insert_landmarks(int L, const WinDev* dstp, EditSpec sp,
                                                                 const unsigned long long* lscan, const uint32_t* cscan) {
    const WinDev& d = *dstp;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= sp.ins_lm) return;
    const unsigned long long a = lscan[L + j], b = lscan[L + j + 1];
    const int ln = (int)(uint32_t)a, rank = (int)(a >> 32);
    const bool fr = (b >> 32) != (a >> 32);
    const size_t Ld = (size_t)d.L, Ni = (size_t)sp.ins_lm;
    writable(d.lm_fc)[ln] = sp.ins_lm_fc[j];
    writable(d.lm_col)[ln] = fr ? sp.nr + rank : -1;
    writable(d.lm_off)[ln] = (int32_t)cscan[L + j];
#pragma unroll
    for (int q = 0; q < LM_PLANES; ++q) writable(d.lmd)[q * Ld + ln] = sp.ins_lmd[q * Ni + j];
    d.rho[ln] = sp.ins_rho[j];
    d.crho[ln] = sp.ins_rho[j];
    d.lm_uid[ln] = sp.ins_uid[j];
}

// k_edit_factors: the kept observations at their new index (landmark order: the landmark's
// offset plus the observation's rank among its kept ones), and the new landmark of each (for the
// pair-order sort).
__global__ __launch_bounds__(kEditBlock) void k_edit_factors(const WinDev* srcp, const WinDev* dstp, EditSpec sp,
                                                             const unsigned long long* lscan, const uint32_t* fscan,
                                                             const uint32_t* cscan, const int32_t* lm_of, int32_t* lmn) {
    const WinDev& s = *srcp;
    const WinDev& d = *dstp;
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= s.F) return;
    if (fscan[f + 1] == fscan[f]) return;  // removed
    const int l = lm_of[f];
    const uint32_t fn = cscan[l] + (fscan[f] - fscan[s.lm_off[l]]);
    const size_t Fs = (size_t)s.F, Fd = (size_t)d.F;
    const int fc = s.ob_fc[f];
    writable(d.ob_fc)[fn] = (sp.renum[fc >> 8] << 8) | (fc & 255);
#pragma unroll
    for (int q = 0; q < OB_PLANES; ++q) writable(d.ob)[q * Fd + fn] = s.ob[q * Fs + f];
    lmn[fn] = (int32_t)(uint32_t)lscan[l];
}

// k_insert_factors: the inserted observations after their landmark's kept ones.
__global__ __launch_bounds__(kEditBlock) void k_insert_factors(const WinDev* srcp, const WinDev* dstp, EditSpec sp,
                                                               const unsigned long long* lscan, const uint32_t* fscan,
                                                               const uint32_t* cscan, int32_t* lmn) {
    const WinDev& s = *srcp;
    const WinDev& d = *dstp;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= sp.ins_obs) return;
    const int l = sp.ins_obs_lm[i];
    if (l < 0 || (uint32_t)lscan[l + 1] == (uint32_t)lscan[l]) return;  // unresolved or removed landmark
    const uint32_t kept = l < s.L ? fscan[s.lm_off[l + 1]] - fscan[s.lm_off[l]] : 0u;
    const uint32_t fn = cscan[l] + kept + (uint32_t)sp.ins_obs_rank[i];
    const size_t Fd = (size_t)d.F, Ni = (size_t)sp.ins_obs;
    writable(d.ob_fc)[fn] = sp.ins_obs_fc[i];
#pragma unroll
    for (int q = 0; q < OB_PLANES; ++q) writable(d.ob)[q * Fd + fn] = sp.ins_ob[q * Ni + i];
    lmn[fn] = (int32_t)(uint32_t)lscan[l];
}

// k_insert_resolve: the landmark of each inserted observation from its landmark id (binary search
// over the source window's ids, then the inserted landmarks' ids; both strictly increasing) and
// the inserted observations per landmark. Unresolved ids (landmarks no longer in the window) are
// counted in unmatched[0] and their observations dropped.
__global__ __launch_bounds__(kEditBlock) void k_insert_resolve(const WinDev* srcp, const int64_t* obs_uid, int nobs,
                                                               const int64_t* new_uid, int nnew, int32_t* obs_lm,
                                                               int32_t* add_cnt, int* unmatched) {
    const WinDev& s = *srcp;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nobs) return;
    const int64_t u = obs_uid[i];
    auto find = [&](const int64_t* ids, int n) {
        int lo = 0, hi = n;
        while (lo < hi) {
            const int mid = (lo + hi) / 2;
            if (ids[mid] < u) lo = mid + 1;
            else hi = mid;
        }
        return lo < n && ids[lo] == u ? lo : -1;
    };
    int l = find(s.lm_uid, s.L);
    if (l < 0) {
        const int j = find(new_uid, nnew);
        l = j < 0 ? -1 : s.L + j;
    }
    obs_lm[i] = l;
    if (l < 0) atomicAdd(unmatched, 1);
    else atomicAdd(&add_cnt[l], 1);
}

// Stable rank of each lane's key among the lanes of its wave with the same key (lanes in order),
// and the count of each key in the wave into wcnt (keys < 0: no element).
__device__ int wave_key_rank(int key, int lane, int* wcnt) {
    unsigned long long pending = __ballot(key >= 0);
    int rank = 0;
    while (pending) {
        const int leader = __ffsll(pending) - 1;
        const int k = __shfl(key, leader, 64);
        const unsigned long long m = __ballot(key == k);
        if (key == k) rank = __popcll(m & ((1ull << lane) - 1));
        if (lane == leader) wcnt[k] = __popcll(m);
        pending &= ~m;
    }
    return rank;
}

__device__ __forceinline__ int pair_key(const WinDev& d, const int32_t* lmn, int f) {
    return (d.lm_fc[lmn[f]] >> 8) * kMaxFrames + (d.ob_fc[f] >> 8);
}

// k_pair_hist: per workgroup of kSortBlock new observations, the count of each pair key, stored
// key-major (counts[key * blocks + block]) so that one exclusive scan gives every (key, block)
// its first position in pair order.
__global__ __launch_bounds__(kSortBlock) void k_pair_hist(const WinDev* dstp, const int32_t* lmn, uint32_t* counts,
                                                          int blocks) {
    __shared__ int wcnt[kSortBlock / 64][kPairKeys];
    const WinDev& d = *dstp;
    const int t = threadIdx.x, lane = t % 64, wave = t / 64;
    for (int i = t; i < (kSortBlock / 64) * kPairKeys; i += blockDim.x) (&wcnt[0][0])[i] = 0;
    __syncthreads();
    const int f = blockIdx.x * kSortBlock + t;
    const int key = f < d.F ? pair_key(d, lmn, f) : -1;
    (void)wave_key_rank(key, lane, wcnt[wave]);
    __syncthreads();
    for (int k = t; k < kPairKeys; k += blockDim.x) {
        uint32_t c = 0;
        for (int w = 0; w < kSortBlock / 64; ++w) c += wcnt[w][k];
        counts[(size_t)k * blocks + blockIdx.x] = c;
    }
}

// k_pair_scatter: each new observation's place in pair order (stable: by key, then landmark
// order) and the pair-order planes (observation, anchor, ids, landmark, position).
__global__ __launch_bounds__(kSortBlock) void k_pair_scatter(const WinDev* dstp, const int32_t* lmn,
                                                             const uint32_t* offs, int blocks) {
    __shared__ int wcnt[kSortBlock / 64][kPairKeys];
    const WinDev& d = *dstp;
    const int t = threadIdx.x, lane = t % 64, wave = t / 64;
    for (int i = t; i < (kSortBlock / 64) * kPairKeys; i += blockDim.x) (&wcnt[0][0])[i] = 0;
    __syncthreads();
    const int f = blockIdx.x * kSortBlock + t;
    const int key = f < d.F ? pair_key(d, lmn, f) : -1;
    const int rank = wave_key_rank(key, lane, wcnt[wave]);
    __syncthreads();
    if (key < 0) return;
    int before = 0;
    for (int w = 0; w < wave; ++w) before += wcnt[w][key];
    const size_t p = offs[(size_t)key * blocks + blockIdx.x] + before + rank;
    const size_t F = (size_t)d.F, L = (size_t)d.L;
    const int l = lmn[f];
#pragma unroll
    for (int q = 0; q < OB_PLANES; ++q) writable(d.pf_ob)[q * F + p] = d.ob[q * F + f];
#pragma unroll
    for (int q = 0; q < LM_PLANES; ++q) writable(d.pf_an)[q * F + p] = d.lmd[q * L + l];
    const int a = d.lm_fc[l];
    writable(d.pf_ids)[p] = ((a >> 8) << 24) | ((a & 255) << 16) | d.ob_fc[f];
    writable(d.pf_lm)[p] = l;
    writable(d.pf_pos)[f] = (int32_t)p;
}

// ---------------------------------------------------------------------------------------------
// Selection kernels (one thread per landmark).
// ---------------------------------------------------------------------------------------------
// Observations whose square-root information the gate zeroed are dropped; a landmark that had
// observations and keeps none goes with them. out[0] counts the dropped observations.
__global__ __launch_bounds__(kEditBlock) void k_strip_select(const WinDev* srcp, uint8_t* keep_lm, uint8_t* keep_f,
                                                             int* out) {
    const WinDev& s = *srcp;
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= s.L) return;
    const size_t F = (size_t)s.F;
    const int f0 = s.lm_off[l], f1 = s.lm_off[l + 1];
    int kept = 0;
    for (int f = f0; f < f1; ++f) {
        const bool k = !(s.ob[OB_S00 * F + f] == 0.0 && s.ob[OB_S10 * F + f] == 0.0 && s.ob[OB_S11 * F + f] == 0.0);
        keep_f[f] = k ? 1 : 0;
        kept += k;
    }
    keep_lm[l] = (f1 == f0 || kept) ? 1 : 0;
    if (f1 - f0 - kept) atomicAdd(out, f1 - f0 - kept);
}

// SECOND_NEW classes of the landmarks anchored in the redundant frame v (SlidingWindowManager.cpp
// :312-457): 1 retired (no surviving view, a moving landmark or observation, or a constant
// inverse range), 2 demoted (one surviving view), 3 to be re-anchored; 0 anchored elsewhere.
__global__ __launch_bounds__(kEditBlock) void k_second_new_classify(const WinDev* srcp, int v, uint8_t* cls) {
    const WinDev& s = *srcp;
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= s.L) return;
    if ((s.lm_fc[l] >> 8) != v) { cls[l] = 0; return; }
    const size_t L = (size_t)s.L, F = (size_t)s.F;
    bool eligible = s.lmd[LM_VX * L + l] == 0 && s.lmd[LM_VY * L + l] == 0 && s.lmd[LM_VZ * L + l] == 0;
    int surviving = 0;
    for (int f = s.lm_off[l]; f < s.lm_off[l + 1]; ++f) {
        const double vx = s.ob[OB_VX * F + f], vy = s.ob[OB_VY * F + f], vz = s.ob[OB_VZ * F + f];
        eligible = eligible && vx * vx + vy * vy + vz * vz == 0;
        surviving += (s.ob_fc[f] >> 8) != v;
    }
    cls[l] = (surviving == 0 || !eligible || s.lm_col[l] < 0) ? 1 : (surviving == 1 ? 2 : 3);
}

// SECOND_NEW keep flags of the next window and of the marginal system from the classes and the
// anchor search (chosen[l] >= 0: the new anchor observation). counts: [0] reset, [1] demoted,
// [2] dropped observations of v, [3] retired landmarks.
__global__ __launch_bounds__(kEditBlock) void k_second_new_flags(const WinDev* srcp, int v, const uint8_t* cls,
                                                                 const int32_t* chosen, const double* chosen_rho,
                                                                 uint8_t* keep_lm, uint8_t* keep_f, int32_t* over_f,
                                                                 double* over_rho, uint8_t* marg_lm, int* counts) {
    const WinDev& s = *srcp;
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= s.L) return;
    const int f0 = s.lm_off[l], f1 = s.lm_off[l + 1];
    int c = cls[l];
    if (c == 3 && chosen[l] < 0) c = 1;
    over_f[l] = -1;
    over_rho[l] = 0;
    keep_lm[l] = c == 0 || c == 3;
    marg_lm[l] = c == 1;
    int dropped = 0;
    for (int f = f0; f < f1; ++f) {
        const bool from_v = (s.ob_fc[f] >> 8) == v;
        bool k = true;
        if (c == 0) { k = !from_v; dropped += from_v; }
        else if (c == 3) k = !from_v && f != chosen[l];
        else if (c == 2) k = false;
        keep_f[f] = k ? 1 : 0;
    }
    if (c == 3) {
        over_f[l] = chosen[l];
        over_rho[l] = chosen_rho[l];
        atomicAdd(&counts[0], 1);
    }
    if (c == 2) atomicAdd(&counts[1], 1);
    if (dropped) atomicAdd(&counts[2], dropped);
    if (c == 1) atomicAdd(&counts[3], 1);
}

// Frames touched by the landmarks of `sel` (anchor and observers, bit k of out[0]) and whether any
// of them has an observation (out[1]): the variable blocks of a marginal system.
__global__ __launch_bounds__(kEditBlock) void k_touch_frames(const WinDev* srcp, const uint8_t* sel, int* out) {
    const WinDev& s = *srcp;
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= s.L || !sel[l]) return;
    int mask = 1 << (s.lm_fc[l] >> 8);
    for (int f = s.lm_off[l]; f < s.lm_off[l + 1]; ++f) mask |= 1 << (s.ob_fc[f] >> 8);
    atomicOr(&out[0], mask);
    if (s.lm_off[l + 1] > s.lm_off[l]) atomicOr(&out[1], 1);
}

// Two-frame window (BackendOptimizer.cpp:476-538): observations touching frame prev or last
// (as anchor or observer), the landmarks with any; out[0] counts the landmarks observed from
// `last` among them (the solve's support).
__global__ __launch_bounds__(kEditBlock) void k_two_frame_select(const WinDev* srcp, int prev, int last,
                                                                 uint8_t* keep_lm, uint8_t* keep_f, int* out) {
    const WinDev& s = *srcp;
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= s.L) return;
    const int a = s.lm_fc[l] >> 8;
    const bool af = a == prev || a == last;
    int kept = 0, seen = 0;
    for (int f = s.lm_off[l]; f < s.lm_off[l + 1]; ++f) {
        const int o = s.ob_fc[f] >> 8;
        const bool k = af || o == prev || o == last;
        keep_f[f] = k ? 1 : 0;
        kept += k;
        seen |= k && o == last;
    }
    keep_lm[l] = kept ? 1 : 0;
    if (seen) atomicAdd(out, 1);
}

}  // namespace vib

// ---------------------------------------------------------------------------------------------
// Launch wrappers used by the host driver.
// ---------------------------------------------------------------------------------------------
namespace vib {

// ---------------------------------------------------------------------------------------------
// Batched copies: workgroup (x, y) copies slice y of segment x, in 4-byte words when both ends
// are 4-byte aligned (the tail bytes by the first threads of slice 0), byte by byte otherwise.
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kEditBlock) void k_copy_segments(const CopySeg* segs, int n) {
    if ((int)blockIdx.x >= n) return;
    const CopySeg g = segs[blockIdx.x];
    unsigned char* d = reinterpret_cast<unsigned char*>(g.dst);
    const unsigned char* s = reinterpret_cast<const unsigned char*>(g.src);
    const uint64_t first = (uint64_t)blockIdx.y * kEditBlock + threadIdx.x, step = (uint64_t)gridDim.y * kEditBlock;
    if (((g.dst | g.src) & 3) == 0) {
        const uint64_t words = g.bytes >> 2;
        uint32_t* dw = reinterpret_cast<uint32_t*>(d);
        const uint32_t* sw = reinterpret_cast<const uint32_t*>(s);
        for (uint64_t i = first; i < words; i += step) dw[i] = s ? sw[i] : 0u;
        if (blockIdx.y == 0 && threadIdx.x < (g.bytes & 3)) {
            const uint64_t i = 4 * words + threadIdx.x;
            d[i] = s ? s[i] : 0;
        }
    } else {
        for (uint64_t i = first; i < g.bytes; i += step) d[i] = s ? s[i] : 0;
    }
}

}  // namespace vib

extern "C" {

hipError_t vibk_copy_segments(const vib::CopySeg* segs, int n, size_t max_bytes, hipStream_t s) {
    using namespace vib;
    if (n == 0) return hipSuccess;
    const size_t per = (size_t)4 * kEditBlock;  // bytes of one slice pass
    const int slices = (int)std::min<size_t>(32, std::max<size_t>(1, (max_bytes + per - 1) / per));
    hipLaunchKernelGGL(k_copy_segments, dim3(n, slices), dim3(kEditBlock), 0, s, segs, n);
    return hipGetLastError();
}

int vibk_edit_scan_tiles(int n) { return (n + vib::kScanTile - 1) / vib::kScanTile; }
int vibk_edit_sort_blocks(int n) { return (n + vib::kSortBlock - 1) / vib::kSortBlock; }
int vibk_edit_pair_keys(void) { return vib::kPairKeys; }
size_t vibk_edit_spec_size(void) { return sizeof(vib::EditSpec); }

// Exclusive scan of n uint32 (wide = 0) or packed uint64 (wide = 1) values into out[0..n];
// tile_sum holds vibk_edit_scan_tiles(n) + 1 values of the same type.
hipError_t vibk_edit_scan(const void* in, int n, void* tile_sum, void* out, int wide, hipStream_t s) {
    using namespace vib;
    if (n < 0) return hipErrorInvalidValue;
    const int tiles = (n + kScanTile - 1) / kScanTile;
    if (tiles == 0) {  // out[0] = 0
        return wide ? hipMemsetAsync(out, 0, sizeof(unsigned long long), s) : hipMemsetAsync(out, 0, sizeof(uint32_t), s);
    }
    if (wide) {
        using T = unsigned long long;
        hipLaunchKernelGGL(k_scan_tiles<T>, dim3(tiles), dim3(kScanThreads), 0, s, (const T*)in, n, (T*)tile_sum);
        hipLaunchKernelGGL(k_scan_sums<T>, dim3(1), dim3(kScanThreads), 0, s, (T*)tile_sum, tiles);
        hipLaunchKernelGGL(k_scan_apply<T>, dim3(tiles), dim3(kScanThreads), 0, s, (const T*)in, n, (const T*)tile_sum,
                           (T*)out);
    } else {
        using T = uint32_t;
        hipLaunchKernelGGL(k_scan_tiles<T>, dim3(tiles), dim3(kScanThreads), 0, s, (const T*)in, n, (T*)tile_sum);
        hipLaunchKernelGGL(k_scan_sums<T>, dim3(1), dim3(kScanThreads), 0, s, (T*)tile_sum, tiles);
        hipLaunchKernelGGL(k_scan_apply<T>, dim3(tiles), dim3(kScanThreads), 0, s, (const T*)in, n, (const T*)tile_sum,
                           (T*)out);
    }
    return hipGetLastError();
}

// spec points to a host vib::EditSpec (passed by value to the kernels).
hipError_t vibk_edit_mark(const vib::WinDev* src, int L, const void* spec, unsigned long long* lmv, uint32_t* fv,
                          uint32_t* cnt, int32_t* lm_of, int* bad, hipStream_t s) {
    using namespace vib;
    const EditSpec& sp = *static_cast<const EditSpec*>(spec);
    if (L > 0)
        hipLaunchKernelGGL(k_edit_mark, dim3((L + kEditBlock - 1) / kEditBlock), dim3(kEditBlock), 0, s, src, sp, lmv, fv,
                           cnt, lm_of, bad);
    if (sp.ins_lm > 0)
        hipLaunchKernelGGL(k_insert_mark, dim3((sp.ins_lm + kEditBlock - 1) / kEditBlock), dim3(kEditBlock), 0, s, L, sp,
                           lmv, cnt);
    return hipGetLastError();
}

hipError_t vibk_edit_gather(const vib::WinDev* src, const vib::WinDev* dst, int Ls, int Fs, const void* spec,
                            const unsigned long long* lscan, const uint32_t* fscan, const uint32_t* cscan,
                            const int32_t* lm_of, int32_t* lmn, hipStream_t s) {
    using namespace vib;
    const EditSpec& sp = *static_cast<const EditSpec*>(spec);
    hipLaunchKernelGGL(k_edit_landmarks, dim3((Ls + kEditBlock) / kEditBlock), dim3(kEditBlock), 0, s, src, dst, sp, lscan,
                       cscan);
    if (sp.ins_lm > 0)
        hipLaunchKernelGGL(k_insert_landmarks, dim3((sp.ins_lm + kEditBlock - 1) / kEditBlock), dim3(kEditBlock), 0, s, Ls,
                           dst, sp, lscan, cscan);
    if (Fs > 0)
        hipLaunchKernelGGL(k_edit_factors, dim3((Fs + kEditBlock - 1) / kEditBlock), dim3(kEditBlock), 0, s, src, dst, sp,
                           lscan, fscan, cscan, lm_of, lmn);
    if (sp.ins_obs > 0)
        hipLaunchKernelGGL(k_insert_factors, dim3((sp.ins_obs + kEditBlock - 1) / kEditBlock), dim3(kEditBlock), 0, s, src,
                           dst, sp, lscan, fscan, cscan, lmn);
    return hipGetLastError();
}

hipError_t vibk_insert_resolve(const vib::WinDev* src, const int64_t* obs_uid, int nobs, const int64_t* new_uid, int nnew,
                               int32_t* obs_lm, int32_t* add_cnt, int* unmatched, hipStream_t s) {
    using namespace vib;
    if (nobs == 0) return hipSuccess;
    hipLaunchKernelGGL(k_insert_resolve, dim3((nobs + kEditBlock - 1) / kEditBlock), dim3(kEditBlock), 0, s, src, obs_uid,
                       nobs, new_uid, nnew, obs_lm, add_cnt, unmatched);
    return hipGetLastError();
}

// counts/offs: vibk_edit_pair_keys() * vibk_edit_sort_blocks(F) + 1 values each.
hipError_t vibk_edit_pair_hist(const vib::WinDev* dst, int F, const int32_t* lmn, uint32_t* counts, hipStream_t s) {
    using namespace vib;
    const int blocks = (F + kSortBlock - 1) / kSortBlock;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pair_hist, dim3(blocks), dim3(kSortBlock), 0, s, dst, lmn, counts, blocks);
    return hipGetLastError();
}

hipError_t vibk_edit_pair_scatter(const vib::WinDev* dst, int F, const int32_t* lmn, const uint32_t* offs, hipStream_t s) {
    using namespace vib;
    const int blocks = (F + kSortBlock - 1) / kSortBlock;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pair_scatter, dim3(blocks), dim3(kSortBlock), 0, s, dst, lmn, offs, blocks);
    return hipGetLastError();
}

hipError_t vibk_strip_select(const vib::WinDev* src, int L, uint8_t* keep_lm, uint8_t* keep_f, int* out, hipStream_t s) {
    using namespace vib;
    if (L == 0) return hipSuccess;
    hipLaunchKernelGGL(k_strip_select, dim3((L + kEditBlock - 1) / kEditBlock), dim3(kEditBlock), 0, s, src, keep_lm, keep_f,
                       out);
    return hipGetLastError();
}

hipError_t vibk_second_new_classify(const vib::WinDev* src, int L, int v, uint8_t* cls, hipStream_t s) {
    using namespace vib;
    if (L == 0) return hipSuccess;
    hipLaunchKernelGGL(k_second_new_classify, dim3((L + kEditBlock - 1) / kEditBlock), dim3(kEditBlock), 0, s, src, v, cls);
    return hipGetLastError();
}

hipError_t vibk_second_new_flags(const vib::WinDev* src, int L, int v, const uint8_t* cls, const int32_t* chosen,
                                 const double* chosen_rho, uint8_t* keep_lm, uint8_t* keep_f, int32_t* over_f,
                                 double* over_rho, uint8_t* marg_lm, int* counts, hipStream_t s) {
    using namespace vib;
    if (L == 0) return hipSuccess;
    hipLaunchKernelGGL(k_second_new_flags, dim3((L + kEditBlock - 1) / kEditBlock), dim3(kEditBlock), 0, s, src, v, cls,
                       chosen, chosen_rho, keep_lm, keep_f, over_f, over_rho, marg_lm, counts);
    return hipGetLastError();
}

hipError_t vibk_touch_frames(const vib::WinDev* src, int L, const uint8_t* sel, int* out, hipStream_t s) {
    using namespace vib;
    if (L == 0) return hipSuccess;
    hipLaunchKernelGGL(k_touch_frames, dim3((L + kEditBlock - 1) / kEditBlock), dim3(kEditBlock), 0, s, src, sel, out);
    return hipGetLastError();
}

hipError_t vibk_two_frame_select(const vib::WinDev* src, int L, int prev, int last, uint8_t* keep_lm, uint8_t* keep_f,
                                 int* out, hipStream_t s) {
    using namespace vib;
    if (L == 0) return hipSuccess;
    hipLaunchKernelGGL(k_two_frame_select, dim3((L + kEditBlock - 1) / kEditBlock), dim3(kEditBlock), 0, s, src, prev, last,
                       keep_lm, keep_f, out);
    return hipGetLastError();
}

}  // extern "C"
