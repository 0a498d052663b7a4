// Host-side window bookkeeping of the Schur backend (topology, layout, device slots, batches).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/vib.h"
#include "vib_device.hpp"

namespace vib {

// Cyclic Jacobi eigen-decomposition of a symmetric n x n matrix (row-major, small n); eigenvalues
// ascending, eigenvectors as columns of V (vib_preint.cpp).
void symmetric_eigen(std::vector<double> A, int n, std::vector<double>& ev, std::vector<double>& V);
constexpr int kKernelCount = 14;  // kernels of one LM iteration (vibk_launch ids 0..13; 12 and 13 are the
                                  // refinement passes of a mixed-precision step)
// Batches with at least this many landmark chunks (about 16 per CU) merge kChunkMerge of them
// per workgroup (measured on the 256-window batched workload).
constexpr size_t kMergeChunksAbove = 4096;
constexpr int kChunkMerge = 4;
// Visual factor ranges (kLinFactors each) are merged kLinMerge at a time per workgroup above
// this many ranges in a batch: the per-workgroup fixed costs (zeroing and writing back the
// dense visual-block accumulator, the frame table) are about 7 us per workgroup.
constexpr size_t kMergeLinAbove = 2048;
constexpr int kLinMerge = 4;
constexpr int kFuseReduceWindows = 16;
constexpr int kXcds = 8;  // accelerator complex dies of an MI355X (workgroup i runs on XCD i % 8)

// Host copy of one window: observation streams sorted by landmark, tangent column layout
// [pose (6, free) ..., rig (6), td (1), speed/bias (width) ...] then free inverse ranges,
// prior block table and the chunk work list.
struct HostWindow {
    static int chunk_size;  // landmarks per chunk workgroup
    int K = 0, C = 0, L = 0, F = 0, E = 0;
    uint32_t flags = 0;
    double huber = 1, min_rho = 1e-8, td_lo = -1e300, td_hi = 1e300;
    std::vector<int64_t> frame_id;
    std::vector<double> pose, sb;
    std::vector<uint8_t> pose_free, sb_width;
    double rig[7]{}, td = 0;
    std::vector<double> cam;  // C x 12
    std::vector<int> lm_frame, lm_cam, lm_off;
    std::vector<uint8_t> lm_free;
    std::vector<double> lmd, rho;
    std::vector<int64_t> lm_uid;  // stable landmark ids, strictly increasing
    int64_t uid_next = 0;         // ids at or above this one were never in the window
    std::vector<int> ob_fc, ob_lm;
    std::vector<double> ob;
    std::vector<int> imu_from, imu_to;
    std::vector<double> imu;
    int prior_rows = 0, prior_cols = 0;
    std::vector<int> prior_kind;
    std::vector<int64_t> prior_id;
    std::vector<double> prior_ref, prior_A, prior_r0;
    // layout
    std::vector<int> pose_col, sb_col, lm_col;
    int rig_col = -1, td_col = -1, nv = 0, nr = 0, ne = 0;
    std::vector<PriorBlockDev> pblk;
    std::vector<int32_t> pfree;
    std::vector<int32_t> colmap;  // nr x 4 retained-factor column map
    std::vector<uint32_t> chol_mask;  // structural tiles of the reduced-system factor (WinDev)
    std::vector<Chunk> chunks;
    // factors in (anchor frame, observer frame) pair order
    std::vector<double> pf_ob, pf_an;
    std::vector<int32_t> pf_ids, pf_lm, pf_pos;
    std::vector<Chunk> lin_chunks;  // factor ranges of the visual-normal kernel
    // The per-landmark and per-observation vectors (lm_frame .. rho, lm_uid, ob_fc, ob_lm, ob,
    // lm_col and the pair-order planes) hold the resident window. False after a device edit: the device
    // holds them, and ensure_bulk downloads them when a host path needs them.
    bool bulk = true;

    // VIB_OK, or VIB_ERR_INVALID_ARGUMENT (malformed), VIB_ERR_UNSUPPORTED_DIMENSIONS (beyond the
    // window limits) or VIB_ERR_UNSUPPORTED_LOSS, with the message in err
    int build(const vib_window& w, std::string& err);
    bool layout(std::string& err);
    // the state-column part of layout: columns, prior blocks, retained column map, factor tiles
    bool layout_states(std::string& err);
    // the landmark chunks and visual factor ranges from L and F
    void make_chunks();
};

// Grow-only pinned host buffer (hipHostMalloc): transfers from and to it are plain DMA, with no
// staging copy inside the runtime (whose pageable-memory path serializes the transfers of all
// threads of the process).
struct PinnedBuffer {
    char* p = nullptr;
    size_t cap = 0;
    bool reserve(size_t bytes);
    ~PinnedBuffer();
};

// Staging of a context's batched transfers: a grow-only device buffer, and two pinned host
// buffers used in turn, each reused only after the transfer that last read it has completed
// (`done`), and one for downloads.
struct Staging {
    void* buf = nullptr;
    size_t cap = 0;
    PinnedBuffer host[2], back;
    hipEvent_t done[2] = {nullptr, nullptr};
    bool recorded[2] = {false, false};
    int next = 0;
    ~Staging();
};

// The small copies of one host call as one transfer and one kernel launch. Each hipMemcpyAsync
// of pageable host memory costs several microseconds of host time (a staging copy and a copy
// launch), and a window edit used to make dozens of them. Uploads are packed into one host
// buffer, moved in one host-to-device transfer and scattered to their destinations by one
// launch (k_copy_segments); downloads are gathered by one launch into the staging buffer and
// moved in one device-to-host transfer. The batch owns its host buffers, so its sources may be
// temporaries, and the batch may be dropped once flushed.
class CopyBatch {
  public:
    void up(void* dst, const void* src, size_t bytes);  // host src, copied into the batch now
    void zero(void* dst, size_t bytes);
    void down(void* host, const void* dev, size_t bytes);  // valid after flush returns true
    // Issues the uploads on `s`; with downloads also gathers them, synchronizes `s` and unpacks
    // them into their host destinations. Every pending copy is consumed.
    bool flush(hipStream_t s, Staging& st);

  private:
    struct Up { void* dst; size_t off; size_t bytes; };  // off: payload offset (SIZE_MAX: zeros)
    struct Down { void* host; const void* dev; size_t bytes; };
    std::vector<Up> ups_;
    std::vector<Down> downs_;
    std::vector<char> pay_;
};

// Device pointers of a slot's arena that the window descriptor does not carry.
struct SlotCarve {
    double* reduced_lin = nullptr;
    double* reduced_schur = nullptr;
    WinDev* gram_win = nullptr;  // descriptor and output pointer of the prior Gram launch
    double** gram_out = nullptr;
};

// One device-resident window (a session's window and prior live here across solves).
struct Slot {
    bool used = false;
    HostWindow host;
    WinDev dev{};
    void* arena = nullptr;
    size_t bytes = 0;     // footprint of the resident window
    size_t capacity = 0;  // size of the arena (at least bytes)
    double* reduced_lin = nullptr;
    double* reduced_schur = nullptr;
    int lin_precision = VIB_FP64;  // coupling-stream format of the last vib_linearize
    double* host_terms = nullptr;  // nr x nr + nr: normal equations of host-evaluated factors
    size_t host_cap = 0;           // doubles of host_terms
    bool host_set = false;         // host terms are added at the next linearizations

    ~Slot();
    void release();
    static size_t footprint(const HostWindow& h);
    // Lays window h out over [base, base + bytes): the descriptor d and the other pointers x;
    // false if the footprint does not fit.
    static bool carve(const HostWindow& h, char* base, size_t bytes, WinDev& d, SlotCarve& x);
    // Copies the host-held parts of window h into a carved arena (state and its trial copy,
    // cameras, IMU, prior, retained column map, LM state; with streams also the landmark and
    // observation streams in both orders), then forms the prior's Gram matrix; returns after
    // `stream` has completed them, false if a copy or launch fails.
    // The small copies go through one batched transfer (`cb`, which may already hold the
    // caller's copies).
    static bool fill(const HostWindow& h, const WinDev& d, const SlotCarve& x, char* base, size_t bytes, bool streams,
                     hipStream_t stream, Staging& st, CopyBatch& cb);
    // VIB_OK, VIB_ERR_OUT_OF_MEMORY (allocation) or VIB_ERR_DEVICE (copies, launches)
    int upload(const HostWindow& h, hipStream_t stream, Staging& st, std::string& err);
};

// Descriptors of a set of windows solved together.
struct Batch {
    int nwin = 0, nchunks = 0, nlin = 0, max_len = 1;
    int max_edges = 0;
    int max_landmarks = 0;
    int mixed = 0;  // 1: the mixed-precision forms of the coupling and Schur kernels
    // landmark chunks / visual factor ranges merged per workgroup (0: the built-in rule, merged
    // by kChunkMerge / kLinMerge above kMergeChunksAbove / kMergeLinAbove; vib_autotune sets them)
    int chunk_merge = 0, lin_merge = 0;
    WinDev* d_wins = nullptr;
    Chunk* d_chunks = nullptr;
    Chunk* d_lin = nullptr;
    double** d_red = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    size_t cap_wins = 0, cap_red = 0, cap_chunks = 0, cap_lin = 0;  // grow-only device buffers (rebuilt batches)

    ~Batch();
    bool build(std::vector<Slot*>& slots, hipStream_t stream, Staging& xs, std::string& err);
    bool launch(int kernel, hipStream_t s, int refinement, int mode);
    bool linearize(hipStream_t s);
    bool step(hipStream_t s, int refinement, int mode);
    bool iteration(hipStream_t s, int refinement);
    bool fuse_reduce() const;        // kernel 3 also sums the visual-block partials
    // launches of one LM iteration (iteration = true) or of a step alone, as (kernel, mode)
    std::vector<std::pair<int, int>> sequence(bool iteration, int refinement, int mode) const;
    bool capture(hipStream_t s, int refinement);
};

}  // namespace vib
