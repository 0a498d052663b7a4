// Device-resident window layout shared by the HIP kernels and the host driver.
//
// One window = one sliding window of one session. Observation streams are stored SoA in
// planes of length F, sorted by landmark (CSR offsets per landmark), so a group of lanes walks
// one landmark's factors with coalesced loads. State lives in HBM for the whole solve; the
// host only reads a handful of scalars per LM iteration.
//
// Batched kernels walk a work list of Chunks (a contiguous landmark range of one window);
// per-chunk partial sums are reduced in a fixed order by the per-window kernels, so a
// window's result does not depend on which other windows share the launch.
#pragma once

#include <cstdint>

namespace vib {

constexpr int kMaxFrames = 20;      // LDS accumulator sized for nv = 6*20 + 7 = 127
constexpr int kMaxCameras = 4;
constexpr int kMaxVisual = 6 * kMaxFrames + 7;
constexpr int kMaxRetained = 15 * kMaxFrames + 7;
constexpr int kChunkLanes = 8;      // lanes per landmark in the per-landmark kernels
constexpr int kBlock = 256;         // threads per chunk workgroup
constexpr int kTile = 16;           // tile edge of the reduced-system factorization
constexpr int kMaxTiles = (kMaxRetained + kTile - 1) / kTile;

// Tile store of the reduced system / Cholesky factor: the lower-triangle 16x16 tiles of the
// padded system (16 T x 16 T, T = ceil(nr / 16), padding rows and columns = identity), tile
// (ti, tj <= ti) at (ti (ti + 1) / 2 + tj) * 256, element (i, j) of a tile at
// (i / 4) * 64 + 16 (i % 4) + j: the accumulator layout of v_mfma_f64_16x16x4f64, so a wave
// moves a tile with four fully coalesced 512-byte accesses.
__host__ __device__ inline int tile_base(int ti, int tj) { return (ti * (ti + 1) / 2 + tj) * kTile * kTile; }
// Off-diagonal tiles below the first tile column (ti > tj >= 1) of a T x T tile system in the
// order of decreasing tj, then increasing ti: index e -> (ti, tj). The factorization keeps the
// leading ones in registers for the whole elimination (they are updated most often).
__host__ __device__ inline void register_tile(int T, int e, int& ti, int& tj) {
    int d = 1;
    while (d * (d + 1) / 2 <= e) ++d;
    tj = T - 1 - d;
    ti = tj + 1 + (e - d * (d - 1) / 2);
}
// The same tiles in reverse order (increasing tj, then decreasing ti): index f -> (ti, tj).
__host__ __device__ inline void store_tile(int T, int f, int& ti, int& tj) {
    register_tile(T, (T - 1) * (T - 2) / 2 - 1 - f, ti, tj);
}
__host__ __device__ inline size_t factor_doubles(int nr) {
    const size_t T = (size_t)(nr + kTile - 1) / kTile;
    return T * (T + 1) / 2 * kTile * kTile;
}

// observation planes (doubles, each F long)
enum ObsPlane { OB_BX, OB_BY, OB_BZ, OB_VX, OB_VY, OB_VZ, OB_T, OB_S00, OB_S10, OB_S11, OB_PLANES };
// landmark planes (doubles, each L long)
enum LmPlane { LM_BX, LM_BY, LM_BZ, LM_VX, LM_VY, LM_VZ, LM_T, LM_PLANES };
// per-landmark linearization outputs (doubles, each L long)
enum LmOut { LO_D, LO_G, LO_CR0, LO_CR1, LO_CR2, LO_CR3, LO_CR4, LO_CR5, LO_CT, LO_PLANES };
// per-factor landmark terms, summed per landmark into LmOut (doubles, each F long)
enum FacLm { FL_D, FL_G, FL_CR0, FL_CR1, FL_CR2, FL_CR3, FL_CR4, FL_CR5, FL_CT, FL_PLANES };
// per-chunk scalars
enum ChunkScalar { CS_COST, CS_INVALID, CS_PRED, CS_MAXINC, CS_TRIAL, CS_TINVALID, CS_NONFINITE, CS_PLANES };

// LM bookkeeping per window (BackendOptimizer.cpp:167-307), updated on the device.
struct LmState {
    double lambda, increase, cost, initial_cost;
    double lin_cost;        // cost reported by the last linearization
    double predicted, trial_cost, max_increment;
    double last_gain;
    double ret_cost;        // retained (IMU + prior) cost at the linearization point
    int32_t ret_invalid;
    int32_t new_point;      // linearize at the start of the next iteration
    int32_t status;         // 0 running, 1 finished (converged / budget), <0 -status error
    int32_t iteration;      // iterations started
    int32_t accepted;
    int32_t converged;
    int32_t lin_valid, chol_ok, trial_valid;
    int32_t step_nonfinite; // some entry of the last step (states or inverse ranges) is not finite
    int32_t max_iterations;
    int32_t evaluate_only;  // piecewise API: compute metrics, never accept
    int32_t continuous;     // benchmark mode: relinearize every iteration, never terminate
    int32_t accept_pending; // the last update accepted the trial state (k_accept copies it)
    // marginalization (landmark-first Schur with a pseudo-inverse cutoff; the state block and
    // the prior's square root follow from k_sym_eigen; Marginalizer.cpp:356-474)
    int32_t marginal;       // 1: build the marginal system instead of an LM step
    double marg_cut;        // landmark cutoff max(1e-12, 1e-9 max d)
};

struct PriorBlockDev {
    int32_t kind;      // VIB_KEY_*
    int32_t slot;      // frame slot (POSE/SB)
    int32_t col;       // tangent column or -1
    int32_t width;     // free width (0 constant)
    int32_t tsize;     // tangent size 6/9/6/1
    int32_t chart;     // chart offset in A's columns
    int32_t ref;       // offset into the reference values
    int32_t pad;
};

// Device code sees the buffers of a window descriptor in the global address space: loads and
// stores through them compile to global memory instructions, which the compiler can schedule
// around LDS accesses (a generic pointer might address LDS). The host sees plain pointers; the
// layout is the same. Host-only translation units define VIB_HOST_VIEW (their device pass then
// sees the host's types).
#if defined(__HIP_DEVICE_COMPILE__) && !defined(VIB_HOST_VIEW)
#define VIB_G __attribute__((address_space(1)))
#else
#define VIB_G
#endif
// The edit kernels fill buffers that the solve kernels only read.
template <class T>
__host__ __device__ inline VIB_G T* writable(VIB_G const T* p) { return const_cast<VIB_G T*>(p); }

struct WinDev {
    int32_t K, C, L, F, E;
    int32_t nv, nr, ne;
    int32_t rig_col, td_col;
    uint32_t flags;
    int32_t tri;               // nv*(nv+1)/2
    double huber, min_rho, td_lo, td_hi;
    int32_t pose_col[kMaxFrames], sb_col[kMaxFrames], sb_w[kMaxFrames];
    // Reduced-system factorization order: the speed/bias columns first, then the visual columns
    // (a rotation of the tangent layout: column c < nv goes to c + nr - nv, c >= nv to c - nv).
    // chol_mask[k] bit i: tile (i, k) of the factor in that order may be nonzero (i >= k).
    uint32_t chol_mask[kMaxTiles];
    // state and trial candidate
    VIB_G double* pose; VIB_G double* sb; VIB_G double* rig; VIB_G double* td; VIB_G double* rho;
    VIB_G double* cpose; VIB_G double* csb; VIB_G double* crig; VIB_G double* ctd; VIB_G double* crho;
    VIB_G const double* cam;         // C x 12: R (9, row-major R_RC) then t (3)
    // landmarks
    VIB_G const int32_t* lm_fc;      // anchor frame << 8 | camera
    VIB_G const int32_t* lm_off;     // L + 1 observation offsets
    VIB_G const int32_t* lm_col;     // rho column or -1
    VIB_G const double* lmd;         // LM_PLANES x L
    // observations (sorted by landmark)
    VIB_G const int32_t* ob_fc;      // observer frame << 8 | camera
    VIB_G const double* ob;          // OB_PLANES x F
    // observations again in (anchor frame, observer frame) pair order, for the visual normal
    // equations (runs of one pair reduce on MFMA)
    VIB_G const double* pf_ob;       // OB_PLANES x F
    VIB_G const double* pf_an;       // LM_PLANES x F : the factor's anchor bearing, velocity, t_ref
    VIB_G const int32_t* pf_ids;     // F : anchor frame << 24 | anchor cam << 16 | observer frame << 8 | cam
    VIB_G const int32_t* pf_lm;      // F : landmark
    VIB_G const int32_t* pf_pos;     // F : pair-order position of each landmark-order factor (the per-factor
                               //     coupling planes fac_u / fac_lm are stored in pair order)
    // linearization outputs
    VIB_G double* fac_u;             // 6 x F : J_anchor^T J_rho per factor (weighted, landmark order)
    VIB_G double* fac_lm;            // FL_PLANES x F : per-factor landmark terms (landmark order)
    VIB_G double* lm_out;            // LO_PLANES x L
    VIB_G double* lm_delta;          // L (tangent step of rho; 0 for constant)
    VIB_G int64_t* lm_uid;           // L stable landmark ids (strictly increasing; kept through edits)
    // IMU
    VIB_G const int32_t* imu_from; VIB_G const int32_t* imu_to; VIB_G const double* imu;
    // prior
    int32_t prior_rows, prior_cols, prior_blocks, pad0;
    VIB_G const PriorBlockDev* pblk;
    VIB_G const double* pA;          // rows x cols
    VIB_G const double* pr0;
    VIB_G const double* pG;          // cols x cols = A^T A
    VIB_G const double* pref;
    VIB_G const int32_t* pfree;      // free prior columns: block << 8 | index within the block
    int32_t prior_free, pad2;
    // retained-factor linearization (written by k_retained, assembled row-parallel)
    VIB_G double* imu_J;             // E x 15 x 30 whitened Jacobians [pose0 | sb0 | pose1 | sb1]
    VIB_G double* imu_r;             // E x 15 whitened residuals
    VIB_G double* pD;                // prior_blocks x 36 chart derivatives (6 x 6, row-major)
    VIB_G double* patr;              // prior_cols : A^T r
    VIB_G double* ret_sc;            // 4 E + 4 : IMU costs (lin, trial), invalid (lin, trial); prior same
    VIB_G const int32_t* colmap;     // nr x 4 : IMU (edge << 8 | local) x2, prior (block << 8 | index), -1
    // dense per-window buffers
    VIB_G const double* hostH;       // nr x nr normal matrix of host-evaluated factors, or null
    VIB_G const double* hostG;       // nr gradient of host-evaluated factors, or null
    VIB_G double* Hfull;             // nr x nr undamped full normal matrix of the states
    VIB_G double* gfull;             // nr
    VIB_G double* fac;               // factor_doubles(nr): tile store of the damped reduced system,
                                // overwritten by its lower Cholesky factor
    VIB_G double* Scopy;             // nr x nr damped reduced system (row-major), kept for refinement
    VIB_G double* rhs;               // nr
    VIB_G double* dstate;            // nr state step
    VIB_G double* scratch;           // 4 * nr
    // partials
    VIB_G double* lin_part;          // nlin x (tri + nv) (one per factor range of the visual kernel)
    VIB_G double* lin_sc;            // nlin x CS_PLANES
    VIB_G double* schur_part;        // nchunks x (tri + nv)
    VIB_G double* chunk_sc;          // nchunks x CS_PLANES
    int32_t nlin, nchunks;
    VIB_G LmState* lm;               // persistent per-window LM state (device)
};

struct Chunk {
    int32_t win;        // window index in the batch
    int32_t local;      // chunk index within the window
    int32_t lm_begin, lm_end;  // landmark range (factor range for the visual-normal work list)
};

constexpr int kLinFactors = 512;  // factors per visual-normal workgroup (pair order)

// One symmetric eigendecomposition of a batched k_sym_eigen launch (one workgroup per job).
struct EigenJob {
    VIB_G double* A;     // n x n (leading dimension ld), overwritten
    VIB_G double* V;     // eigenvectors as rows
    VIB_G double* ev;    // eigenvalues (unsorted)
    VIB_G double* rot;   // rotation log (eigen_log_doubles(n))
    VIB_G int32_t* ok;   // 1: converged
    int32_t n, ld;
};

// One copy of a batched transfer (k_copy_segments): `bytes` from src to dst, both device
// addresses; src 0 fills dst with zeros.
struct CopySeg {
    uint64_t dst, src, bytes;
};

// One window edit (vib_edit.hip): which landmarks and observations of a resident window the
// edited window keeps, and how its frames are renumbered.
struct EditSpec {
    int32_t renum[kMaxFrames];  // new slot of each source frame; -1: the frame is removed
    int32_t lm_mode;            // inverse ranges of kept landmarks: 0 as before, 1 all free, 2 all constant
    int32_t nr;                 // retained dimension of the edited window (free rho columns follow it)
    int32_t keep_lm_invert;     // 1: keep_lm marks the landmarks removed instead
    int32_t pad;
    const uint8_t* keep_lm;     // L keep flags (null: every landmark)
    const uint8_t* keep_f;      // F keep flags in landmark order (null: every observation of a kept landmark)
    const int32_t* over_f;      // L (null: none): >= 0 re-anchors the landmark in that observation
    const double* over_rho;     // L: the inverse range in the new anchor
    // inserted landmarks (after the kept ones) and observations (an appended keyframe's), all in
    // the edited window's frame numbering
    int32_t ins_lm, ins_obs;
    const int32_t* add_cnt;      // L + ins_lm: inserted observations per landmark (null: none)
    const int32_t* ins_lm_fc;    // ins_lm: anchor frame << 8 | camera
    const double* ins_lmd;       // LM_PLANES x ins_lm
    const double* ins_rho;       // ins_lm
    const uint8_t* ins_free;     // ins_lm
    const int64_t* ins_uid;      // ins_lm
    const int32_t* ins_obs_lm;   // ins_obs: source landmark, or L + j for inserted landmark j; -1 dropped
    const int32_t* ins_obs_rank; // ins_obs: position among the inserted observations of its landmark
    const int32_t* ins_obs_fc;   // ins_obs: observer frame << 8 | camera
    const double* ins_ob;        // OB_PLANES x ins_obs
};

}  // namespace vib
