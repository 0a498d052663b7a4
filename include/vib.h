/*
 * vib.h -- C-ABI of the MI355X sliding-window visual-inertial Schur backend.
 *
 * This library is a drop-in for the `backend_visual_batch: schur` estimator option of
 * CGMLI-SLAM (config/hilti_robot_estimator.yaml:224-225, parsed at
 * src/common/ViParametersReader.cpp:354-359, mapped at src/pipeline/ThreadedSlam.cpp:33-38).
 * It replaces the device layer `cgmli::backend::detail::VisualCudaBatch` /
 * `MarginalCudaWorkspace` (src/backend/VisualCudaBatch.hpp:43-104) together with the host
 * drivers that call it: the native LM loop `solveVisualSchur`
 * (src/backend/BackendOptimizer.cpp:27-331), the two-frame solve `optimizeLatestTwo`
 * (src/backend/BackendOptimizer.cpp:476-590) and the marginalizer
 * (src/backend/Marginalizer.cpp:284-474).
 *
 * Conventions (identical to the reference, src/backend/README.md:45-54,111):
 *   pose        7 doubles  [p_x p_y p_z q_x q_y q_z q_w]  (T_WB, Eigen xyzw quaternion)
 *   speed/bias  9 doubles  [v_W(3) b_g(3) b_a(3)]
 *   tangent     pose [rho(3) theta(3)] left perturbation; everything else additive
 *   rig         T_BR, one shared extrinsic; fixed per-camera T_RC
 *   landmark    inverse range rho = 1/|p_C| along the anchor bearing, rho >= 1e-8
 *   td          camera-IMU time offset, t_IMU = t_sensor + td
 *
 * Every function returns an int status (VIB_OK = 0). A non-zero status corresponds to
 * one of the reference's fallback reasons (src/backend/BackendOptimizer.cpp:32-308);
 * vib_last_error() gives the message. No function throws.
 *
 * Threading: every entry point takes the context's lock, so threads may share a context; the
 * request server (vib_server_*) is a client of these entry points driven by one thread, and runs
 * its workers' contexts on threads of its own.
 */
#ifndef VIB_H
#define VIB_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VIB_ABI_VERSION 3

/* ---- status codes (reference fallback classes, src/backend/BackendOptimizer.cpp) ---- */
enum {
    VIB_OK = 0,
    VIB_ERR_INVALID_ARGUMENT = 1,      /* malformed window / options                       */
    VIB_ERR_UNSUPPORTED_DIMENSIONS = 2,/* "unsupported Schur dimensions"  (:67-68)          */
    VIB_ERR_INVALID_RESIDUAL = 3,      /* invalid factor (rho<=0, |p_C|<1e-10, ...) (:231)   */
    VIB_ERR_NOT_POSITIVE_DEFINITE = 4, /* damped RCS Cholesky failed (potrf info != 0)       */
    VIB_ERR_COST_MISMATCH = 5,         /* "joint cost disagrees with accepted trial" (:236)  */
    VIB_ERR_NO_ACCEPTABLE_STEP = 6,    /* "no acceptable CUDA Schur trial" (:308)            */
    VIB_ERR_UNSUPPORTED_LOSS = 7,      /* "unsupported visual loss" (:112)                   */
    VIB_ERR_DEVICE = 8,                /* HIP runtime failure                               */
    VIB_ERR_OUT_OF_MEMORY = 9,         /* HBM budget / allocation failure                   */
    VIB_ERR_NONFINITE_STEP = 10,       /* "nonfinite CUDA Schur step" (:254)                 */
    VIB_ERR_INDEFINITE = 11,           /* "Schur prior is indefinite" (Marginalizer.cpp:444) */
    VIB_ERR_EIGEN_FAILED = 12,         /* eigensolver did not converge                      */
    VIB_ERR_NO_DEVICE = 13,            /* no HIP device                                     */
    VIB_ERR_NOT_FOUND = 14,            /* unknown slot / session                            */
    VIB_ERR_UNSUPPORTED = 15,          /* unsupported configuration                          */
    VIB_ERR_LIGHTWEIGHT_SUPPORT = 16,  /* "missing pose or velocity block" / "fewer than six
                                          current landmarks in two-frame solve" (:500-514)   */
    VIB_ERR_IMU_BRIDGE = 17,           /* "two-frame IMU bridge missing" / "... invalid before
                                          / after solve" (:528-567)                          */
    VIB_ERR_IMU_GATE = 18              /* "new IMU edge residual exceeded lightweight health
                                          gate" (NIS > 60, :572-577)                         */
};

/* window flags */
#define VIB_RIG_FREE 1u   /* estimate the shared T_BR online */
#define VIB_TD_FREE 2u    /* estimate the camera-IMU time offset online */

/* prior key kinds (reference BlockKind order, src/backend/BackendTypes.hpp:30-31) */
#define VIB_KEY_POSE 0
#define VIB_KEY_SPEED_BIAS 1
#define VIB_KEY_RIG 2
#define VIB_KEY_TD 3

/* precision of a solve */
#define VIB_FP64 0   /* plain FP64 everywhere */
#define VIB_MIXED 1  /* Hessian-side coupling streams as FP32 (hi, lo) pairs; the landmark Schur
                        products of the reduced system from the hi parts (FP32 MFMA, FP64 sums
                        over 16-landmark groups); factor evaluation, gradients, costs, the
                        reduced-system Cholesky and the back-substitution stay FP64. With
                        refinement_steps > 0 the state step is refined against the exact FP64
                        reduced operator (matrix-free landmark pass over hi + lo), which brings
                        it to the FP64 step; with 0 it keeps the FP32 products' error. */

/* marginalization policies (src/backend/SlidingWindowManager.cpp:503) */
#define VIB_MARGIN_OLD 0
#define VIB_MARGIN_SECOND_NEW 1

/* Frozen IMU preintegration record, 301 doubles (reference FrozenImuData,
 * src/backend/GpuRetainedFactors.hpp:7-14), every matrix row-major:
 *   [0] duration  [1] gravity  [2..4] p_BS  [5..8] q_BS(xyzw)  [9..14] gyro at start/end
 *   [15..18] dq(xyzw) [19..21] dv  [22..24] dp  [25..30] reference bias (b_g, b_a)
 *   [31..39] d(alpha)/db_g  [40..48] dv/db_g  [49..57] dp/db_g  [58..66] dv/db_a (C integral)
 *   [67..75] dp/db_a (C double integral)  [76..300] whitening W (15x15)                       */
#define VIB_IMU_STRIDE 301

/* Raw IMU sample: time [s], gyro (3) [rad/s], accel (3) [m/s^2]. */
#define VIB_IMU_SAMPLE_STRIDE 7

/* IMU noise model (reference ImuParameters, config/hilti_robot_estimator.yaml:120-140). */
typedef struct vib_imu_params {
    double sigma_g_c, sigma_a_c, sigma_gw_c, sigma_aw_c;
    double g;                 /* gravity magnitude (9.81) */
    double g_max, a_max;      /* saturation thresholds */
    double p_BS[3];
    double q_BS[4];           /* xyzw */
} vib_imu_params;

/* One sliding window, described by host pointers (copied on upload). Observations are the
 * non-anchor visual factors (one per (landmark, observing camera) pair other than the anchor).
 * Arrays marked (in/out) receive the solved state in vib_read_state(). */
typedef struct vib_window {
    int32_t num_frames;        /* K (window slots, oldest first)                      */
    int32_t num_cameras;       /* C (fixed T_RC per camera)                           */
    int32_t num_landmarks;     /* L                                                   */
    int32_t num_observations;  /* F                                                   */
    int32_t num_imu;           /* E (IMU edges)                                       */
    uint32_t flags;            /* VIB_RIG_FREE | VIB_TD_FREE                          */
    double huber_delta;        /* whitened Huber scale; 0 disables the robust loss    */
    double min_inverse_depth;  /* lower bound of rho (reference default 1e-8)         */
    double td_lower, td_upper; /* bounds of td when free                              */
    /* frames */
    const int64_t* frame_id;   /* [K] stable ids; prior keys refer to them            */
    double* pose;              /* [K*7] (in/out)                                      */
    double* speed_bias;        /* [K*9] (in/out)                                      */
    const uint8_t* pose_free;  /* [K] 1 free / 0 constant; NULL = all free            */
    const uint8_t* sb_width;   /* [K] 9 full, 3 velocity only, 0 constant; NULL = 9   */
    /* calibration */
    double* rig;               /* [7] T_BR (in/out)                                   */
    double* td;                /* [1] (in/out)                                        */
    const double* cam_rotation;    /* [C*9] R_RC row-major                            */
    const double* cam_translation; /* [C*3] t_RC                                      */
    /* landmarks (anchor observation + inverse range) */
    const int32_t* lm_anchor_frame; /* [L] window slot of the anchor                   */
    const int32_t* lm_anchor_cam;   /* [L]                                            */
    const double* lm_bearing;       /* [L*3] raw anchor bearing                       */
    const double* lm_velocity;      /* [L*3] anchor bearing velocity                  */
    const double* lm_tref;          /* [L] anchor reference time offset               */
    double* lm_rho;                 /* [L] (in/out)                                   */
    const uint8_t* lm_free;         /* [L] NULL = all free                            */
    /* visual factors */
    const int32_t* obs_landmark;    /* [F]                                            */
    const int32_t* obs_frame;       /* [F] window slot of the observer                */
    const int32_t* obs_cam;         /* [F]                                            */
    const double* obs_bearing;      /* [F*3] raw measured bearing                     */
    const double* obs_velocity;     /* [F*3] measured bearing velocity                */
    const double* obs_tref;         /* [F] reference time offset                      */
    const double* obs_sqrt_info;    /* [F*3] (s00, s10, s11) of the lower-triangular 2x2
                                       square-root information in the sphere basis      */
    /* IMU factors */
    const int32_t* imu_from;        /* [E] window slot of the first state             */
    const int32_t* imu_to;          /* [E] window slot of the second state            */
    const double* imu_data;         /* [E*VIB_IMU_STRIDE]                             */
    /* square-root prior r = r0 + A * (x [-] x_ref) (MarginalizationPrior.cpp:20-66) */
    int32_t prior_rows;             /* m (0 = no prior)                               */
    int32_t prior_blocks;           /* number of keys                                 */
    const int32_t* prior_key_kind;  /* [B] VIB_KEY_*                                  */
    const int64_t* prior_key_id;    /* [B] frame id (POSE/SPEED_BIAS), 0 otherwise    */
    const double* prior_ref;        /* concatenated reference values (7/9/7/1 per key) */
    const double* prior_A;          /* [m * n] row-major, n = sum of tangent sizes     */
    const double* prior_r0;         /* [m]                                            */
    /* landmark ids (the reference's LandmarkId): vib_window_append's observations name landmarks
       by id, and the ids follow the landmarks through slides */
    const int64_t* lm_id;           /* [L] strictly increasing; NULL = 0 .. L-1         */
} vib_window;

/* A new keyframe for vib_window_append: its state, the IMU edge from the window's newest frame,
 * the landmarks it anchors and the observations it makes of resident or new landmarks. */
typedef struct vib_keyframe {
    int64_t frame_id;
    const double* pose;             /* [7] initial estimate                           */
    const double* speed_bias;       /* [9]                                            */
    int32_t pose_free;              /* 1 free / 0 constant                            */
    int32_t sb_width;               /* 9, 3 or 0                                      */
    const double* imu_data;         /* [VIB_IMU_STRIDE] edge newest -> this frame; NULL: none */
    int32_t num_landmarks;          /* N landmarks anchored in this frame              */
    int32_t num_observations;       /* M observations made from this frame             */
    const int64_t* lm_id;           /* [N] strictly increasing, above every id the window held */
    const int32_t* lm_anchor_cam;   /* [N]                                            */
    const double* lm_bearing;       /* [N*3]                                          */
    const double* lm_velocity;      /* [N*3]                                          */
    const double* lm_tref;          /* [N]                                            */
    const double* lm_rho;           /* [N]                                            */
    const uint8_t* lm_free;         /* [N] NULL = all free                            */
    const int64_t* obs_landmark_id; /* [M] id of the observed landmark                 */
    const int32_t* obs_cam;         /* [M]                                            */
    const double* obs_bearing;      /* [M*3]                                          */
    const double* obs_velocity;     /* [M*3]                                          */
    const double* obs_tref;         /* [M]                                            */
    const double* obs_sqrt_info;    /* [M*3]                                          */
} vib_keyframe;

/* solve option flags */
#define VIB_SOLVE_CONTINUOUS 1u  /* relinearize every iteration and run the full iteration
                                    budget without terminating on convergence (benchmarks,
                                    serving load)                                          */

typedef struct vib_solve_options {
    int32_t max_iterations;     /* LM iteration budget (reference: maxIterations)       */
    int32_t min_iterations;
    double initial_lambda;      /* 1e-4 in the reference (:167)                          */
    double time_limit_seconds;  /* <= 0: unlimited                                       */
    int32_t precision;          /* VIB_FP64 / VIB_MIXED                                  */
    int32_t refinement_steps;   /* iterative-refinement steps of the RCS solve (VIB_MIXED:
                                   against the exact operator; 2 reach the FP64 step)    */
    uint32_t flags;             /* VIB_SOLVE_*                                           */
} vib_solve_options;

typedef struct vib_step_report {    /* reference VisualSchurStep (VisualCudaBatch.hpp:38-42) */
    double cost;                /* robust cost at the linearization point              */
    double predicted_reduction; /* model decrease of the clamped step                  */
    double trial_cost;          /* cost at x [+] delta                                 */
    double max_increment;       /* max |delta|                                         */
    int32_t trial_valid;
    int32_t status;
} vib_step_report;

typedef struct vib_solve_report {
    int32_t status;
    int32_t iterations;
    int32_t accepted_steps;
    int32_t converged;
    int32_t effective_parameters; /* retained + eliminated tangent dimension */
    int32_t retained_dimension;   /* nr */
    double initial_cost;
    double final_cost;
    double final_lambda;
    double seconds;               /* wall time of the solve on the host */
} vib_solve_report;

/* IMU handling of a SECOND_NEW slide ([Window] imu_mode, SlidingWindowManager.cpp:275-311) */
#define VIB_IMU_MERGE_NONE 0          /* no IMU edge touched the frame                         */
#define VIB_IMU_MERGE_REINTEGRATED 1  /* merged edge re-integrated at the current bias          */
#define VIB_IMU_MERGE_INCREMENTAL 2   /* merged edge kept the left edge's reference bias        */
#define VIB_IMU_MERGE_SUPPRESSED 3    /* merged interval longer than 10 s: no IMU factor left   */
#define VIB_IMU_MERGE_BRIDGE 4        /* no replay (or its SB in the prior): edges go to Schur  */

typedef struct vib_marginal_report {
    int32_t status;
    int32_t prior_reused;        /* SECOND_NEW with an untouched prior                   */
    int32_t eliminated_dimension;
    int32_t retained_dimension;
    int32_t prior_rank;
    int32_t eliminated_landmarks;
    int32_t consumed_factors;
    double eigen_cutoff;
    int32_t imu_mode;            /* VIB_IMU_MERGE_* (SECOND_NEW)                          */
    int32_t reset_landmarks;     /* re-anchored on a surviving view (reset_tracks)        */
    int32_t demoted_landmarks;   /* one surviving ray: left the window (demoted_tracks)   */
    int32_t dropped_factors;     /* the frame's image measurements discarded              */
} vib_marginal_report;

typedef struct vib_latest_two_report {  /* SolveReport lightweight fields (:541-562) */
    vib_solve_report solve;
    double old_imu_nis_before;   /* squared whitened residual of the edge ending at K-2   */
    double old_imu_nis_after;
    double new_imu_nis_before;   /* edge K-2 -> K-1                                        */
    double new_imu_nis_after;
    int32_t support_landmarks;   /* distinct landmarks observed in the latest frame       */
} vib_latest_two_report;

typedef struct vib_context vib_context;

/* ---------------- context ---------------- */
int vib_abi_version(void);
int vib_create(vib_context** ctx, int device);
/* The device a context was created on. */
int vib_context_device(const vib_context* ctx, int32_t* device);
int vib_destroy(vib_context* ctx);
/* Wait for all work of the context's device (hipDeviceSynchronize). */
int vib_device_synchronize(vib_context* ctx);
/* Last error message of the context (reference takeVisualCudaRuntimeError, VisualCudaBatch.cu:399-402). */
const char* vib_last_error(const vib_context* ctx);

/* ---------------- window slots (device-resident sessions) ---------------- */
/* Upload a window into slot `slot` (>=0); replaces any previous content. Topology, measurement
 * streams, IMU and prior are copied to HBM (reference linearizeSchur(newTopology) +
 * configureRetained + configureResident, VisualCudaBatch.cu:433-511,606-624,673-691). */
int vib_window_upload(vib_context* ctx, int32_t slot, const vib_window* window);
/* Errors of an upload: VIB_ERR_INVALID_ARGUMENT (malformed window), VIB_ERR_UNSUPPORTED_DIMENSIONS
 * (more than 20 frames or 4 cameras: the reference's "unsupported Schur dimensions",
 * BackendOptimizer.cpp:67-68, at this library's limits), VIB_ERR_UNSUPPORTED_LOSS (Huber scale
 * negative or not finite; BackendOptimizer.cpp:112), VIB_ERR_OUT_OF_MEMORY (allocation, or the
 * HBM budget of vib_set_hbm_budget), VIB_ERR_DEVICE (copies). A failed upload leaves the slot
 * empty (VIB_ERR_NOT_FOUND afterwards). A landmark may have any number of factors. */
/* HBM admission (no counterpart in the reference; for callers that serve many sessions): an
 * upload fails with VIB_ERR_OUT_OF_MEMORY when the resident windows' footprint plus the new one
 * would exceed `bytes` (0: no budget). vib_resident_bytes reports the current footprint. */
int vib_set_hbm_budget(vib_context* ctx, int64_t bytes);
int vib_resident_bytes(vib_context* ctx, int64_t* bytes);
/* Footprint of one resident slot's window (part of vib_resident_bytes). Budgets count these
 * footprints (vib_window_footprint of the resident windows); a slot's arena keeps up to a
 * quarter more, so that the window's appends are edited in place. */
int vib_slot_bytes(vib_context* ctx, int32_t slot, int64_t* bytes);
/* Bytes of HBM a window of this shape occupies (admission accounting). */
int64_t vib_window_footprint(const vib_window* window, int32_t precision);
int vib_window_release(vib_context* ctx, int32_t slot);
/* Append a keyframe as the window's newest frame (FixedLagBackend::append,
 * src/backend/FixedLagBackend.cpp:17-22, with the topology change that the next solve uploads,
 * BackendOptimizer.cpp:169-209). A device edit inserts its landmarks and observations into the
 * resident streams in the layout of an upload; nothing of the window is re-uploaded. Observations
 * of ids no longer in the window (landmarks marginalized meanwhile) are skipped and counted in
 * *unmatched (may be NULL). Errors: VIB_ERR_UNSUPPORTED_DIMENSIONS (the window already holds 20
 * frames), VIB_ERR_INVALID_ARGUMENT (missing arrays, ids not increasing, camera out of range),
 * VIB_ERR_OUT_OF_MEMORY (HBM budget); a failed append leaves the window as it was. */
int vib_window_append(vib_context* ctx, int32_t slot, const vib_keyframe* keyframe, int32_t* unmatched);
/* Copy the resident state back into the (in/out) arrays of `window` (readResident, :704-713). */
int vib_read_state(vib_context* ctx, int32_t slot, vib_window* window);
/* Shape of the resident window after marginalizations. */
int vib_window_shape(vib_context* ctx, int32_t slot, int32_t* frames, int32_t* landmarks,
                     int32_t* observations, int32_t* imu, int32_t* prior_rows, int32_t* prior_cols);

/* ---------------- one LM iteration, exposed piecewise (VisualCudaBatch API) ---------------- */
/* Linearize visual + IMU + prior factors at the resident state (linearizeResident, :693-700). */
int vib_linearize(vib_context* ctx, int32_t slot, int32_t precision);
/* Damped Schur step, back-substitution, resident trial and trial cost (solveOnline, :530-604).
 * Returns VIB_ERR_NONFINITE_STEP when an entry of the step (vib_read_step) is not finite. */
int vib_solve_step(vib_context* ctx, int32_t slot, double lambda, int32_t precision,
                   int32_t refinement_steps, vib_step_report* report);
/* Normal equations of factors the caller evaluates on the host (the retainedH / retainedG
 * arguments of solveOnline, VisualCudaBatch.hpp:61-62; BackendOptimizer.cpp:210-228): H is the
 * full nr x nr matrix (row-major, the slot's tangent column layout, vib_retained_dimension) and g
 * the nr gradient of those factors at the current state. They are added to the window's own
 * terms at every following vib_linearize (damping then uses diag(H + H_host), as the reference)
 * until replaced, cleared (H = g = NULL) or the slot is re-uploaded. As in the reference, whose
 * solve leaves the device-resident mode when host factors exist, such a window takes the
 * piecewise API: the caller adds its factors' cost to the reported costs, and vib_solve,
 * vib_solve_batch and the request API return VIB_ERR_UNSUPPORTED for it. */
int vib_set_host_retained(vib_context* ctx, int32_t slot, const double* H, const double* g);
/* Adopt the last trial state (acceptResident, :702). */
int vib_accept(vib_context* ctx, int32_t slot);
/* Debug/parity read-outs of the last step: full tangent step [states, rho] (length nr + L)
 * and the damped reduced camera system (nr*nr row-major, upper triangle meaningful) + rhs. */
int vib_read_step(vib_context* ctx, int32_t slot, double* delta, int32_t capacity);
int vib_read_reduced(vib_context* ctx, int32_t slot, double* S, double* rhs, int32_t capacity);
/* Undamped normal equations of the retained block (readNormal, :651-671): H (nr*nr) and g. */
int vib_read_normal(vib_context* ctx, int32_t slot, double* H, double* g, double* cost, int32_t capacity);
int vib_retained_dimension(vib_context* ctx, int32_t slot, int32_t* nr, int32_t* nv);

/* ---------------- full solves ---------------- */
/* Native LM (solveVisualSchur, BackendOptimizer.cpp:27-331). */
int vib_solve(vib_context* ctx, int32_t slot, const vib_solve_options* options, vib_solve_report* report);
/* Many independent windows in lock-step batched launches (one graph per LM iteration). */
int vib_solve_batch(vib_context* ctx, int32_t count, const int32_t* slots,
                    const vib_solve_options* options, vib_solve_report* reports);
/* Two-frame lightweight solve (optimizeLatestTwo, BackendOptimizer.cpp:476-590, scoped
 * pose/velocity manifold :594,783-790): free blocks are the poses and velocities of slots K-2
 * and K-1; landmarks, biases, rig, td and the other frames are constant; only factors touching
 * a free block take part; >= 6 distinct landmarks observed in frame K-1 and exactly the two IMU
 * edges ending at K-2 and K-2 -> K-1 are required; at most 3 iterations and 8 ms. With
 * gate_new_imu_nis the result is rejected when the new edge's NIS exceeds 60. On success the
 * poses and velocities of K-2 and K-1 in the resident window are updated. */
int vib_solve_latest_two(vib_context* ctx, int32_t slot, const vib_solve_options* options,
                         int32_t gate_new_imu_nis, vib_latest_two_report* report);

/* ---------------- serving: continuous batching (no counterpart in the reference) ----------
 * A request is a short LM solve on a resident window. vib_request_begin starts it (fresh LM
 * state with the options' iteration budget and initial lambda); vib_request_iterate then runs
 * one LM iteration for every listed slot in one batched launch sequence. done[i] = 1 once slot
 * i's request has finished (converged, budget spent or failed; reports[i] then holds its
 * result, status included). The set of slots may change between calls, so requests that arrive
 * while others are running join at the next iteration: each window's result does not depend on
 * which other windows share its launches. All slots of one call use one precision. */
int vib_request_begin(vib_context* ctx, int32_t slot, const vib_solve_options* options);
int vib_request_iterate(vib_context* ctx, int32_t count, const int32_t* slots, int32_t precision,
                        int32_t refinement_steps, int32_t* done, vib_solve_report* reports);

/* The request server: many sessions, each a sliding window whose requests bring the next
 * keyframe. Serving a request is the reference's per-frame cycle (FixedLagBackend::append, then
 * step: the solve, then the slide once the window is over its size;
 * src/backend/FixedLagBackend.cpp:17-22, 405-451): vib_window_append, LM iterations batched
 * with the other running requests of its size bucket, and vib_marginalize back to the
 * session's window size, the prior staying resident. Every session belongs to one worker (the
 * least loaded by window footprint when it is added); each worker runs passes on a thread of its
 * own: (1) admit the first queued request of every idle session of the worker whose window is
 * resident or fits the HBM budget (idle sessions with nothing queued are evicted to host memory,
 * least recently used first), (2) start them (append, LM state), (3) run one LM iteration for its
 * running requests of one size bucket (the buckets take turns), and (4) slide the requests that
 * finished. Requests submitted while others run join at the next pass. A session whose request
 * finished waits until vib_server_turn has handed the result out and has been called again, so
 * between two turn calls the caller may read the windows of the sessions just reported. The
 * server uses context slots from first_slot on; one thread drives it (submit, turn). */
typedef struct vib_server vib_server;
typedef struct vib_server_options {
    int64_t hbm_budget;          /* bytes of resident windows the server admits (0: no limit); a
                                    request whose window alone exceeds it fails with
                                    VIB_ERR_OUT_OF_MEMORY                                      */
    int32_t precision;           /* VIB_FP64 or VIB_MIXED                                       */
    int32_t refinement_steps;    /* refinement passes of a mixed-precision step                */
    int32_t slide_policy;        /* VIB_MARGIN_OLD or VIB_MARGIN_SECOND_NEW after a request that
                                    grew the window; -1: no slide                             */
    int32_t full_iterations;     /* 1: every request runs its whole budget (VIB_SOLVE_CONTINUOUS) */
    int32_t first_slot;          /* context slots first_slot, first_slot + 1, ... are the server's */
    int32_t num_buckets;         /* size buckets by retained dimension (<= 8; 0: one bucket)   */
    int32_t bucket_limit[8];     /* largest retained dimension of each bucket (ascending)      */
    int32_t workers;             /* contexts serving disjoint sessions concurrently, each on a
                                    thread and stream of its own (<= VIB_SERVER_MAX_WORKERS;
                                    0 or 1: the caller's context alone). Worker 0 is the
                                    caller's context, the others are created on its device;
                                    hbm_budget bounds the windows resident in all of them.   */
} vib_server_options;
#define VIB_SERVER_MAX_WORKERS 16
typedef struct vib_served {      /* one finished request */
    int64_t request;
    int32_t session;
    int32_t status;              /* VIB_OK or the first error of its append, solve or slide   */
    int32_t iterations;          /* LM iterations run                                          */
    int32_t unmatched;           /* keyframe observations of landmarks no longer resident      */
    double final_cost;
    double queued_seconds;       /* submission to admission                                   */
    double latency_seconds;      /* submission to completion, slide included                  */
} vib_served;
typedef struct vib_server_stats {
    int64_t submitted, finished, failed, turns, launches, window_iterations;
    int64_t uploads, evictions, appends, slides, slide_launches, resident_bytes;
    int32_t queued, running, resident, pending_results;
    /* host wall time spent in each phase */
    double admit_seconds;        /* uploads, evictions, appends, request starts               */
    double iterate_seconds;      /* batched LM iterations                                    */
    double slide_seconds;        /* slides                                                   */
} vib_server_stats;
int vib_server_create(vib_context* ctx, const vib_server_options* options, vib_server** server);
/* Releases the server's resident windows. */
int vib_server_destroy(vib_server* server);
/* A session and its first window (copied; uploaded when the session's first request is admitted).
 * Its frame count is the size the session's window slides back to. */
int vib_server_add_session(vib_server* server, int32_t session, const vib_window* window);
/* Queues a request of `session`: append `keyframe` (copied; NULL: none), then at most
 * `iterations` LM iterations, then the slide. Requests of one session are served in order. */
int vib_server_submit(vib_server* server, int64_t request, int32_t session, int32_t iterations,
                      const vib_keyframe* keyframe);
/* Collects finished requests: up to `capacity` are copied to `finished` (the rest are kept for the
 * next call), their number in *count; waits up to a millisecond when none has finished. */
int vib_server_turn(vib_server* server, vib_served* finished, int32_t capacity, int32_t* count);
/* Makes the sessions' windows resident in session order until the HBM budget is full. */
int vib_server_preload(vib_server* server);
int vib_server_get_stats(const vib_server* server, vib_server_stats* stats);
/* The context slot holding the session's window, -1 while it is evicted. */
int vib_server_session_slot(const vib_server* server, int32_t session, int32_t* slot);
/* The same, with the worker context the slot belongs to (the caller's context with one worker). */
int vib_server_session_context(const vib_server* server, int32_t session, vib_context** ctx, int32_t* slot);

/* ---------------- measurement harness (no counterpart in the reference) ---------------- */
/* Number of distinct kernels of one LM iteration and their names (for per-kernel timing). */
int vib_kernel_count(void);
const char* vib_kernel_name(int32_t kernel);
/* Runs `iterations` continuous-mode LM iterations on the given slots (one graph replay per
 * iteration, no host synchronization in between) and writes each iteration's device time,
 * measured with events on the library's stream, to iteration_ms. Then runs `iterations` more
 * iterations kernel by kernel with events around every launch and writes the average duration
 * of each kernel (vib_kernel_count entries) to kernel_ms. Either output may be NULL. For up to
 * 16 slots kernels 1 (the IMU edges and prior linearized) and 2 (the visual-block sums) report
 * 0: kernels 0 and 3 then run them as extra workgroups. Kernels 12 and 13 (the refinement
 * passes of a VIB_MIXED step) run refinement_steps times per iteration and report the sum of
 * their launches; 0 otherwise. */
int vib_bench(vib_context* ctx, int32_t count, const int32_t* slots, const vib_solve_options* options,
              int32_t iterations, double* iteration_ms, double* kernel_ms);
/* The work-list shapes of batched solves, chosen by measurement (no counterpart in the
 * reference): every combination of landmark chunks merged per workgroup (1, 2, 4, 8 chunks of
 * landmarks_per_chunk) and visual factor ranges merged per workgroup (1, 2, 4, 8 ranges of
 * factors_per_range) runs `iterations` LM iterations of the listed windows (captured graph,
 * continuous mode, timed with events); the fastest is kept by the context for batches of the
 * same precision and size class (log2 of their visual factor ranges) and used by vib_solve,
 * vib_solve_batch, the request API and vib_bench. The windows' states are restored afterwards.
 * Merged work lists change the order of some sums: a window's result in a tuned batch can
 * differ in the last bits from its result alone. */
typedef struct vib_tune_report {
    int32_t candidates;          /* combinations measured (16)                               */
    int32_t best;                /* index of the chosen one                                  */
    int32_t size_class;          /* the batch's size class                                   */
    int32_t landmarks_per_chunk; /* layout constants the merges multiply                     */
    int32_t factors_per_range;
    int32_t chunk_merge[16];
    int32_t range_merge[16];
    double ms_per_iteration[16];
} vib_tune_report;
int vib_autotune(vib_context* ctx, int32_t count, const int32_t* slots, const vib_solve_options* options,
                 int32_t iterations, vib_tune_report* report);
/* One launch of a plain streaming copy of `bytes` (multiple of 16) between two fresh device
 * buffers: the known-size reference that scales the FETCH_SIZE / WRITE_SIZE counters. */
int vib_copy_probe(vib_context* ctx, int64_t bytes);

/* ---------------- marginalization ---------------- */
/* OLD: Schur-eliminate slot 0 (pose, speed/bias) plus every landmark anchored in or observed
 * by it into a new square-root prior (Marginalizer::compute, Marginalizer.cpp:284-474).
 * SECOND_NEW, the redundant-frame path (SlidingWindowManager.cpp:263-482, 587-680): slot K-2's
 * two IMU edges become one edge K-3 -> K-1 re-integrated from the raw samples given here
 * (sample_count rows of VIB_IMU_SAMPLE_STRIDE covering [t_start, t_end], the times of slots K-3
 * and K-1; mergeBodyImuRecords, BodyImuFactor.cpp:306-388); without samples, or when K-2's
 * speed/bias is in the prior, the edges go into the Schur prior instead. Landmarks anchored in
 * K-2 whose bearings do not move are re-anchored on their newest surviving view (inverse range
 * from the projection of the current point, <= 0.03 rad off the ray), or leave the window with
 * one surviving ray; the others are Schur-eliminated with the slot. K-2's other image
 * measurements are discarded. When no factor is left on K-2 and the prior does not hold it, the
 * prior is reused unchanged. Either way the resident window slides by one frame. */
int vib_marginalize(vib_context* ctx, int32_t slot, int32_t policy,
                    const double* imu_samples, int32_t sample_count, double t_start, double t_end,
                    const vib_imu_params* imu_params, vib_marginal_report* report);
/* The slides of many windows at once (no counterpart in the reference; for the request server,
 * whose sessions slide together): with VIB_MARGIN_OLD the marginal systems linearize and reduce
 * in one batch of launches and their eigendecompositions run as one launch per stage (one
 * workgroup per window); each window's result equals vib_marginalize's. SECOND_NEW slides the
 * windows one by one without IMU samples. reports[i] holds window i's result; the return value
 * is the first failure. */
int vib_marginalize_batch(vib_context* ctx, int32_t count, const int32_t* slots, int32_t policy,
                          vib_marginal_report* reports);
/* Copy the whole resident window (topology, measurements, state and prior) into `window`, whose
 * sizes (num_* and prior_rows) must match vib_window_shape and whose arrays the caller
 * allocates at those sizes (prior keys: up to prior_cols entries, prior_ref: up to
 * 7 * prior_cols / 6 + 1 doubles); prior_blocks is written. Observations come in landmark
 * order. Lets the host mirror the window after a slide (no counterpart in the reference, whose
 * window manager keeps the graph on the host). */
int vib_read_window(vib_context* ctx, int32_t slot, vib_window* window);
/* Read the resident prior: keys, references, A (rows x cols row-major) and r0. */
int vib_read_prior(vib_context* ctx, int32_t slot, int32_t* rows, int32_t* cols, int32_t* blocks,
                   int32_t* key_kind, int64_t* key_id, double* ref, double* A, double* r0,
                   int32_t capacity_rows, int32_t capacity_cols, int32_t capacity_blocks);

/* ---------------- statistical gate and outlier refits ---------------- */
/* Residual statistics of one factor set (ResidualStatistics, BackendTypes.hpp:209-214), visual
 * category. NIS = squared whitened residual before the robust loss. worst_factor is a factor
 * index in landmark order (the order vib_read_window returns), -1 for an empty set. */
typedef struct vib_residual_stats {
    double factors, components, over_gate;
    double squared_norm, robust_cost;   /* robust cost: sum of the Huber 0.5 rho(NIS)       */
    double median_nis, p95_nis, max_nis; /* sorted[n/2], sorted[ceil(0.95 n) - 1], maximum    */
    int32_t worst_factor, pad;
} vib_residual_stats;

typedef struct vib_gate_report {
    int32_t status;        /* VIB_OK, or VIB_ERR_INVALID_RESIDUAL when a factor failed to
                              evaluate or was nonfinite (the reference rolls the solve back)   */
    int32_t rejected;      /* factors rejected by this call (0 unless applied)                */
    int32_t skipped;       /* factors rejected earlier (not participating)                    */
    int32_t invalid;       /* factors that failed to evaluate                                 */
    vib_residual_stats evaluated, admitted;
} vib_gate_report;

/* Statistical gate of the resident window at its current state
 * (BackendOptimizer::statisticalGate, BackendOptimizer.cpp:394-472): every participating visual
 * factor's NIS against `threshold` (the reference's visualNisThreshold 9.210340371976184,
 * BackendTypes.hpp:186). With apply != 0 and no invalid factor, the factors over the gate are
 * rejected for the rest of the window's life: their square-root information becomes zero on the
 * device and in the host mirror (FactorLifecycle::Rejected), and the next slide removes them
 * (window maintenance archives rejected factors before marginalization). decisions (nullable, F
 * bytes, landmark order): 0 admitted, 1 over the gate, 2 invalid, 3 not participating. */
int vib_nis_gate(vib_context* ctx, int32_t slot, double threshold, int32_t apply, vib_gate_report* report,
                 uint8_t* decisions);
/* vib_nis_gate without device allocation. No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under stream capture, tight serving loops,
 * harnesses that forbid hipMalloc). The same kernels and results as vib_nis_gate; the
 * per-call descriptors, energies, decisions and statistics live in one file-scope device array
 * of the library (about 11.5 MB per GPU, held until the process ends) instead of fresh
 * allocations. Windows beyond 2^20 factors or 2^17 landmark chunks return
 * VIB_ERR_UNSUPPORTED_DIMENSIONS. Every call on one GPU shares that array: calls that overlap on
 * one GPU (from several contexts or threads) race. */
int vib_nis_gate_static(vib_context* ctx, int32_t slot, double threshold, int32_t apply, vib_gate_report* report,
                        uint8_t* decisions);

/* SolveTermination (BackendTypes.hpp:216) */
#define VIB_TERM_FAILURE 0
#define VIB_TERM_SOLVER_FINISHED 1
#define VIB_TERM_TIME_BUDGET 2
#define VIB_TERM_ITERATION_BUDGET 3
#define VIB_TERM_REFIT_BUDGET 4

typedef struct vib_refit_options {  /* BackendOptions (BackendTypes.hpp:186-190)             */
    double nis_threshold;           /* visualNisThreshold                                     */
    int32_t max_outlier_refits;     /* maxOutlierRefits (3)                                   */
    int32_t skip_pre_nis;           /* SolveBudget::skipPreNis                                */
} vib_refit_options;

typedef struct vib_refit_report {
    vib_solve_report solve;         /* the last pass; iterations and seconds summed over passes,
                                       initial_cost of the first pass                          */
    int32_t passes, pre_rejected, post_rejected, unrefitted_outliers;
    int32_t termination;            /* VIB_TERM_*                                             */
    int32_t pad;
    vib_gate_report pre, post;      /* pre-gate (evaluated/admitted), last post-gate          */
} vib_refit_report;

/* The solve with the pre-gate and the bounded outlier-refit loop of BackendOptimizer::solve
 * (BackendOptimizer.cpp:680-690, 720-737, 1036-1079): gate, then up to max_outlier_refits + 1
 * LM passes sharing max_iterations (each pass reserves one iteration per possible later refit
 * and is charged at least one), each followed by a post-gate that rejects live outliers; the
 * loop ends when a post-gate rejects nothing or a budget is spent. */
int vib_solve_refit(vib_context* ctx, int32_t slot, const vib_solve_options* options,
                    const vib_refit_options* refit, vib_refit_report* report);

/* Device preintegration of `count` intervals in one launch (one wavefront per interval), each
 * the computation of vib_preintegrate (ImuError::redoPreintegration, ImuError.cpp:254-462,
 * which the reference calls once per edge on the host: ImuError.cpp:864, BodyImuFactor.cpp:103,
 * and the re-integration of mergeBodyImuRecords, BodyImuFactor.cpp:306-388). Interval e
 * integrates samples[sample_begin[e] .. sample_end[e]) (rows of VIB_IMU_SAMPLE_STRIDE inside
 * one shared array of num_samples rows) over [t0[e], t1[e]] at speed_bias[9e..9e+8] with gyro
 * readings gyro_start/gyro_end[3e..3e+2], writing out[e * VIB_IMU_STRIDE ...]. status[e] is
 * VIB_OK or VIB_ERR_INVALID_ARGUMENT (the checks of vib_preintegrate; that record is zeroed).
 * kernel_ms (nullable) receives the kernel's device time. Host buffers in and out. */
int vib_preintegrate_batch(vib_context* ctx, int32_t count, const double* samples, int32_t num_samples,
                           const int32_t* sample_begin, const int32_t* sample_end, const double* t0, const double* t1,
                           const double* speed_bias, const vib_imu_params* params, const double* gyro_start,
                           const double* gyro_end, double* out, int32_t* status, double* kernel_ms);

/* Device observation whitening (FrameStore::observation, FrameStore.cpp:102-124; undistorted
 * PinholeCamera::backProject, PinholeCamera.hpp:778-807) of `count` pixels (u, v interleaved)
 * from the cameras camera[i] (nullable when num_cameras == 1) described by intrinsics
 * (num_cameras x 6: fx, fy, cx, cy, width, height), pixel sigma[i] (nullable: sigma_all for
 * all). Writes unit bearings (count x 3) and (s00, s10, s11) of the lower-triangular
 * square-root information in the sphere basis (count x 3): the observation record of
 * vib_window. status[i]: 0 ok, 1 outside the image ("observation is outside valid camera
 * image"), 2 back-projection failed, 3 singular covariance / invalid whitening, 4 invalid sigma
 * or camera index; failed records are zero. Returns VIB_ERR_INVALID_ARGUMENT when any status is
 * nonzero. kernel_ms (nullable) receives the kernel's device time. Host buffers in and out. */
int vib_whiten_observations(vib_context* ctx, int32_t count, const double* pixels, const int32_t* camera,
                            const double* intrinsics, int32_t num_cameras, const double* sigma, double sigma_all,
                            double* bearings, double* sqrt_info, int32_t* status, double* kernel_ms);

/* ---------------- host-side helpers (CPU only, no device) ---------------- */
/* IMU preintegration between t0 and t1 at the given speed/bias linearization point,
 * written as one VIB_IMU_STRIDE record (ImuError::redoPreintegration, ImuError.cpp:254-462). */
int vib_preintegrate(const double* samples, int32_t count, double t0, double t1,
                     const double* speed_bias, const vib_imu_params* params,
                     const double* gyro_start, const double* gyro_end, double* out);
/* Observation whitening for `count` pixels (u, v interleaved) of an undistorted pinhole camera
 * with pixel sigma: unit bearings (count*3) and (s00, s10, s11) of the lower-triangular
 * square-root information in the sphere basis (FrameStore::observation, FrameStore.cpp:102-124;
 * PinholeCamera::backProject, PinholeCamera.hpp:778-807). */
int vib_pinhole_observations(double fx, double fy, double cx, double cy, double sigma, int32_t count,
                             const double* pixels, double* bearings, double* sqrt_info);

#ifdef __cplusplus
}
#endif

#endif /* VIB_H */
