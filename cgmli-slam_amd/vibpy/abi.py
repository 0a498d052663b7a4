"""ctypes binding of include/vib.h.

`Window` owns numpy arrays for every field of `vib_window`; `Window.cstruct()` returns a
`VibWindow` whose pointers alias those arrays (the Window must outlive the call).
`load_library()` opens the in-tree `libvib.so` and fails loudly if it is missing.
"""
import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "build", "libvib.so")

VIB_OK = 0
STATUS_NAMES = {
    0: "OK", 1: "INVALID_ARGUMENT", 2: "UNSUPPORTED_DIMENSIONS", 3: "INVALID_RESIDUAL",
    4: "NOT_POSITIVE_DEFINITE", 5: "COST_MISMATCH", 6: "NO_ACCEPTABLE_STEP", 7: "UNSUPPORTED_LOSS",
    8: "DEVICE", 9: "OUT_OF_MEMORY", 10: "NONFINITE_STEP", 11: "INDEFINITE", 12: "EIGEN_FAILED",
    13: "NO_DEVICE", 14: "NOT_FOUND", 15: "UNSUPPORTED", 16: "LIGHTWEIGHT_SUPPORT", 17: "IMU_BRIDGE",
    18: "IMU_GATE",
}
VIB_ERR_INVALID_ARGUMENT = 1
VIB_ERR_UNSUPPORTED_DIMENSIONS = 2
VIB_ERR_INVALID_RESIDUAL = 3
VIB_ERR_NOT_POSITIVE_DEFINITE = 4
VIB_ERR_COST_MISMATCH = 5
VIB_ERR_NO_ACCEPTABLE_STEP = 6
VIB_ERR_UNSUPPORTED_LOSS = 7
VIB_ERR_DEVICE = 8
VIB_ERR_OUT_OF_MEMORY = 9
VIB_ERR_NONFINITE_STEP = 10
VIB_ERR_NOT_FOUND = 14
VIB_ERR_UNSUPPORTED = 15
VIB_ERR_LIGHTWEIGHT_SUPPORT = 16
VIB_ERR_IMU_BRIDGE = 17
VIB_ERR_IMU_GATE = 18
VIB_RIG_FREE = 1
VIB_TD_FREE = 2
VIB_KEY_POSE, VIB_KEY_SPEED_BIAS, VIB_KEY_RIG, VIB_KEY_TD = 0, 1, 2, 3
VIB_FP64, VIB_MIXED = 0, 1
VIB_MARGIN_OLD, VIB_MARGIN_SECOND_NEW = 0, 1
IMU_STRIDE = 301

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int32)
lp = C.POINTER(C.c_int64)
up = C.POINTER(C.c_uint8)


class VibImuParams(C.Structure):
    _fields_ = [("sigma_g_c", C.c_double), ("sigma_a_c", C.c_double), ("sigma_gw_c", C.c_double),
                ("sigma_aw_c", C.c_double), ("g", C.c_double), ("g_max", C.c_double), ("a_max", C.c_double),
                ("p_BS", C.c_double * 3), ("q_BS", C.c_double * 4)]


class VibWindow(C.Structure):
    _fields_ = [
        ("num_frames", C.c_int32), ("num_cameras", C.c_int32), ("num_landmarks", C.c_int32),
        ("num_observations", C.c_int32), ("num_imu", C.c_int32), ("flags", C.c_uint32),
        ("huber_delta", C.c_double), ("min_inverse_depth", C.c_double),
        ("td_lower", C.c_double), ("td_upper", C.c_double),
        ("frame_id", lp), ("pose", dp), ("speed_bias", dp), ("pose_free", up), ("sb_width", up),
        ("rig", dp), ("td", dp), ("cam_rotation", dp), ("cam_translation", dp),
        ("lm_anchor_frame", ip), ("lm_anchor_cam", ip), ("lm_bearing", dp), ("lm_velocity", dp),
        ("lm_tref", dp), ("lm_rho", dp), ("lm_free", up),
        ("obs_landmark", ip), ("obs_frame", ip), ("obs_cam", ip), ("obs_bearing", dp),
        ("obs_velocity", dp), ("obs_tref", dp), ("obs_sqrt_info", dp),
        ("imu_from", ip), ("imu_to", ip), ("imu_data", dp),
        ("prior_rows", C.c_int32), ("prior_blocks", C.c_int32), ("prior_key_kind", ip),
        ("prior_key_id", lp), ("prior_ref", dp), ("prior_A", dp), ("prior_r0", dp), ("lm_id", lp),
    ]


class VibTuneReport(C.Structure):
    _fields_ = [("candidates", C.c_int32), ("best", C.c_int32), ("size_class", C.c_int32),
                ("landmarks_per_chunk", C.c_int32), ("factors_per_range", C.c_int32),
                ("chunk_merge", C.c_int32 * 16), ("range_merge", C.c_int32 * 16), ("ms_per_iteration", C.c_double * 16)]


class VibServerOptions(C.Structure):
    _fields_ = [("hbm_budget", C.c_int64), ("precision", C.c_int32), ("refinement_steps", C.c_int32),
                ("slide_policy", C.c_int32), ("full_iterations", C.c_int32), ("first_slot", C.c_int32),
                ("num_buckets", C.c_int32), ("bucket_limit", C.c_int32 * 8), ("workers", C.c_int32)]


class VibServed(C.Structure):
    _fields_ = [("request", C.c_int64), ("session", C.c_int32), ("status", C.c_int32), ("iterations", C.c_int32),
                ("unmatched", C.c_int32), ("final_cost", C.c_double), ("queued_seconds", C.c_double),
                ("latency_seconds", C.c_double)]


class VibServerStats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("submitted", "finished", "failed", "turns", "launches", "window_iterations",
                                         "uploads", "evictions", "appends", "slides", "slide_launches",
                                         "resident_bytes")] + \
               [(n, C.c_int32) for n in ("queued", "running", "resident", "pending_results")] + \
               [(n, C.c_double) for n in ("admit_seconds", "iterate_seconds", "slide_seconds")]


class VibKeyframe(C.Structure):
    _fields_ = [
        ("frame_id", C.c_int64), ("pose", dp), ("speed_bias", dp), ("pose_free", C.c_int32), ("sb_width", C.c_int32),
        ("imu_data", dp), ("num_landmarks", C.c_int32), ("num_observations", C.c_int32),
        ("lm_id", lp), ("lm_anchor_cam", ip), ("lm_bearing", dp), ("lm_velocity", dp), ("lm_tref", dp),
        ("lm_rho", dp), ("lm_free", up), ("obs_landmark_id", lp), ("obs_cam", ip), ("obs_bearing", dp),
        ("obs_velocity", dp), ("obs_tref", dp), ("obs_sqrt_info", dp),
    ]


class VibSolveOptions(C.Structure):
    _fields_ = [("max_iterations", C.c_int32), ("min_iterations", C.c_int32),
                ("initial_lambda", C.c_double), ("time_limit_seconds", C.c_double),
                ("precision", C.c_int32), ("refinement_steps", C.c_int32), ("flags", C.c_uint32)]


class VibStepReport(C.Structure):
    _fields_ = [("cost", C.c_double), ("predicted_reduction", C.c_double), ("trial_cost", C.c_double),
                ("max_increment", C.c_double), ("trial_valid", C.c_int32), ("status", C.c_int32)]


class VibSolveReport(C.Structure):
    _fields_ = [("status", C.c_int32), ("iterations", C.c_int32), ("accepted_steps", C.c_int32),
                ("converged", C.c_int32), ("effective_parameters", C.c_int32),
                ("retained_dimension", C.c_int32), ("initial_cost", C.c_double),
                ("final_cost", C.c_double), ("final_lambda", C.c_double), ("seconds", C.c_double)]


class VibResidualStats(C.Structure):
    _fields_ = [("factors", C.c_double), ("components", C.c_double), ("over_gate", C.c_double),
                ("squared_norm", C.c_double), ("robust_cost", C.c_double), ("median_nis", C.c_double),
                ("p95_nis", C.c_double), ("max_nis", C.c_double), ("worst_factor", C.c_int32), ("pad", C.c_int32)]


class VibGateReport(C.Structure):
    _fields_ = [("status", C.c_int32), ("rejected", C.c_int32), ("skipped", C.c_int32), ("invalid", C.c_int32),
                ("evaluated", VibResidualStats), ("admitted", VibResidualStats)]


class VibRefitOptions(C.Structure):
    _fields_ = [("nis_threshold", C.c_double), ("max_outlier_refits", C.c_int32), ("skip_pre_nis", C.c_int32)]


class VibRefitReport(C.Structure):
    _fields_ = [("solve", VibSolveReport), ("passes", C.c_int32), ("pre_rejected", C.c_int32),
                ("post_rejected", C.c_int32), ("unrefitted_outliers", C.c_int32), ("termination", C.c_int32),
                ("pad", C.c_int32), ("pre", VibGateReport), ("post", VibGateReport)]


VISUAL_NIS_THRESHOLD = 9.210340371976184  # chi-square, 2 dof, 0.99 (BackendTypes.hpp:186)
TERM_FAILURE, TERM_SOLVER_FINISHED, TERM_TIME_BUDGET, TERM_ITERATION_BUDGET, TERM_REFIT_BUDGET = range(5)


class VibLatestTwoReport(C.Structure):
    _fields_ = [("solve", VibSolveReport), ("old_imu_nis_before", C.c_double), ("old_imu_nis_after", C.c_double),
                ("new_imu_nis_before", C.c_double), ("new_imu_nis_after", C.c_double),
                ("support_landmarks", C.c_int32)]


class VibMarginalReport(C.Structure):
    _fields_ = [("status", C.c_int32), ("prior_reused", C.c_int32), ("eliminated_dimension", C.c_int32),
                ("retained_dimension", C.c_int32), ("prior_rank", C.c_int32),
                ("eliminated_landmarks", C.c_int32), ("consumed_factors", C.c_int32),
                ("eigen_cutoff", C.c_double), ("imu_mode", C.c_int32), ("reset_landmarks", C.c_int32),
                ("demoted_landmarks", C.c_int32), ("dropped_factors", C.c_int32)]


VIB_IMU_MERGE_NONE, VIB_IMU_MERGE_REINTEGRATED, VIB_IMU_MERGE_INCREMENTAL = 0, 1, 2
VIB_IMU_MERGE_SUPPRESSED, VIB_IMU_MERGE_BRIDGE = 3, 4


VIB_SOLVE_CONTINUOUS = 1


def options(max_iterations=8, min_iterations=0, initial_lambda=1e-4, time_limit=0.0,
            precision=VIB_FP64, refinement_steps=0, flags=0):
    return VibSolveOptions(max_iterations, min_iterations, initial_lambda, time_limit, precision,
                           refinement_steps, flags)


def _ptr(a, ctype):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


@dataclass
class Window:
    """One sliding window in the C-ABI layout (all arrays C-contiguous)."""
    frame_id: np.ndarray
    pose: np.ndarray            # (K, 7)
    speed_bias: np.ndarray      # (K, 9)
    rig: np.ndarray             # (7,)
    td: np.ndarray              # (1,)
    cam_rotation: np.ndarray    # (C, 3, 3)
    cam_translation: np.ndarray  # (C, 3)
    lm_anchor_frame: np.ndarray
    lm_anchor_cam: np.ndarray
    lm_bearing: np.ndarray      # (L, 3)
    lm_velocity: np.ndarray     # (L, 3)
    lm_tref: np.ndarray
    lm_rho: np.ndarray
    obs_landmark: np.ndarray
    obs_frame: np.ndarray
    obs_cam: np.ndarray
    obs_bearing: np.ndarray     # (F, 3)
    obs_velocity: np.ndarray    # (F, 3)
    obs_tref: np.ndarray
    obs_sqrt_info: np.ndarray   # (F, 3)
    imu_from: np.ndarray
    imu_to: np.ndarray
    imu_data: np.ndarray        # (E, 301)
    flags: int = 0
    huber_delta: float = 1.0
    min_inverse_depth: float = 1e-8
    td_lower: float = -0.05
    td_upper: float = 0.05
    pose_free: np.ndarray = None
    sb_width: np.ndarray = None
    lm_free: np.ndarray = None
    prior_key_kind: np.ndarray = None
    prior_key_id: np.ndarray = None
    prior_ref: np.ndarray = None
    prior_A: np.ndarray = None   # (m, n)
    prior_r0: np.ndarray = None
    lm_id: np.ndarray = None     # (L,) strictly increasing landmark ids (None: 0 .. L-1)
    meta: dict = field(default_factory=dict)

    def __post_init__(self):
        f64 = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        i32 = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.int32)
        i64 = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.int64)
        u8 = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.uint8)
        for name in ("pose", "speed_bias", "rig", "td", "cam_rotation", "cam_translation", "lm_bearing",
                     "lm_velocity", "lm_tref", "lm_rho", "obs_bearing", "obs_velocity", "obs_tref",
                     "obs_sqrt_info", "imu_data", "prior_ref", "prior_A", "prior_r0"):
            setattr(self, name, f64(getattr(self, name)))
        for name in ("lm_anchor_frame", "lm_anchor_cam", "obs_landmark", "obs_frame", "obs_cam", "imu_from",
                     "imu_to", "prior_key_kind"):
            setattr(self, name, i32(getattr(self, name)))
        for name in ("frame_id", "prior_key_id", "lm_id"):
            setattr(self, name, i64(getattr(self, name)))
        for name in ("pose_free", "sb_width", "lm_free"):
            setattr(self, name, u8(getattr(self, name)))

    # sizes
    @property
    def K(self):
        return int(self.pose.shape[0])

    @property
    def L(self):
        return int(self.lm_rho.shape[0])

    @property
    def F(self):
        return int(self.obs_landmark.shape[0])

    @property
    def E(self):
        return int(self.imu_from.shape[0])

    @property
    def prior_rows(self):
        return 0 if self.prior_A is None else int(self.prior_A.shape[0])

    def copy(self):
        kw = {}
        for k, v in self.__dict__.items():
            kw[k] = v.copy() if isinstance(v, np.ndarray) else (dict(v) if isinstance(v, dict) else v)
        return Window(**kw)

    def layout(self):
        """Tangent column layout shared by the library and the oracle:
        [pose_k (6, free), rig (6), td (1), sb_k (width)] then free inverse ranges."""
        K = self.K
        pf = np.ones(K, np.uint8) if self.pose_free is None else self.pose_free
        sw = np.full(K, 9, np.uint8) if self.sb_width is None else self.sb_width
        lf = np.ones(self.L, np.uint8) if self.lm_free is None else self.lm_free
        c = 0
        pose_col = np.full(K, -1)
        for k in range(K):
            if pf[k]:
                pose_col[k] = c
                c += 6
        rig_col = td_col = -1
        if self.flags & VIB_RIG_FREE:
            rig_col = c
            c += 6
        if self.flags & VIB_TD_FREE:
            td_col = c
            c += 1
        nv = c
        sb_col = np.full(K, -1)
        for k in range(K):
            if sw[k]:
                sb_col[k] = c
                c += int(sw[k])
        nr = c
        lm_col = np.full(self.L, -1)
        idx = np.nonzero(lf)[0]
        lm_col[idx] = nr + np.arange(idx.size)
        return dict(pose=pose_col, rig=rig_col, td=td_col, sb=sb_col, lm=lm_col, nv=nv, nr=nr,
                    ne=int(idx.size))

    def cstruct(self):
        w = VibWindow()
        w.num_frames = self.K
        w.num_cameras = int(self.cam_rotation.shape[0])
        w.num_landmarks = self.L
        w.num_observations = self.F
        w.num_imu = self.E
        w.flags = self.flags
        w.huber_delta = self.huber_delta
        w.min_inverse_depth = self.min_inverse_depth
        w.td_lower, w.td_upper = self.td_lower, self.td_upper
        w.frame_id = _ptr(self.frame_id, C.c_int64)
        w.lm_id = _ptr(self.lm_id, C.c_int64)
        w.pose = _ptr(self.pose, C.c_double)
        w.speed_bias = _ptr(self.speed_bias, C.c_double)
        w.pose_free = _ptr(self.pose_free, C.c_uint8)
        w.sb_width = _ptr(self.sb_width, C.c_uint8)
        w.rig = _ptr(self.rig, C.c_double)
        w.td = _ptr(self.td, C.c_double)
        w.cam_rotation = _ptr(self.cam_rotation, C.c_double)
        w.cam_translation = _ptr(self.cam_translation, C.c_double)
        w.lm_anchor_frame = _ptr(self.lm_anchor_frame, C.c_int32)
        w.lm_anchor_cam = _ptr(self.lm_anchor_cam, C.c_int32)
        w.lm_bearing = _ptr(self.lm_bearing, C.c_double)
        w.lm_velocity = _ptr(self.lm_velocity, C.c_double)
        w.lm_tref = _ptr(self.lm_tref, C.c_double)
        w.lm_rho = _ptr(self.lm_rho, C.c_double)
        w.lm_free = _ptr(self.lm_free, C.c_uint8)
        w.obs_landmark = _ptr(self.obs_landmark, C.c_int32)
        w.obs_frame = _ptr(self.obs_frame, C.c_int32)
        w.obs_cam = _ptr(self.obs_cam, C.c_int32)
        w.obs_bearing = _ptr(self.obs_bearing, C.c_double)
        w.obs_velocity = _ptr(self.obs_velocity, C.c_double)
        w.obs_tref = _ptr(self.obs_tref, C.c_double)
        w.obs_sqrt_info = _ptr(self.obs_sqrt_info, C.c_double)
        w.imu_from = _ptr(self.imu_from, C.c_int32)
        w.imu_to = _ptr(self.imu_to, C.c_int32)
        w.imu_data = _ptr(self.imu_data, C.c_double)
        m = self.prior_rows
        w.prior_rows = m
        w.prior_blocks = 0 if not m else int(self.prior_key_kind.shape[0])
        w.prior_key_kind = _ptr(self.prior_key_kind, C.c_int32)
        w.prior_key_id = _ptr(self.prior_key_id, C.c_int64)
        w.prior_ref = _ptr(self.prior_ref, C.c_double)
        w.prior_A = _ptr(self.prior_A, C.c_double)
        w.prior_r0 = _ptr(self.prior_r0, C.c_double)
        return w


_LIB = None


def load_library(path=LIB_PATH):
    """Open libvib.so. Raises if the native library has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise RuntimeError(f"native library missing: {path} (run __graft_entry__.build())")
    lib = C.CDLL(path)
    vp = C.c_void_p
    sig = {
        "vib_abi_version": ([], C.c_int),
        "vib_window_append": ([vp, C.c_int32, C.POINTER(VibKeyframe), ip], C.c_int),
        "vib_marginalize_batch": ([vp, C.c_int32, ip, C.c_int32, C.POINTER(VibMarginalReport)], C.c_int),
        "vib_autotune": ([vp, C.c_int32, ip, C.POINTER(VibSolveOptions), C.c_int32, C.POINTER(VibTuneReport)], C.c_int),
        "vib_server_create": ([vp, C.POINTER(VibServerOptions), C.POINTER(vp)], C.c_int),
        "vib_server_destroy": ([vp], C.c_int),
        "vib_server_add_session": ([vp, C.c_int32, C.POINTER(VibWindow)], C.c_int),
        "vib_server_submit": ([vp, C.c_int64, C.c_int32, C.c_int32, C.POINTER(VibKeyframe)], C.c_int),
        "vib_server_turn": ([vp, C.POINTER(VibServed), C.c_int32, ip], C.c_int),
        "vib_server_get_stats": ([vp, C.POINTER(VibServerStats)], C.c_int),
        "vib_server_preload": ([vp], C.c_int),
        "vib_server_session_slot": ([vp, C.c_int32, ip], C.c_int),
        "vib_server_session_context": ([vp, C.c_int32, C.POINTER(vp), ip], C.c_int),
        "vib_context_device": ([vp, ip], C.c_int),
        "vib_slot_bytes": ([vp, C.c_int32, C.POINTER(C.c_int64)], C.c_int),
        "vib_create": ([C.POINTER(vp), C.c_int], C.c_int),
        "vib_destroy": ([vp], C.c_int),
        "vib_last_error": ([vp], C.c_char_p),
        "vib_device_synchronize": ([vp], C.c_int),
        "vib_window_upload": ([vp, C.c_int32, C.POINTER(VibWindow)], C.c_int),
        "vib_window_footprint": ([C.POINTER(VibWindow), C.c_int32], C.c_int64),
        "vib_window_release": ([vp, C.c_int32], C.c_int),
        "vib_set_hbm_budget": ([vp, C.c_int64], C.c_int),
        "vib_set_host_retained": ([vp, C.c_int32, dp, dp], C.c_int),
        "vib_resident_bytes": ([vp, lp], C.c_int),
        "vib_read_state": ([vp, C.c_int32, C.POINTER(VibWindow)], C.c_int),
        "vib_window_shape": ([vp, C.c_int32] + [ip] * 6, C.c_int),
        "vib_linearize": ([vp, C.c_int32, C.c_int32], C.c_int),
        "vib_solve_step": ([vp, C.c_int32, C.c_double, C.c_int32, C.c_int32, C.POINTER(VibStepReport)], C.c_int),
        "vib_accept": ([vp, C.c_int32], C.c_int),
        "vib_read_step": ([vp, C.c_int32, dp, C.c_int32], C.c_int),
        "vib_read_reduced": ([vp, C.c_int32, dp, dp, C.c_int32], C.c_int),
        "vib_read_normal": ([vp, C.c_int32, dp, dp, dp, C.c_int32], C.c_int),
        "vib_retained_dimension": ([vp, C.c_int32, ip, ip], C.c_int),
        "vib_solve": ([vp, C.c_int32, C.POINTER(VibSolveOptions), C.POINTER(VibSolveReport)], C.c_int),
        "vib_solve_batch": ([vp, C.c_int32, ip, C.POINTER(VibSolveOptions), C.POINTER(VibSolveReport)], C.c_int),
        "vib_solve_latest_two": ([vp, C.c_int32, C.POINTER(VibSolveOptions), C.c_int32, C.POINTER(VibLatestTwoReport)],
                                 C.c_int),
        "vib_marginalize": ([vp, C.c_int32, C.c_int32, dp, C.c_int32, C.c_double, C.c_double,
                             C.POINTER(VibImuParams), C.POINTER(VibMarginalReport)], C.c_int),
        "vib_read_window": ([vp, C.c_int32, C.POINTER(VibWindow)], C.c_int),
        "vib_request_begin": ([vp, C.c_int32, C.POINTER(VibSolveOptions)], C.c_int),
        "vib_request_iterate": ([vp, C.c_int32, ip, C.c_int32, C.c_int32, ip, C.POINTER(VibSolveReport)], C.c_int),
        "vib_read_prior": ([vp, C.c_int32, ip, ip, ip, ip, lp, dp, dp, dp, C.c_int32, C.c_int32, C.c_int32], C.c_int),
        "vib_kernel_count": ([], C.c_int),
        "vib_kernel_name": ([C.c_int32], C.c_char_p),
        "vib_bench": ([vp, C.c_int32, ip, C.POINTER(VibSolveOptions), C.c_int32, dp, dp], C.c_int),
        "vib_copy_probe": ([vp, C.c_int64], C.c_int),
        "vib_preintegrate": ([dp, C.c_int32, C.c_double, C.c_double, dp, C.POINTER(VibImuParams), dp, dp, dp], C.c_int),
        "vib_nis_gate": ([vp, C.c_int32, C.c_double, C.c_int32, C.POINTER(VibGateReport), C.POINTER(C.c_uint8)], C.c_int),
        "vib_nis_gate_static": ([vp, C.c_int32, C.c_double, C.c_int32, C.POINTER(VibGateReport), C.POINTER(C.c_uint8)],
                                C.c_int),
        "vib_solve_refit": ([vp, C.c_int32, C.POINTER(VibSolveOptions), C.POINTER(VibRefitOptions),
                             C.POINTER(VibRefitReport)], C.c_int),
        "vib_whiten_observations": ([vp, C.c_int32, dp, ip, dp, C.c_int32, dp, C.c_double, dp, dp, ip, dp], C.c_int),
        "vib_preintegrate_batch": ([vp, C.c_int32, dp, C.c_int32, ip, ip, dp, dp, dp, C.POINTER(VibImuParams), dp, dp,
                                    dp, ip, dp], C.c_int),
        "vib_pinhole_observations": ([C.c_double] * 5 + [C.c_int32, dp, dp, dp], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _LIB = lib
    return lib


def f64p(a):
    return a.ctypes.data_as(dp)


def i32p(a):
    return a.ctypes.data_as(ip)


@dataclass
class Keyframe:
    """A new keyframe for vib_window_append: state, IMU edge from the newest frame, the landmarks
    it anchors (ids above every id of the window) and its observations (by landmark id)."""
    frame_id: int
    pose: np.ndarray                 # (7,)
    speed_bias: np.ndarray           # (9,)
    lm_id: np.ndarray
    lm_anchor_cam: np.ndarray
    lm_bearing: np.ndarray           # (N, 3)
    lm_velocity: np.ndarray          # (N, 3)
    lm_tref: np.ndarray
    lm_rho: np.ndarray
    obs_landmark_id: np.ndarray
    obs_cam: np.ndarray
    obs_bearing: np.ndarray          # (M, 3)
    obs_velocity: np.ndarray         # (M, 3)
    obs_tref: np.ndarray
    obs_sqrt_info: np.ndarray        # (M, 3)
    imu_data: np.ndarray = None      # (IMU_STRIDE,)
    lm_free: np.ndarray = None
    pose_free: int = 1
    sb_width: int = 9

    def __post_init__(self):
        for name in ("pose", "speed_bias", "lm_bearing", "lm_velocity", "lm_tref", "lm_rho", "obs_bearing",
                     "obs_velocity", "obs_tref", "obs_sqrt_info", "imu_data"):
            a = getattr(self, name)
            setattr(self, name, None if a is None else np.ascontiguousarray(a, dtype=np.float64))
        for name in ("lm_anchor_cam", "obs_cam"):
            setattr(self, name, np.ascontiguousarray(getattr(self, name), dtype=np.int32))
        for name in ("lm_id", "obs_landmark_id"):
            setattr(self, name, np.ascontiguousarray(getattr(self, name), dtype=np.int64))
        if self.lm_free is not None:
            self.lm_free = np.ascontiguousarray(self.lm_free, dtype=np.uint8)

    def cstruct(self):
        k = VibKeyframe()
        k.frame_id = int(self.frame_id)
        k.pose = _ptr(self.pose, C.c_double)
        k.speed_bias = _ptr(self.speed_bias, C.c_double)
        k.pose_free = int(self.pose_free)
        k.sb_width = int(self.sb_width)
        k.imu_data = _ptr(self.imu_data, C.c_double)
        k.num_landmarks = int(self.lm_id.shape[0])
        k.num_observations = int(self.obs_landmark_id.shape[0])
        k.lm_id = _ptr(self.lm_id, C.c_int64)
        k.lm_anchor_cam = _ptr(self.lm_anchor_cam, C.c_int32)
        k.lm_bearing = _ptr(self.lm_bearing, C.c_double)
        k.lm_velocity = _ptr(self.lm_velocity, C.c_double)
        k.lm_tref = _ptr(self.lm_tref, C.c_double)
        k.lm_rho = _ptr(self.lm_rho, C.c_double)
        k.lm_free = _ptr(self.lm_free, C.c_uint8)
        k.obs_landmark_id = _ptr(self.obs_landmark_id, C.c_int64)
        k.obs_cam = _ptr(self.obs_cam, C.c_int32)
        k.obs_bearing = _ptr(self.obs_bearing, C.c_double)
        k.obs_velocity = _ptr(self.obs_velocity, C.c_double)
        k.obs_tref = _ptr(self.obs_tref, C.c_double)
        k.obs_sqrt_info = _ptr(self.obs_sqrt_info, C.c_double)
        return k


def window_buffers(K, C_, L, F, E, prior_rows, prior_cols, prior_blocks, prior_ref):
    """Zeroed arrays of a window at the given capacities, for the calls that write a window."""
    z = np.zeros
    return Window(frame_id=z(K, np.int64), pose=z((K, 7)), speed_bias=z((K, 9)), rig=z(7), td=z(1),
                  cam_rotation=z((C_, 3, 3)), cam_translation=z((C_, 3)),
                  lm_anchor_frame=z(L, np.int32), lm_anchor_cam=z(L, np.int32), lm_bearing=z((L, 3)),
                  lm_velocity=z((L, 3)), lm_tref=z(L), lm_rho=z(L),
                  obs_landmark=z(F, np.int32), obs_frame=z(F, np.int32), obs_cam=z(F, np.int32),
                  obs_bearing=z((F, 3)), obs_velocity=z((F, 3)), obs_tref=z(F), obs_sqrt_info=z((F, 3)),
                  imu_from=z(E, np.int32), imu_to=z(E, np.int32), imu_data=z((E, IMU_STRIDE)),
                  pose_free=z(K, np.uint8), sb_width=z(K, np.uint8), lm_free=z(L, np.uint8),
                  prior_key_kind=z(max(prior_blocks, 1), np.int32), prior_key_id=z(max(prior_blocks, 1), np.int64),
                  prior_ref=z(max(prior_ref, 1)), prior_A=z((max(prior_rows, 1), max(prior_cols, 1))),
                  prior_r0=z(max(prior_rows, 1)), lm_id=z(L, np.int64))


def trimmed_window(buf, cs, K, L, F, E, rows, cols, blocks, reflen):
    """The window of `buf` (filled through its C struct `cs`) cut to the written sizes."""
    prior = {}
    if rows > 0:
        prior = dict(prior_key_kind=buf.prior_key_kind[:blocks].copy(), prior_key_id=buf.prior_key_id[:blocks].copy(),
                     prior_ref=buf.prior_ref[:reflen].copy(),
                     prior_A=buf.prior_A.reshape(-1)[:rows * cols].reshape(rows, cols).copy(),
                     prior_r0=buf.prior_r0[:rows].copy())
    return Window(frame_id=buf.frame_id[:K].copy(), pose=buf.pose[:K].copy(), speed_bias=buf.speed_bias[:K].copy(),
                  rig=buf.rig.copy(), td=buf.td.copy(), cam_rotation=buf.cam_rotation.copy(),
                  cam_translation=buf.cam_translation.copy(), lm_anchor_frame=buf.lm_anchor_frame[:L].copy(),
                  lm_anchor_cam=buf.lm_anchor_cam[:L].copy(), lm_bearing=buf.lm_bearing[:L].copy(),
                  lm_velocity=buf.lm_velocity[:L].copy(), lm_tref=buf.lm_tref[:L].copy(), lm_rho=buf.lm_rho[:L].copy(),
                  obs_landmark=buf.obs_landmark[:F].copy(), obs_frame=buf.obs_frame[:F].copy(),
                  obs_cam=buf.obs_cam[:F].copy(), obs_bearing=buf.obs_bearing[:F].copy(),
                  obs_velocity=buf.obs_velocity[:F].copy(), obs_tref=buf.obs_tref[:F].copy(),
                  obs_sqrt_info=buf.obs_sqrt_info[:F].copy(), imu_from=buf.imu_from[:E].copy(),
                  imu_to=buf.imu_to[:E].copy(), imu_data=buf.imu_data[:E].copy(), flags=int(cs.flags),
                  huber_delta=cs.huber_delta, min_inverse_depth=cs.min_inverse_depth, td_lower=cs.td_lower,
                  td_upper=cs.td_upper, pose_free=buf.pose_free[:K].copy(), sb_width=buf.sb_width[:K].copy(),
                  lm_free=buf.lm_free[:L].copy(), lm_id=buf.lm_id[:L].copy(), **prior)


def window_footprint(win, precision=VIB_FP64, lib=None):
    """Device bytes the library allocates for `win` in one slot (vib_window_footprint)."""
    lib = lib or load_library()
    s = win.cstruct()
    n = lib.vib_window_footprint(C.byref(s), precision)
    if n < 0:
        raise RuntimeError("vib_window_footprint rejected the window")
    return int(n)


def imu_params(sigma_g_c=1e-3, sigma_a_c=1e-2, sigma_gw_c=2.6867875817855036e-06,
               sigma_aw_c=4.8555923907289822e-06, g=9.81, g_max=35.0, a_max=160.0,
               p_BS=(0.0, 0.0, 0.0), q_BS=(0.0, 0.0, 0.0, 1.0)):
    p = VibImuParams()
    p.sigma_g_c, p.sigma_a_c, p.sigma_gw_c, p.sigma_aw_c = sigma_g_c, sigma_a_c, sigma_gw_c, sigma_aw_c
    p.g, p.g_max, p.a_max = g, g_max, a_max
    p.p_BS[:] = list(p_BS)
    p.q_BS[:] = list(q_BS)
    return p


class Context:
    """Owning wrapper of a vib_context (device `device`)."""

    def __init__(self, device=0, lib=None):
        self.lib = lib or load_library()
        h = C.c_void_p()
        st = self.lib.vib_create(C.byref(h), device)
        if st != VIB_OK:
            raise RuntimeError(f"vib_create failed: {STATUS_NAMES.get(st, st)}")
        self.h = h
        self._owned = True
        self._cams = {}  # cameras per uploaded slot (the shape query does not report them)

    @classmethod
    def borrowed(cls, handle, lib=None):
        """A non-owning view of a context that something else created (a server's worker)."""
        self = cls.__new__(cls)
        self.lib = lib or load_library()
        self.h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle
        self._owned = False
        self._cams = {}
        return self

    def close(self):
        if self.h:
            if self._owned:
                self.lib.vib_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        return self.check(self.lib.vib_device_synchronize(self.h), "vib_device_synchronize")

    def copy_probe(self, nbytes):
        """One plain device copy of nbytes (counter calibration)."""
        return self.check(self.lib.vib_copy_probe(self.h, int(nbytes)), "vib_copy_probe")

    def whiten(self, pixels, intrinsics, camera=None, sigma=1.0, timed=False):
        """Device whitening of pixel observations: (bearings (n,3), sqrt_info (n,3), status (n,)) and,
        with timed=True, the kernel's device milliseconds. intrinsics: (C, 6) fx fy cx cy width height;
        sigma: a scalar or one value per observation."""
        px = np.ascontiguousarray(pixels, np.float64).reshape(-1, 2)
        n = px.shape[0]
        intr = np.ascontiguousarray(intrinsics, np.float64).reshape(-1, 6)
        cam = None if camera is None else np.ascontiguousarray(camera, np.int32)
        sg = None if np.isscalar(sigma) else np.ascontiguousarray(sigma, np.float64)
        b = np.zeros((n, 3))
        s = np.zeros((n, 3))
        status = np.zeros(n, np.int32)
        ms = C.c_double(0)
        st = self.lib.vib_whiten_observations(self.h, n, f64p(px), None if cam is None else i32p(cam), f64p(intr),
                                              intr.shape[0], None if sg is None else f64p(sg),
                                              float(sigma) if sg is None else 0.0, f64p(b), f64p(s), i32p(status),
                                              C.byref(ms) if timed else None)
        if st not in (VIB_OK, VIB_ERR_INVALID_ARGUMENT):
            self.check(st, "vib_whiten_observations")
        return (b, s, status, ms.value) if timed else (b, s, status)

    def preintegrate_batch(self, samples, begin, end, t0, t1, speed_bias, params, gyro_start, gyro_end, timed=False):
        """Device preintegration of many intervals of one sample array (one launch).

        Returns (records (n, IMU_STRIDE), status (n,)) and, with timed=True, the kernel's
        device milliseconds as a third value. Raises only on a device failure; an interval that
        fails the argument checks has status VIB_ERR_INVALID_ARGUMENT and a zero record."""
        s = np.ascontiguousarray(samples, np.float64)
        b = np.ascontiguousarray(begin, np.int32)
        e = np.ascontiguousarray(end, np.int32)
        n = int(b.shape[0])
        a0 = np.ascontiguousarray(t0, np.float64)
        a1 = np.ascontiguousarray(t1, np.float64)
        sb = np.ascontiguousarray(speed_bias, np.float64).reshape(n, 9)
        g0 = np.ascontiguousarray(gyro_start, np.float64).reshape(n, 3)
        g1 = np.ascontiguousarray(gyro_end, np.float64).reshape(n, 3)
        out = np.zeros((n, IMU_STRIDE))
        status = np.zeros(n, np.int32)
        ms = C.c_double(0)
        st = self.lib.vib_preintegrate_batch(self.h, n, f64p(s), int(s.shape[0]), i32p(b), i32p(e), f64p(a0), f64p(a1),
                                             f64p(sb), C.byref(params), f64p(g0), f64p(g1), f64p(out), i32p(status),
                                             C.byref(ms) if timed else None)
        if st not in (VIB_OK, VIB_ERR_INVALID_ARGUMENT):
            self.check(st, "vib_preintegrate_batch")
        return (out, status, ms.value) if timed else (out, status)

    def error(self):
        m = self.lib.vib_last_error(self.h)
        return m.decode() if m else ""

    def check(self, st, what):
        if st != VIB_OK:
            raise RuntimeError(f"{what}: {STATUS_NAMES.get(st, st)}: {self.error()}")
        return st

    def upload(self, slot, win):
        s = win.cstruct()
        self._cams[slot] = int(win.cam_rotation.shape[0])
        return self.check(self.lib.vib_window_upload(self.h, slot, C.byref(s)), "vib_window_upload")

    def append(self, slot, kf):
        """vib_window_append: the keyframe becomes the window's newest frame; returns the number of
        its observations whose landmark is no longer in the window (skipped)."""
        k = kf.cstruct()
        miss = C.c_int32()
        self.check(self.lib.vib_window_append(self.h, slot, C.byref(k), C.byref(miss)), "vib_window_append")
        return int(miss.value)

    def set_host_retained(self, slot, H=None, g=None):
        """Normal equations of host-evaluated factors (None clears them)."""
        if H is None:
            return self.check(self.lib.vib_set_host_retained(self.h, slot, None, None), "vib_set_host_retained")
        H = np.ascontiguousarray(H, np.float64)
        g = np.ascontiguousarray(g, np.float64)
        return self.check(self.lib.vib_set_host_retained(self.h, slot, f64p(H), f64p(g)), "vib_set_host_retained")

    def set_hbm_budget(self, nbytes):
        return self.check(self.lib.vib_set_hbm_budget(self.h, int(nbytes)), "vib_set_hbm_budget")

    def resident_bytes(self):
        v = C.c_int64()
        self.check(self.lib.vib_resident_bytes(self.h, C.byref(v)), "vib_resident_bytes")
        return v.value

    def try_upload(self, slot, win):
        """vib_window_upload returning its status instead of raising."""
        s = win.cstruct()
        return self.lib.vib_window_upload(self.h, slot, C.byref(s))

    def release_if_present(self, slot):
        """vib_window_release that ignores an empty slot."""
        st = self.lib.vib_window_release(self.h, slot)
        if st not in (VIB_OK, VIB_ERR_NOT_FOUND):
            self.check(st, "vib_window_release")

    def release(self, slot):
        return self.check(self.lib.vib_window_release(self.h, slot), "vib_window_release")

    def read_state(self, slot, win):
        s = win.cstruct()
        return self.check(self.lib.vib_read_state(self.h, slot, C.byref(s)), "vib_read_state")

    def linearize(self, slot, precision=VIB_FP64):
        return self.lib.vib_linearize(self.h, slot, precision)

    def solve_step(self, slot, lam, precision=VIB_FP64, refinement=0):
        rep = VibStepReport()
        st = self.lib.vib_solve_step(self.h, slot, lam, precision, refinement, C.byref(rep))
        return st, rep

    def dims(self, slot):
        nr, nv = C.c_int32(), C.c_int32()
        self.check(self.lib.vib_retained_dimension(self.h, slot, C.byref(nr), C.byref(nv)), "dims")
        return nr.value, nv.value

    def shape(self, slot):
        """(frames, landmarks, observations, imu edges, prior rows, prior cols) of a resident window."""
        v = [C.c_int32() for _ in range(6)]
        self.check(self.lib.vib_window_shape(self.h, slot, *[C.byref(x) for x in v]), "vib_window_shape")
        return tuple(x.value for x in v)

    def read_step(self, slot, n):
        out = np.zeros(n)
        self.check(self.lib.vib_read_step(self.h, slot, f64p(out), n), "vib_read_step")
        return out

    def read_reduced(self, slot, nr):
        S = np.zeros((nr, nr))
        rhs = np.zeros(nr)
        self.check(self.lib.vib_read_reduced(self.h, slot, f64p(S), f64p(rhs), nr), "vib_read_reduced")
        return S, rhs

    def read_normal(self, slot, nr):
        H = np.zeros((nr, nr))
        g = np.zeros(nr)
        cost = C.c_double()
        self.check(self.lib.vib_read_normal(self.h, slot, f64p(H), f64p(g), C.byref(cost), nr), "vib_read_normal")
        return H, g, cost.value

    def solve(self, slot, opts):
        rep = VibSolveReport()
        st = self.lib.vib_solve(self.h, slot, C.byref(opts), C.byref(rep))
        return st, rep

    def nis_gate(self, slot, threshold=VISUAL_NIS_THRESHOLD, apply=False):
        """Statistical gate of a resident window: (status, VibGateReport, decisions (F,) uint8 in
        landmark order: 0 admitted, 1 over the gate, 2 invalid, 3 not participating)."""
        n = self.shape(slot)[2]
        dec = np.zeros(max(n, 1), np.uint8)
        rep = VibGateReport()
        st = self.lib.vib_nis_gate(self.h, slot, float(threshold), 1 if apply else 0, C.byref(rep),
                                   dec.ctypes.data_as(C.POINTER(C.c_uint8)))
        return st, rep, dec[:n]

    def nis_gate_static(self, slot, threshold=VISUAL_NIS_THRESHOLD, apply=False):
        """vib_nis_gate_static (the gate without device allocation), returned like nis_gate."""
        n = self.shape(slot)[2]
        dec = np.zeros(max(n, 1), np.uint8)
        rep = VibGateReport()
        st = self.lib.vib_nis_gate_static(self.h, slot, float(threshold), 1 if apply else 0, C.byref(rep),
                                          dec.ctypes.data_as(C.POINTER(C.c_uint8)))
        return st, rep, dec[:n]

    def solve_refit(self, slot, opts, threshold=VISUAL_NIS_THRESHOLD, max_refits=3, skip_pre=False):
        """Solve with the pre-gate and the bounded outlier-refit loop: (status, VibRefitReport)."""
        ro = VibRefitOptions(float(threshold), int(max_refits), 1 if skip_pre else 0)
        rep = VibRefitReport()
        st = self.lib.vib_solve_refit(self.h, slot, C.byref(opts), C.byref(ro), C.byref(rep))
        return st, rep

    def solve_latest_two(self, slot, opts, gate=False):
        rep = VibLatestTwoReport()
        st = self.lib.vib_solve_latest_two(self.h, slot, C.byref(opts), 1 if gate else 0, C.byref(rep))
        return st, rep

    def solve_batch(self, slots, opts):
        arr = np.ascontiguousarray(slots, dtype=np.int32)
        reps = (VibSolveReport * len(arr))()
        st = self.lib.vib_solve_batch(self.h, len(arr), arr.ctypes.data_as(ip), C.byref(opts), reps)
        return st, list(reps)

    def bench(self, slots, opts, iterations, kernels=True):
        arr = np.ascontiguousarray(slots, dtype=np.int32)
        it_ms = np.zeros(iterations)
        nk = self.lib.vib_kernel_count()
        k_ms = np.zeros(nk)
        st = self.lib.vib_bench(self.h, len(arr), arr.ctypes.data_as(ip), C.byref(opts), iterations, f64p(it_ms),
                                f64p(k_ms) if kernels else None)
        self.check(st, "vib_bench")
        names = [self.lib.vib_kernel_name(i).decode() for i in range(nk)]
        # kernels that run inside other launches of an iteration report no time of their own
        return it_ms, {n: v for n, v in zip(names, k_ms) if not (kernels and v == 0.0)}

    def marginalize(self, slot, policy, samples=None, t_start=0.0, t_end=0.0, params=None):
        """vib_marginalize; SECOND_NEW takes the raw IMU samples over [t_start, t_end]."""
        rep = VibMarginalReport()
        smp = None if samples is None else np.ascontiguousarray(samples, np.float64)
        cnt = 0 if smp is None else int(smp.shape[0])
        st = self.lib.vib_marginalize(self.h, slot, policy, None if smp is None else f64p(smp), cnt, t_start, t_end,
                                      None if params is None else C.byref(params), C.byref(rep))
        return st, rep

    def autotune(self, slots, opts, iterations=3):
        """vib_autotune over `slots`: the measured work-list shapes and the chosen one (dict)."""
        ids = np.ascontiguousarray(slots, np.int32)
        r = VibTuneReport()
        self.check(self.lib.vib_autotune(self.h, len(slots), ids.ctypes.data_as(ip), C.byref(opts), iterations,
                                         C.byref(r)), "vib_autotune")
        n = r.candidates
        table = [dict(chunk_merge=r.chunk_merge[i], range_merge=r.range_merge[i], ms=r.ms_per_iteration[i])
                 for i in range(n)]
        return dict(best=table[r.best], size_class=r.size_class, landmarks_per_chunk=r.landmarks_per_chunk,
                    factors_per_range=r.factors_per_range, candidates=table)

    def marginalize_batch(self, slots, policy):
        """vib_marginalize_batch: (first failure, one report per slot)."""
        n = len(slots)
        ids = np.ascontiguousarray(slots, np.int32)
        reps = (VibMarginalReport * n)()
        st = self.lib.vib_marginalize_batch(self.h, n, ids.ctypes.data_as(ip), policy, reps)
        return st, list(reps)

    def request_begin(self, slot, opts):
        return self.check(self.lib.vib_request_begin(self.h, slot, C.byref(opts)), "vib_request_begin")

    def request_iterate(self, slots, precision=VIB_FP64, refinement=0):
        """One batched LM iteration of the running requests on `slots`: (done flags, reports)."""
        n = len(slots)
        ids = np.ascontiguousarray(slots, np.int32)
        done = np.zeros(n, np.int32)
        reps = (VibSolveReport * n)()
        self.check(self.lib.vib_request_iterate(self.h, n, ids.ctypes.data_as(ip), precision, refinement,
                                                done.ctypes.data_as(ip), reps), "vib_request_iterate")
        return done.astype(bool), list(reps)

    def read_window(self, slot, cameras=None):
        """The whole resident window (vib_read_window) as a Window; observations in landmark order.
        cameras: the window's camera count (known from the upload through this object otherwise)."""
        K, L, F, E, rows, cols = self.shape(slot)
        ncam = int(self._cams.get(slot, 0) if cameras is None else cameras)
        blocks = max(cols, 1)
        buf = window_buffers(K, ncam, L, F, E, rows, cols, blocks, 2 * cols + 2)
        cs = buf.cstruct()
        cs.num_frames, cs.num_cameras, cs.num_landmarks, cs.num_observations, cs.num_imu = K, ncam, L, F, E
        cs.prior_rows = rows
        self.check(self.lib.vib_read_window(self.h, slot, C.byref(cs)), "vib_read_window")
        nb = int(cs.prior_blocks)
        reflen = sum(7 if k in (VIB_KEY_POSE, VIB_KEY_RIG) else 9 if k == VIB_KEY_SPEED_BIAS else 1
                     for k in buf.prior_key_kind[:nb])
        return trimmed_window(buf, cs, K, L, F, E, rows, cols, nb, reflen)

    def read_prior(self, slot, max_rows=2048, max_cols=2048, max_blocks=256):
        rows, cols, blocks = C.c_int32(), C.c_int32(), C.c_int32()
        kind = np.zeros(max_blocks, np.int32)
        kid = np.zeros(max_blocks, np.int64)
        ref = np.zeros(max_blocks * 9)
        A = np.zeros(max_rows * max_cols)
        r0 = np.zeros(max_rows)
        st = self.lib.vib_read_prior(self.h, slot, C.byref(rows), C.byref(cols), C.byref(blocks),
                                     kind.ctypes.data_as(ip), kid.ctypes.data_as(lp), f64p(ref), f64p(A),
                                     f64p(r0), max_rows, max_cols, max_blocks)
        self.check(st, "vib_read_prior")
        m, n, b = rows.value, cols.value, blocks.value
        return dict(A=A[:m * n].reshape(m, n), r0=r0[:m], kind=kind[:b], id=kid[:b], ref=ref)


def preintegrate(samples, t0, t1, speed_bias, params, gyro_start, gyro_end, lib=None):
    lib = lib or load_library()
    out = np.zeros(IMU_STRIDE)
    s = np.ascontiguousarray(samples, np.float64)
    sb = np.ascontiguousarray(speed_bias, np.float64)
    g0 = np.ascontiguousarray(gyro_start, np.float64)
    g1 = np.ascontiguousarray(gyro_end, np.float64)
    st = lib.vib_preintegrate(f64p(s), int(s.shape[0]), t0, t1, f64p(sb), C.byref(params), f64p(g0), f64p(g1),
                              f64p(out))
    if st != VIB_OK:
        raise RuntimeError(f"vib_preintegrate: {STATUS_NAMES.get(st, st)}")
    return out


def pinhole_observations(fx, fy, cx, cy, sigma, pixels, lib=None):
    """Vectorized whitening through the native library: returns (bearings (n,3), sqrt_info (n,3))."""
    lib = lib or load_library()
    px = np.ascontiguousarray(pixels, np.float64).reshape(-1, 2)
    n = px.shape[0]
    b = np.zeros((n, 3))
    s = np.zeros((n, 3))
    st = lib.vib_pinhole_observations(fx, fy, cx, cy, sigma, n, f64p(px), f64p(b), f64p(s))
    if st != VIB_OK:
        raise RuntimeError("vib_pinhole_observations failed")
    return b, s
