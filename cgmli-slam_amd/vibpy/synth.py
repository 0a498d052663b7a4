"""Seeded synthetic sliding windows for the benchmark workloads.

A rigid rig of pinhole cameras (720x540, f=350, c=(360,270)) facing front/left/back/right with
small random extrinsic perturbations, a smooth random trajectory, landmarks in a box, noisy
pixel observations whitened into S^2 bearings, 200 Hz IMU samples preintegrated per keyframe
edge (through the native `vib_preintegrate`), and a gauge prior on the first state.

Every window is a pure function of its seed and configuration.
"""
import numpy as np

from . import abi

FX = FY = 350.0
CX, CY = 360.0, 270.0
WIDTH, HEIGHT = 720, 540
IMU_RATE = 200.0
KEYFRAME_DT = 0.1


def rot_x(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])


def rot_y(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])


def rot_z(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


def quat_from_matrix(R):
    """xyzw unit quaternion of a rotation matrix (w >= 0)."""
    t = np.trace(R)
    if t > 0:
        s = np.sqrt(t + 1.0) * 2
        w, x, y, z = 0.25 * s, (R[2, 1] - R[1, 2]) / s, (R[0, 2] - R[2, 0]) / s, (R[1, 0] - R[0, 1]) / s
    elif R[0, 0] > R[1, 1] and R[0, 0] > R[2, 2]:
        s = np.sqrt(1.0 + R[0, 0] - R[1, 1] - R[2, 2]) * 2
        w, x, y, z = (R[2, 1] - R[1, 2]) / s, 0.25 * s, (R[0, 1] + R[1, 0]) / s, (R[0, 2] + R[2, 0]) / s
    elif R[1, 1] > R[2, 2]:
        s = np.sqrt(1.0 + R[1, 1] - R[0, 0] - R[2, 2]) * 2
        w, x, y, z = (R[0, 2] - R[2, 0]) / s, (R[0, 1] + R[1, 0]) / s, 0.25 * s, (R[1, 2] + R[2, 1]) / s
    else:
        s = np.sqrt(1.0 + R[2, 2] - R[0, 0] - R[1, 1]) * 2
        w, x, y, z = (R[1, 0] - R[0, 1]) / s, (R[0, 2] + R[2, 0]) / s, (R[1, 2] + R[2, 1]) / s, 0.25 * s
    q = np.array([x, y, z, w])
    q /= np.linalg.norm(q)
    return q if q[3] >= 0 else -q


def matrix_from_quat(q):
    x, y, z, w = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def exp_so3(v):
    a = np.linalg.norm(v)
    K = np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])
    if a < 1e-12:
        return np.eye(3) + K
    return np.eye(3) + np.sin(a) / a * K + (1 - np.cos(a)) / a ** 2 * K @ K


def random_small_rotation(rng, max_angle):
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    return exp_so3(axis * rng.uniform(0, max_angle))


def make_rig(rng, n_cams, max_rot_deg=2.0, max_trans=0.1):
    """Cameras facing front/left/back/right (body x forward, z up; camera z optical, y down)."""
    R_front = np.array([[0.0, 0.0, 1.0], [-1.0, 0.0, 0.0], [0.0, -1.0, 0.0]])
    Rs, ts = [], []
    for c in range(n_cams):
        yaw = np.pi / 2 * c
        base = rot_z(yaw) @ R_front
        R = base @ random_small_rotation(rng, np.deg2rad(max_rot_deg))
        t = 0.05 * np.array([np.cos(yaw), np.sin(yaw), 0.0]) + rng.uniform(-1, 1, 3) * max_trans / np.sqrt(3)
        Rs.append(R)
        ts.append(t)
    return np.array(Rs), np.array(ts)


class Trajectory:
    """p(t) = c + sum_j a_j sin(w_j tau + phi_j); yaw/pitch/roll smooth sinusoids of tau.

    tau(t) is a time warp: tau' = 1 before `slow_after`, `slow_factor` after it (a near-static
    stretch). Velocity, acceleration and body rate are analytic in t."""

    def __init__(self, rng, speed=1.5, duration=2.0, slow_after=None, slow_factor=1.0):
        self.a = rng.uniform(0.5, 1.5, (3, 3)) * np.array([[1.0], [1.0], [0.2]])
        self.w = rng.uniform(0.4, 1.6, (3, 3))
        self.phi = rng.uniform(0, 2 * np.pi, (3, 3))
        self.c = np.zeros(3)
        self.slow_after = np.inf if slow_after is None else slow_after
        self.slow_factor = slow_factor
        ts = np.linspace(0, duration, 200)
        sp = np.linalg.norm(self._dpos(ts), axis=0).mean()
        self.scale = speed / sp
        self.ang_a = np.array([rng.uniform(0.1, 0.3), rng.uniform(0.02, 0.06), rng.uniform(0.02, 0.06)])
        self.ang_w = rng.uniform(0.5, 1.5, 3)
        self.ang_phi = rng.uniform(0, 2 * np.pi, 3)
        self.yaw0 = rng.uniform(-np.pi, np.pi)

    def warp(self, t):
        t = np.atleast_1d(np.asarray(t, dtype=np.float64))
        late = t > self.slow_after
        tau = t.copy()
        tau[late] = self.slow_after + self.slow_factor * (t[late] - self.slow_after)
        rate = np.where(late, self.slow_factor, 1.0)
        return tau, rate

    def _arg(self, tau):
        return self.w[:, :, None] * tau + self.phi[:, :, None]

    def _pos(self, tau):
        return (self.a[:, :, None] * np.sin(self._arg(tau))).sum(1)

    def _dpos(self, tau):
        return (self.a[:, :, None] * self.w[:, :, None] * np.cos(self._arg(tau))).sum(1)

    def _ddpos(self, tau):
        return -(self.a[:, :, None] * self.w[:, :, None] ** 2 * np.sin(self._arg(tau))).sum(1)

    def position(self, t):
        tau, _ = self.warp(t)
        return (self.c[:, None] + self.scale * self._pos(tau)).T

    def velocity(self, t):
        tau, r = self.warp(t)
        return (self.scale * self._dpos(tau) * r).T

    def acceleration(self, t):
        tau, r = self.warp(t)
        return (self.scale * self._ddpos(tau) * r * r).T

    def angles(self, t):
        tau, r = self.warp(t)
        ang = self.ang_a[:, None] * np.sin(self.ang_w[:, None] * tau + self.ang_phi[:, None])
        dang = self.ang_a[:, None] * self.ang_w[:, None] * np.cos(self.ang_w[:, None] * tau + self.ang_phi[:, None]) * r
        ang[0] += self.yaw0
        return ang.T, dang.T  # (yaw, pitch, roll)

    def rotation(self, t):
        ang, _ = self.angles(t)
        return np.array([rot_z(y) @ rot_y(p) @ rot_x(r) for y, p, r in ang])

    def body_rate(self, t):
        ang, dang = self.angles(t)
        out = []
        for (y, p, r), (dy, dp_, dr) in zip(ang, dang):
            out.append([dr - dy * np.sin(p),
                        dp_ * np.cos(r) + dy * np.cos(p) * np.sin(r),
                        -dp_ * np.sin(r) + dy * np.cos(p) * np.cos(r)])
        return np.array(out)


def imu_samples(traj, t0, t1, rng, sigma_g_c, sigma_a_c, bg, ba, g=9.81):
    n0 = int(np.floor(t0 * IMU_RATE)) - 2
    n1 = int(np.ceil(t1 * IMU_RATE)) + 2
    idx = np.arange(n0, n1 + 1)
    t = idx / IMU_RATE
    R = traj.rotation(t)
    a = traj.acceleration(t)
    w = traj.body_rate(t)
    f = np.einsum('nji,nj->ni', R, a + np.array([0, 0, g]))
    sd_g = sigma_g_c * np.sqrt(IMU_RATE)
    sd_a = sigma_a_c * np.sqrt(IMU_RATE)
    gyro = w + bg + rng.normal(0, sd_g, w.shape)
    acc = f + ba + rng.normal(0, sd_a, f.shape)
    return np.concatenate([t[:, None], gyro, acc], axis=1), idx


def gyro_at(samples, t):
    ts = samples[:, 0]
    j = int(np.searchsorted(ts, t))
    if j == 0:
        return samples[0, 1:4]
    if j >= len(ts):
        return samples[-1, 1:4]
    t0, t1 = ts[j - 1], ts[j]
    al = (t - t0) / (t1 - t0)
    return (1 - al) * samples[j - 1, 1:4] + al * samples[j, 1:4]


def generate(K=10, L=2000, n_cams=4, obs_range=(2, 8), seed=20261204, rig_free=False, td_free=False,
             pixel_sigma=1.0, pixel_noise=None, outlier_fraction=0.0, outlier_px=20.0, speed=1.5, ill_conditioned=False,
             sigma_g_c=1e-3, sigma_a_c=1e-2, init_pos=0.05, init_rot_deg=1.0, init_rho=0.05,
             td_init=0.002, imu_noise=True, static_bearings=False, lib=None):
    """Generate one window. Returns a vibpy.abi.Window with the initial (perturbed) estimate; the
    ground truth is in window.meta. With static_bearings the measured bearing velocities are zero
    (a window without time-offset information, whose tracks a SECOND_NEW slide may re-anchor)."""
    lib = lib or abi.load_library()
    rng = np.random.default_rng(seed)
    cam_R, cam_t = make_rig(rng, n_cams)
    R_BR_true = random_small_rotation(rng, np.deg2rad(1.0))
    t_BR_true = rng.uniform(-0.02, 0.02, 3)
    duration = KEYFRAME_DT * (K - 1)
    frame_t = np.arange(K) * round(KEYFRAME_DT * IMU_RATE) / IMU_RATE
    # ill-conditioned: the second half of the window is a near-static stretch at 0.05 m/s
    traj = Trajectory(rng, speed=speed, duration=max(duration, 0.5),
                      slow_after=(0.5 * duration if ill_conditioned else None),
                      slow_factor=(0.05 / speed if ill_conditioned else 1.0))
    P = traj.position(frame_t)
    R = traj.rotation(frame_t)
    V = traj.velocity(frame_t)
    Wb = traj.body_rate(frame_t)
    # world camera poses per (frame, cam)
    R_WC = np.einsum('kij,jl,clm->kcim', R, R_BR_true, cam_R)
    t_WC = P[:, None, :] + np.einsum('kij,cj->kci', R, (R_BR_true @ cam_t.T).T + t_BR_true)
    center = P.mean(0)
    # landmarks
    pts_list, obs_list = [], []
    need = L
    tries = 0
    while need > 0 and tries < 50:
        tries += 1
        m = int(need * 1.6) + 64
        if ill_conditioned:
            # depth log-uniform 0.5..200 m along random directions from a random frame/camera
            kk = rng.integers(0, K, m)
            cc = rng.integers(0, n_cams, m)
            depth = np.exp(rng.uniform(np.log(0.5), np.log(200.0), m))
            u = rng.uniform(20, WIDTH - 20, m)
            v = rng.uniform(20, HEIGHT - 20, m)
            ray = np.stack([(u - CX) / FX, (v - CY) / FY, np.ones(m)], 1)
            ray /= np.linalg.norm(ray, axis=1, keepdims=True)
            pts = t_WC[kk, cc] + np.einsum('nij,nj->ni', R_WC[kk, cc], ray * depth[:, None])
        else:
            pts = center + rng.uniform(-10.0, 10.0, (m, 3))
        # visibility of every (frame, cam)
        d = pts[:, None, None, :] - t_WC[None]  # m,K,C,3
        pc = np.einsum('kcji,mkcj->mkci', R_WC, d)
        z = pc[..., 2]
        with np.errstate(divide='ignore', invalid='ignore'):
            uu = FX * pc[..., 0] / z + CX
            vv = FY * pc[..., 1] / z + CY
        vis = (z > 0.3) & (uu >= 1) & (uu < WIDTH - 1) & (vv >= 1) & (vv < HEIGHT - 1)
        vis = vis.reshape(m, K * n_cams)
        nvis = vis.sum(1)
        ok = np.nonzero(nvis >= 2)[0]
        for i in ok:
            if need <= 0:
                break
            pairs = np.nonzero(vis[i])[0]
            n = int(rng.integers(obs_range[0], obs_range[1] + 1))
            n = min(n, pairs.size)
            chosen = np.sort(rng.choice(pairs, n, replace=False))
            pts_list.append(pts[i])
            obs_list.append(chosen)
            need -= 1
    L = len(pts_list)
    pts = np.array(pts_list)
    # low-parallax tracks for the ill-conditioned workload: collapse 30% of tracks onto one frame
    # pair set that sees the point from nearly the same place (consecutive frames, same camera)
    # is already implied by depth; push 30% of points far away along their anchor ray.
    anchor_pair = np.array([o[0] for o in obs_list])
    a_frame = anchor_pair // n_cams
    a_cam = anchor_pair % n_cams
    if ill_conditioned:
        sel = rng.random(L) < 0.3
        for i in np.nonzero(sel)[0]:
            k, c = a_frame[i], a_cam[i]
            dvec = pts[i] - t_WC[k, c]
            # parallax < 1 deg over the window baseline: depth >= baseline / tan(1 deg)
            base = np.linalg.norm(P[-1] - P[0]) + 0.1
            pts[i] = t_WC[k, c] + dvec / np.linalg.norm(dvec) * max(np.linalg.norm(dvec), base / np.tan(np.deg2rad(1.0)))
    # measurements
    all_pairs = np.concatenate(obs_list)
    lm_of = np.concatenate([np.full(len(o), i) for i, o in enumerate(obs_list)])
    fk = all_pairs // n_cams
    fc = all_pairs % n_cams
    d = pts[lm_of] - t_WC[fk, fc]
    pc = np.einsum('nji,nj->ni', R_WC[fk, fc], d)
    px = np.stack([FX * pc[:, 0] / pc[:, 2] + CX, FY * pc[:, 1] / pc[:, 2] + CY], 1)
    noisy = px + rng.normal(0, pixel_sigma if pixel_noise is None else pixel_noise, px.shape)
    if outlier_fraction > 0:
        out = rng.random(len(noisy)) < outlier_fraction
        ang = rng.uniform(0, 2 * np.pi, out.sum())
        noisy[out] += outlier_px * np.stack([np.cos(ang), np.sin(ang)], 1)
    bearing, sqrt_info = abi.pinhole_observations(FX, FY, CX, CY, pixel_sigma, noisy, lib=lib)
    # bearing velocity from the true camera motion at each frame
    w_W = np.einsum('kij,kj->ki', R, Wb)  # world angular rate
    R_BC = np.einsum('ij,cjk->cik', R_BR_true, cam_R)
    t_BC = (R_BR_true @ cam_t.T).T + t_BR_true
    w_C = np.einsum('kcji,kj->kci', R_WC, w_W)
    v_WC = V[:, None, :] + np.cross(w_W[:, None, :], np.einsum('kij,cj->kci', R, t_BC))
    v_C = np.einsum('kcji,kcj->kci', R_WC, v_WC)
    pdot = -np.cross(w_C[fk, fc], pc) - v_C[fk, fc]
    rn = np.linalg.norm(pc, axis=1)
    bt = pc / rn[:, None]
    bvel = (pdot - bt * (bt * pdot).sum(1, keepdims=True)) / rn[:, None]
    # a feature-track bearing derivative is tangent to the measured bearing itself
    bvel = bvel - bearing * (bearing * bvel).sum(1, keepdims=True)
    if static_bearings:
        bvel = np.zeros_like(bvel)
    # split anchors / factors
    first_idx = np.concatenate([[0], np.cumsum([len(o) for o in obs_list])[:-1]])
    is_anchor = np.zeros(len(all_pairs), bool)
    is_anchor[first_idx] = True
    lm_bearing = bearing[is_anchor]
    lm_vel = bvel[is_anchor]
    true_rho = 1.0 / rn[is_anchor]
    fac = ~is_anchor
    obs_landmark = lm_of[fac]
    obs_frame = fk[fac]
    obs_cam = fc[fac]
    # IMU
    bg_true = rng.normal(0, 1e-3, 3)
    ba_true = rng.normal(0, 1e-2, 3)
    samples, _ = imu_samples(traj, frame_t[0], frame_t[-1], rng, sigma_g_c if imu_noise else 0.0,
                             sigma_a_c if imu_noise else 0.0, bg_true, ba_true)
    params = abi.imu_params(sigma_g_c=sigma_g_c, sigma_a_c=sigma_a_c)
    # initial estimate: perturbed truth (the gauge frame 0 is kept at truth)
    q_true = np.array([quat_from_matrix(r) for r in R])
    pose = np.zeros((K, 7))
    sb = np.zeros((K, 9))
    for k in range(K):
        if k == 0:
            Rk, pk = R[k], P[k]
        else:
            Rk = random_small_rotation(rng, np.deg2rad(init_rot_deg)) @ R[k]
            pk = P[k] + rng.uniform(-1, 1, 3) * init_pos
        pose[k, :3] = pk
        pose[k, 3:] = quat_from_matrix(Rk)
        sb[k, :3] = V[k] + (0 if k == 0 else rng.normal(0, 0.02, 3))
    rho = true_rho * (1 + rng.uniform(-init_rho, init_rho, L))
    E = K - 1
    imu_data = np.zeros((E, abi.IMU_STRIDE))
    for e in range(E):
        imu_data[e] = abi.preintegrate(samples, frame_t[e], frame_t[e + 1], sb[e], params,
                                       gyro_at(samples, frame_t[e]), gyro_at(samples, frame_t[e + 1]), lib=lib)
    rig_true = np.concatenate([t_BR_true, quat_from_matrix(R_BR_true)])
    rig = rig_true.copy()
    if rig_free:
        Rr = random_small_rotation(rng, np.deg2rad(0.3)) @ R_BR_true
        rig = np.concatenate([t_BR_true + rng.uniform(-0.005, 0.005, 3), quat_from_matrix(Rr)])
    td = np.array([td_init if td_free else 0.0])
    flags = (abi.VIB_RIG_FREE if rig_free else 0) | (abi.VIB_TD_FREE if td_free else 0)
    # gauge prior on frame 0 (pose + speed/bias), r0 = 0 at the initial estimate
    sig = np.array([1e-3] * 3 + [1e-3] * 3 + [0.1] * 3 + [0.01] * 3 + [0.1] * 3)
    A = np.diag(1.0 / sig)
    win = abi.Window(
        frame_id=np.arange(K) + 1000, pose=pose, speed_bias=sb, rig=rig, td=td,
        cam_rotation=cam_R, cam_translation=cam_t,
        lm_anchor_frame=a_frame, lm_anchor_cam=a_cam, lm_bearing=lm_bearing, lm_velocity=lm_vel,
        lm_tref=np.zeros(L), lm_rho=rho,
        obs_landmark=obs_landmark, obs_frame=obs_frame, obs_cam=obs_cam, obs_bearing=bearing[fac],
        obs_velocity=bvel[fac], obs_tref=np.zeros(fac.sum()), obs_sqrt_info=sqrt_info[fac],
        imu_from=np.arange(E), imu_to=np.arange(E) + 1, imu_data=imu_data,
        flags=flags, huber_delta=1.0, min_inverse_depth=1e-8, td_lower=-0.05, td_upper=0.05,
        prior_key_kind=np.array([abi.VIB_KEY_POSE, abi.VIB_KEY_SPEED_BIAS]),
        prior_key_id=np.array([1000, 1000]),
        prior_ref=np.concatenate([pose[0], sb[0]]), prior_A=A, prior_r0=np.zeros(15),
    )
    win.meta = dict(seed=seed, true_pose=np.concatenate([P, q_true], 1),
                    true_sb=np.concatenate([V, np.tile(bg_true, (K, 1)), np.tile(ba_true, (K, 1))], 1),
                    true_rho=true_rho, true_rig=rig_true, samples=samples, frame_t=frame_t, imu_params=params,
                    trajectory=traj)
    return by_anchor_frame(win)


def by_anchor_frame(win):
    """The window with its landmarks in anchor-frame order (stable) and the observations grouped
    by landmark: the order in which a backend creates landmarks as frames arrive."""
    order = np.argsort(win.lm_anchor_frame, kind="stable")
    new = np.empty_like(order)
    new[order] = np.arange(order.size)
    w = win.copy()
    for k in ("lm_anchor_frame", "lm_anchor_cam", "lm_bearing", "lm_velocity", "lm_tref", "lm_rho"):
        setattr(w, k, getattr(win, k)[order].copy())
    if win.lm_free is not None:
        w.lm_free = win.lm_free[order].copy()
    lm = new[win.obs_landmark]
    fo = np.argsort(lm, kind="stable")
    w.obs_landmark = lm[fo].astype(np.int32)
    for k in ("obs_frame", "obs_cam", "obs_bearing", "obs_velocity", "obs_tref", "obs_sqrt_info"):
        setattr(w, k, getattr(win, k)[fo].copy())
    if "true_rho" in w.meta:
        w.meta["true_rho"] = np.asarray(win.meta["true_rho"])[order]
    return w


def camera_world(win, k, c):
    """World pose (R, t) of camera c of slot k: T_WB T_BR T_RC."""
    Rb = matrix_from_quat(win.pose[k, 3:])
    Re = matrix_from_quat(win.rig[3:])
    return Rb @ Re @ win.cam_rotation[c], Rb @ (Re @ win.cam_translation[c] + win.rig[:3]) + win.pose[k, :3]


def reanchor_at(win, frame, min_views=2):
    """Re-anchor the tracks that are observed in `frame` and have at least `min_views` other views
    on their view in `frame`: the anchor takes that view's measured bearing and the inverse range
    of the current point along it, and the former anchor ray becomes an observation carrying the
    re-anchored view's square-root information. Returns the new window and the re-anchored
    landmark indices (the tracks of a window whose frames were anchored in a later slot)."""
    w = win.copy()
    obs_of = [[] for _ in range(w.L)]
    for f in range(w.F):
        obs_of[w.obs_landmark[f]].append(f)
    moved = []
    for l in range(w.L):
        if w.lm_anchor_frame[l] == frame:
            continue
        here = [f for f in obs_of[l] if w.obs_frame[f] == frame]
        if not here or len(obs_of[l]) - len(here) + 1 < min_views + 1:
            continue
        f = here[0]
        Ra, ta = camera_world(w, w.lm_anchor_frame[l], w.lm_anchor_cam[l])
        b = w.lm_bearing[l] / np.linalg.norm(w.lm_bearing[l])
        world = Ra @ (b / w.lm_rho[l]) + ta
        Rc, tc = camera_world(w, frame, w.obs_cam[f])
        ray = w.obs_bearing[f] / np.linalg.norm(w.obs_bearing[f])
        rng_ = float(ray @ (Rc.T @ (world - tc)))
        if not rng_ > 0:
            continue
        old = (w.lm_anchor_frame[l], w.lm_anchor_cam[l], w.lm_bearing[l].copy(), w.lm_velocity[l].copy(), w.lm_tref[l])
        w.lm_anchor_frame[l], w.lm_anchor_cam[l] = frame, w.obs_cam[f]
        w.lm_bearing[l], w.lm_velocity[l], w.lm_tref[l] = w.obs_bearing[f], w.obs_velocity[f], w.obs_tref[f]
        w.lm_rho[l] = 1.0 / rng_
        w.obs_frame[f], w.obs_cam[f] = old[0], old[1]
        w.obs_bearing[f], w.obs_velocity[f], w.obs_tref[f] = old[2], old[3], old[4]
        moved.append(l)
    return w, np.array(moved, int)


def drop_landmarks(win, drop):
    """The window without the landmarks in `drop` and their observations."""
    w = win.copy()
    keep = np.ones(w.L, bool)
    keep[list(drop)] = False
    new = np.cumsum(keep) - 1
    fk = keep[w.obs_landmark]
    lm = dict(lm_anchor_frame=w.lm_anchor_frame[keep], lm_anchor_cam=w.lm_anchor_cam[keep],
              lm_bearing=w.lm_bearing[keep], lm_velocity=w.lm_velocity[keep], lm_tref=w.lm_tref[keep],
              lm_rho=w.lm_rho[keep], lm_free=None if w.lm_free is None else w.lm_free[keep])
    ob = dict(obs_landmark=new[w.obs_landmark[fk]], obs_frame=w.obs_frame[fk], obs_cam=w.obs_cam[fk],
              obs_bearing=w.obs_bearing[fk], obs_velocity=w.obs_velocity[fk], obs_tref=w.obs_tref[fk],
              obs_sqrt_info=w.obs_sqrt_info[fk])
    kw = {k: v for k, v in w.__dict__.items()}
    kw.update(lm)
    kw.update(ob)
    return abi.Window(**kw)


def quantize_fp32(win):
    """Round every measurement stream to FP32-representable values (both precisions then solve
    the identical problem)."""
    w = win.copy()
    for name in ("lm_bearing", "lm_velocity", "lm_tref", "obs_bearing", "obs_velocity", "obs_tref",
                 "obs_sqrt_info"):
        setattr(w, name, getattr(w, name).astype(np.float32).astype(np.float64))
    return w


CONFIGS = {
    # C1: CPU-reference window
    "c1": dict(K=10, L=2000, n_cams=4, obs_range=(2, 8), seed=20261204),
    # C2: FP64 GPU window at real size, calibration online
    "c2": dict(K=20, L=20000, n_cams=4, obs_range=(4, 8), seed=20261205, rig_free=True, td_free=True),
    # C3: ill-conditioned mixed-precision window
    "c3": dict(K=20, L=40000, n_cams=4, obs_range=(2, 8), seed=20261206, rig_free=True, td_free=True,
               ill_conditioned=True),
}


def batch_config(i, seed_base=20261207):
    """C4: window i of the 256-window batch (sizes drawn from the seeded stream)."""
    rng = np.random.default_rng(seed_base * 1000 + i)
    return dict(K=int(rng.integers(5, 21)), L=int(rng.integers(500, 30001)), n_cams=int(rng.integers(2, 5)),
                obs_range=(2, 10), seed=seed_base * 1000 + i, outlier_fraction=0.05)


def session_config(i, seed_base=20261208):
    """C5: session i of the serving workload (its own 4-camera rig, 8-20 keyframes, 500-30,000
    landmarks, 2-10 observations per landmark)."""
    rng = np.random.default_rng(seed_base * 1000 + i)
    return dict(K=int(rng.integers(8, 21)), L=int(rng.integers(500, 30001)), n_cams=4, obs_range=(2, 10),
                seed=seed_base * 1000 + i)


def stream(K, n_frames, lm_per_frame, n_cams=4, obs_range=(2, 8), span=8, seed=20261211, rig_free=False,
           td_free=False, pixel_sigma=1.0, outlier_fraction=0.0, outlier_px=20.0, speed=1.5, static_bearings=False,
           sigma_g_c=1e-3, sigma_a_c=1e-2, init_pos=0.05, init_rot_deg=1.0, init_rho=0.05, lib=None):
    """A session's stream of keyframes for the sliding-window request loop: the first window
    (frames 0 .. K-1) and the keyframes K .. n_frames-1 that the window appends one per request.

    Every frame gives birth to `lm_per_frame` landmarks: a random pixel of a random camera at a
    depth of 2-20 m, anchored in that (frame, camera) and observed in up to obs_range[1] - 1 more
    (frame, camera) pairs among the following `span` frames that see it. Landmark ids are global
    and increase with the anchor frame (the order in which a backend creates them). Measurements,
    noise, IMU samples and the perturbed initial estimates follow `generate`. Returns (window,
    keyframes) with window.meta holding the truth, the IMU samples, frame times and parameters."""
    lib = lib or abi.load_library()
    rng = np.random.default_rng(seed)
    cam_R, cam_t = make_rig(rng, n_cams)
    R_BR_true = random_small_rotation(rng, np.deg2rad(1.0))
    t_BR_true = rng.uniform(-0.02, 0.02, 3)
    N = n_frames
    frame_t = np.arange(N) * round(KEYFRAME_DT * IMU_RATE) / IMU_RATE
    traj = Trajectory(rng, speed=speed, duration=max(KEYFRAME_DT * (N - 1), 0.5))
    P, R, V, Wb = traj.position(frame_t), traj.rotation(frame_t), traj.velocity(frame_t), traj.body_rate(frame_t)
    R_WC = np.einsum('kij,jl,clm->kcim', R, R_BR_true, cam_R)
    t_WC = P[:, None, :] + np.einsum('kij,cj->kci', R, (R_BR_true @ cam_t.T).T + t_BR_true)
    pts, pairs = [], []  # per landmark: world point, (frame, cam) pairs with the anchor first
    for j in range(N):
        m = lm_per_frame
        c = rng.integers(0, n_cams, m)
        u = rng.uniform(20, WIDTH - 20, m)
        v = rng.uniform(20, HEIGHT - 20, m)
        depth = rng.uniform(2.0, 20.0, m)
        ray = np.stack([(u - CX) / FX, (v - CY) / FY, np.ones(m)], 1)
        ray /= np.linalg.norm(ray, axis=1, keepdims=True)
        world = t_WC[j, c] + np.einsum('nij,nj->ni', R_WC[j, c], ray * depth[:, None])
        later = np.arange(j, min(N, j + span + 1))
        d = world[:, None, None, :] - t_WC[later][None]
        pc = np.einsum('kcji,mkcj->mkci', R_WC[later], d)
        z = pc[..., 2]
        with np.errstate(divide='ignore', invalid='ignore'):
            uu = FX * pc[..., 0] / z + CX
            vv = FY * pc[..., 1] / z + CY
        vis = (z > 0.3) & (uu >= 1) & (uu < WIDTH - 1) & (vv >= 1) & (vv < HEIGHT - 1)
        vis[np.arange(m), 0, c] = False  # the anchor view itself
        vis = vis.reshape(m, -1)         # (frame, camera) pairs in increasing order
        # n - 1 further views per landmark, uniformly among the visible ones: the smallest
        # random keys of the visible pairs
        n = rng.integers(obs_range[0], obs_range[1] + 1, m) - 1
        keys = np.where(vis, rng.random(vis.shape), np.inf)
        rank = np.argsort(np.argsort(keys, axis=1), axis=1)
        pick = vis & (rank < n[:, None])
        for i in np.nonzero(pick.any(1))[0]:
            sel = np.nonzero(pick[i])[0]
            pts.append(world[i])
            pairs.append([(j, int(c[i]))] + [(int(later[q // n_cams]), int(q % n_cams)) for q in sel])
    L = len(pts)
    pts = np.array(pts)
    lm_of = np.concatenate([np.full(len(p), i) for i, p in enumerate(pairs)])
    fk = np.concatenate([[f for f, _ in p] for p in pairs])
    fc = np.concatenate([[cc for _, cc in p] for p in pairs])
    d = pts[lm_of] - t_WC[fk, fc]
    pc = np.einsum('nji,nj->ni', R_WC[fk, fc], d)
    px = np.stack([FX * pc[:, 0] / pc[:, 2] + CX, FY * pc[:, 1] / pc[:, 2] + CY], 1)
    noisy = px + rng.normal(0, pixel_sigma, px.shape)
    if outlier_fraction > 0:
        out = rng.random(len(noisy)) < outlier_fraction
        ang = rng.uniform(0, 2 * np.pi, out.sum())
        noisy[out] += outlier_px * np.stack([np.cos(ang), np.sin(ang)], 1)
    bearing, sqrt_info = abi.pinhole_observations(FX, FY, CX, CY, pixel_sigma, noisy, lib=lib)
    w_W = np.einsum('kij,kj->ki', R, Wb)
    t_BC = (R_BR_true @ cam_t.T).T + t_BR_true
    w_C = np.einsum('kcji,kj->kci', R_WC, w_W)
    v_WC = V[:, None, :] + np.cross(w_W[:, None, :], np.einsum('kij,cj->kci', R, t_BC))
    v_C = np.einsum('kcji,kcj->kci', R_WC, v_WC)
    pdot = -np.cross(w_C[fk, fc], pc) - v_C[fk, fc]
    rn = np.linalg.norm(pc, axis=1)
    bt = pc / rn[:, None]
    bvel = (pdot - bt * (bt * pdot).sum(1, keepdims=True)) / rn[:, None]
    bvel = bvel - bearing * (bearing * bvel).sum(1, keepdims=True)
    if static_bearings:
        bvel = np.zeros_like(bvel)
    first = np.concatenate([[0], np.cumsum([len(p) for p in pairs])[:-1]])
    is_anchor = np.zeros(len(fk), bool)
    is_anchor[first] = True
    a_frame, a_cam = fk[first], fc[first]
    true_rho = 1.0 / rn[first]
    rho0 = true_rho * (1 + rng.uniform(-init_rho, init_rho, L))
    # IMU and the perturbed initial estimate of every frame (frame 0 at truth: the gauge)
    bg_true = rng.normal(0, 1e-3, 3)
    ba_true = rng.normal(0, 1e-2, 3)
    samples, _ = imu_samples(traj, frame_t[0], frame_t[-1], rng, sigma_g_c, sigma_a_c, bg_true, ba_true)
    params = abi.imu_params(sigma_g_c=sigma_g_c, sigma_a_c=sigma_a_c)
    pose = np.zeros((N, 7))
    sb = np.zeros((N, 9))
    for k in range(N):
        Rk = R[k] if k == 0 else random_small_rotation(rng, np.deg2rad(init_rot_deg)) @ R[k]
        pose[k, :3] = P[k] + (0 if k == 0 else rng.uniform(-1, 1, 3) * init_pos)
        pose[k, 3:] = quat_from_matrix(Rk)
        sb[k, :3] = V[k] + (0 if k == 0 else rng.normal(0, 0.02, 3))
    imu = np.zeros((max(N - 1, 0), abi.IMU_STRIDE))
    for e in range(N - 1):
        imu[e] = abi.preintegrate(samples, frame_t[e], frame_t[e + 1], sb[e], params,
                                  gyro_at(samples, frame_t[e]), gyro_at(samples, frame_t[e + 1]), lib=lib)
    rig_true = np.concatenate([t_BR_true, quat_from_matrix(R_BR_true)])
    rig = rig_true.copy()
    if rig_free:
        Rr = random_small_rotation(rng, np.deg2rad(0.3)) @ R_BR_true
        rig = np.concatenate([t_BR_true + rng.uniform(-0.005, 0.005, 3), quat_from_matrix(Rr)])
    flags = (abi.VIB_RIG_FREE if rig_free else 0) | (abi.VIB_TD_FREE if td_free else 0)
    # the first window
    lm_in = np.nonzero(a_frame < K)[0]
    ob_in = np.nonzero(~is_anchor & (fk < K))[0]
    ob_in = ob_in[np.argsort(lm_of[ob_in], kind="stable")]
    sig = np.array([1e-3] * 3 + [1e-3] * 3 + [0.1] * 3 + [0.01] * 3 + [0.1] * 3)
    win = abi.Window(
        frame_id=np.arange(K) + 1000, pose=pose[:K], speed_bias=sb[:K], rig=rig, td=np.array([0.002 if td_free else 0.0]),
        cam_rotation=cam_R, cam_translation=cam_t,
        lm_anchor_frame=a_frame[lm_in], lm_anchor_cam=a_cam[lm_in], lm_bearing=bearing[first[lm_in]],
        lm_velocity=bvel[first[lm_in]], lm_tref=np.zeros(lm_in.size), lm_rho=rho0[lm_in],
        obs_landmark=np.searchsorted(lm_in, lm_of[ob_in]), obs_frame=fk[ob_in], obs_cam=fc[ob_in],
        obs_bearing=bearing[ob_in], obs_velocity=bvel[ob_in], obs_tref=np.zeros(ob_in.size),
        obs_sqrt_info=sqrt_info[ob_in], imu_from=np.arange(K - 1), imu_to=np.arange(1, K), imu_data=imu[:K - 1],
        flags=flags, huber_delta=1.0, min_inverse_depth=1e-8, td_lower=-0.05, td_upper=0.05,
        prior_key_kind=np.array([abi.VIB_KEY_POSE, abi.VIB_KEY_SPEED_BIAS]), prior_key_id=np.array([1000, 1000]),
        prior_ref=np.concatenate([pose[0], sb[0]]), prior_A=np.diag(1.0 / sig), prior_r0=np.zeros(15),
        lm_id=lm_in)
    # the keyframes
    kfs = []
    for j in range(K, N):
        new = np.nonzero(a_frame == j)[0]
        ob = np.nonzero(~is_anchor & (fk == j))[0]
        kfs.append(abi.Keyframe(
            frame_id=1000 + j, pose=pose[j], speed_bias=sb[j], imu_data=imu[j - 1],
            lm_id=new, lm_anchor_cam=a_cam[new], lm_bearing=bearing[first[new]], lm_velocity=bvel[first[new]],
            lm_tref=np.zeros(new.size), lm_rho=rho0[new],
            obs_landmark_id=lm_of[ob], obs_cam=fc[ob], obs_bearing=bearing[ob], obs_velocity=bvel[ob],
            obs_tref=np.zeros(ob.size), obs_sqrt_info=sqrt_info[ob]))
    win.meta = dict(seed=seed, true_pose=np.concatenate([P, np.array([quat_from_matrix(r) for r in R])], 1),
                    true_rho=true_rho, true_rig=rig_true, samples=samples, frame_t=frame_t, imu_params=params,
                    lm_anchor_frame=a_frame, obs=(lm_of, fk, fc, is_anchor))
    return win, kfs
