"""ctypes binding of the FP64 CPU oracle (oracle/build/liboracle.so).

Test infrastructure only: loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg to check (or time, as a reported baseline) the native HIP path.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")

_LIB = None


def build():
    subprocess.run(["make", "-B", "-C", HERE], check=True, stdout=subprocess.DEVNULL)


def load(abi):
    """Open liboracle.so; `abi` is the vibpy.abi module (struct definitions)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int32)
    lp = C.POINTER(C.c_int64)
    W = C.POINTER(abi.VibWindow)
    lib.oracle_visual_factor.argtypes = [W, C.c_int, dp, dp, dp]
    lib.oracle_visual_residual_reference.argtypes = [W, C.c_int, dp]
    lib.oracle_imu_factor.argtypes = [W, C.c_int, dp, dp]
    lib.oracle_linearize.argtypes = [W, dp, dp, dp, dp, dp, ip]
    lib.oracle_solve_step.argtypes = [W, C.c_double, dp, dp, dp, dp]
    lib.oracle_cost.argtypes = [W, dp]
    lib.oracle_solve_step_threaded.argtypes = [W, C.c_double, C.c_int, dp, dp]
    lib.oracle_time_steps.argtypes = [W, C.c_double, C.c_int, C.c_int, dp]
    lib.oracle_time_windows.argtypes = [W, C.c_int, C.c_double, C.c_int, C.c_int, dp]
    lib.oracle_solve.argtypes = [W, C.POINTER(abi.VibSolveOptions), C.POINTER(abi.VibSolveReport)]
    lib.oracle_solve_latest_two.argtypes = [W, C.POINTER(abi.VibSolveOptions), C.c_int, dp, ip,
                                            C.POINTER(abi.VibSolveReport)]
    lib.oracle_marginalize_old.argtypes = [W, ip, dp, dp, dp, dp, ip, lp, dp, dp]
    lib.oracle_marginal_system.argtypes = [W, ip, dp, dp, ip]
    lib.oracle_marginalize_second_new.argtypes = [W, dp, C.c_int, C.c_double, C.c_double,
                                                  C.POINTER(abi.VibImuParams), ip, ip, W]
    lib.oracle_second_new_system.argtypes = [W, dp, C.c_int, C.c_double, C.c_double, C.POINTER(abi.VibImuParams),
                                             ip, dp, dp, ip]
    lib.oracle_preintegrate.argtypes = [dp, C.c_int, C.c_double, C.c_double, dp, C.POINTER(abi.VibImuParams), dp, dp, dp]
    lib.oracle_preintegrate_cov.argtypes = [dp, C.c_int, C.c_double, C.c_double, dp, C.POINTER(abi.VibImuParams), dp,
                                            dp, dp, dp]
    lib.oracle_pinhole_observation.argtypes = [C.c_double] * 7 + [dp, dp]
    lib.oracle_eigen_sym.argtypes = [dp, C.c_int, dp, dp]
    lib.oracle_pose_plus.argtypes = [dp, dp, dp]
    lib.oracle_pose_minus.argtypes = [dp, dp, dp]
    for f in ("oracle_pose_plus", "oracle_pose_minus"):
        getattr(lib, f).restype = None
    _LIB = lib
    return lib


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    def __init__(self, abi):
        self.abi = abi
        self.lib = load(abi)

    def visual_factor(self, win, f):
        s = win.cstruct()
        r = np.zeros(2)
        J = np.zeros((2, 20))
        amb = np.zeros((2, 23))
        ok = self.lib.oracle_visual_factor(C.byref(s), f, _p(r), _p(J), _p(amb))
        return bool(ok), r, J, amb

    def visual_residual_reference(self, win, f):
        s = win.cstruct()
        r = np.zeros(2)
        ok = self.lib.oracle_visual_residual_reference(C.byref(s), f, _p(r))
        return bool(ok), r

    def imu_factor(self, win, e):
        s = win.cstruct()
        r = np.zeros(15)
        J = np.zeros((4, 15, 9))
        ok = self.lib.oracle_imu_factor(C.byref(s), e, _p(r), _p(J))
        return bool(ok), r, J

    def linearize(self, win):
        s = win.cstruct()
        ne = C.c_int32()
        nr = self.lib.oracle_linearize(C.byref(s), None, None, None, None, None, C.byref(ne))
        ne = ne.value
        H = np.zeros((nr, nr))
        cross = np.zeros((ne, nr))
        diag = np.zeros(ne)
        g = np.zeros(nr + ne)
        cost = C.c_double()
        st = self.lib.oracle_linearize(C.byref(s), _p(H), _p(cross), _p(diag), _p(g), C.byref(cost), None)
        if st < 0:
            raise RuntimeError("oracle linearization invalid")
        return dict(H=H, cross=cross, diag=diag, g=g, cost=cost.value, nr=nr, ne=ne)

    def solve_step(self, win, lam):
        lay = win.layout()
        nr, ne = lay["nr"], lay["ne"]
        delta = np.zeros(nr + ne)
        S = np.zeros((nr, nr))
        rhs = np.zeros(nr)
        m = np.zeros(5)
        s = win.cstruct()
        st = self.lib.oracle_solve_step(C.byref(s), lam, _p(delta), _p(S), _p(rhs), _p(m))
        return st, dict(delta=delta, S=S, rhs=rhs, cost=m[0], predicted=m[1], trial_cost=m[2], max_increment=m[3],
                        trial_valid=bool(m[4]))

    def solve_step_threaded(self, win, lam, threads):
        """Block-sparse step on `threads` host threads: (status, delta, metrics dict)."""
        lay = win.layout()
        delta = np.zeros(lay["nr"] + lay["ne"])
        m = np.zeros(5)
        s = win.cstruct()
        st = self.lib.oracle_solve_step_threaded(C.byref(s), lam, threads, _p(delta), _p(m))
        return st, delta, dict(cost=m[0], predicted=m[1], trial_cost=m[2], max_increment=m[3], trial_valid=bool(m[4]))

    def time_steps(self, win, lam, threads, steps):
        """Wall milliseconds of `steps` block-sparse steps of one window on `threads` threads."""
        ms = np.zeros(steps)
        s = win.cstruct()
        st = self.lib.oracle_time_steps(C.byref(s), lam, threads, steps, _p(ms))
        return st, ms

    def time_windows(self, wins, lam, threads, rounds=1):
        """Wall seconds of `rounds` single-threaded steps of every window, windows spread over
        `threads` host threads."""
        structs = [w.cstruct() for w in wins]
        arr = (self.abi.VibWindow * len(structs))(*structs)
        sec = C.c_double()
        st = self.lib.oracle_time_windows(arr, len(structs), lam, threads, rounds, C.byref(sec))
        return st, sec.value

    def cost(self, win):
        s = win.cstruct()
        c = C.c_double()
        st = self.lib.oracle_cost(C.byref(s), C.byref(c))
        return st, c.value

    def solve(self, win, opts):
        """Solves in place (win's state arrays receive the result)."""
        s = win.cstruct()
        rep = self.abi.VibSolveReport()
        st = self.lib.oracle_solve(C.byref(s), C.byref(opts), C.byref(rep))
        return st, rep

    def solve_latest_two(self, win, opts, gate=False):
        """Two-frame solve in place (poses and velocities of the two newest slots).
        Returns status, report, NIS {old before, old after, new before, new after}, support."""
        s = win.cstruct()
        rep = self.abi.VibSolveReport()
        nis = np.zeros(4)
        sup = C.c_int32()
        st = self.lib.oracle_solve_latest_two(C.byref(s), C.byref(opts), 1 if gate else 0, _p(nis), C.byref(sup),
                                              C.byref(rep))
        return st, rep, nis, sup.value

    def marginalize_old(self, win):
        s = win.cstruct()
        sizes = np.zeros(5, np.int32)
        ip = C.POINTER(C.c_int32)
        st = self.lib.oracle_marginalize_old(C.byref(s), sizes.ctypes.data_as(ip), None, None, None, None, None, None,
                                             None, None)
        if st != 0:
            return st, None
        n, rank, blocks, reflen = int(sizes[0]), int(sizes[1]), int(sizes[2]), int(sizes[3])
        H = np.zeros((n, n))
        b = np.zeros(n)
        A = np.zeros((rank, n))
        r0 = np.zeros(rank)
        kind = np.zeros(blocks, np.int32)
        kid = np.zeros(blocks, np.int64)
        ref = np.zeros(reflen)
        cut = C.c_double()
        st = self.lib.oracle_marginalize_old(C.byref(s), sizes.ctypes.data_as(ip), _p(H), _p(b), _p(A), _p(r0),
                                             kind.ctypes.data_as(ip), kid.ctypes.data_as(C.POINTER(C.c_int64)), _p(ref),
                                             C.byref(cut))
        return st, dict(H=H, b=b, A=A, r0=r0, kind=kind, id=kid, ref=ref, cutoff=cut.value,
                        landmarks=int(sizes[4]))

    def marginal_system(self, win):
        """Whole system (H, b) of the factors an OLD marginalization consumes, eliminated mask."""
        s = win.cstruct()
        n = C.c_int32()
        st = self.lib.oracle_marginal_system(C.byref(s), C.byref(n), None, None, None)
        if st != 0:
            return st, None
        H = np.zeros((n.value, n.value))
        b = np.zeros(n.value)
        mask = np.zeros(n.value, np.int32)
        st = self.lib.oracle_marginal_system(C.byref(s), C.byref(n), _p(H), _p(b), mask.ctypes.data_as(C.POINTER(C.c_int32)))
        return st, dict(H=H, b=b, mask=mask.astype(bool))

    def marginalize_second_new(self, win, samples, t0, t1, params):
        """SECOND_NEW slide of the second-newest slot. Returns status, report dict and the window
        after the slide (None on failure)."""
        abi = self.abi
        nr = win.layout()["nr"]
        cap = nr + 16
        blocks = 2 * win.K + 2
        buf = abi.window_buffers(win.K - 1, int(win.cam_rotation.shape[0]), win.L, win.F, win.E, cap, cap, blocks,
                                 7 * blocks)
        out = buf.cstruct()
        s = win.cstruct()
        sizes = np.array([win.L, win.F, win.E, cap * cap, cap, blocks, 0, 0], np.int32)
        rep = np.zeros(8, np.int32)
        smp = None if samples is None else np.ascontiguousarray(samples, np.float64)
        ipp = C.POINTER(C.c_int32)
        st = self.lib.oracle_marginalize_second_new(C.byref(s), None if smp is None else _p(smp),
                                                    0 if smp is None else smp.shape[0], t0, t1,
                                                    None if params is None else C.byref(params),
                                                    sizes.ctypes.data_as(ipp), rep.ctypes.data_as(ipp), C.byref(out))
        report = dict(status=int(rep[0]), prior_reused=bool(rep[1]), imu_mode=int(rep[2]), reset=int(rep[3]),
                      demoted=int(rep[4]), retired=int(rep[5]), dropped=int(rep[6]), rank=int(rep[7]))
        if st != 0:
            return st, report, None
        K, L, F, E, rows, cols, nb, reflen = (int(v) for v in sizes)
        return st, report, abi.trimmed_window(buf, out, K, L, F, E, rows, cols, nb, reflen)

    def second_new_system(self, win, samples, t0, t1, params):
        """Whole system (H, b) of the factors a SECOND_NEW slide consumes, eliminated mask (None
        when the prior is reused)."""
        s = win.cstruct()
        smp = None if samples is None else np.ascontiguousarray(samples, np.float64)
        n = C.c_int32()
        args = (C.byref(s), None if smp is None else _p(smp), 0 if smp is None else smp.shape[0], t0, t1,
                None if params is None else C.byref(params))
        st = self.lib.oracle_second_new_system(*args, C.byref(n), None, None, None)
        if st != 0 or n.value == 0:
            return st, None
        H = np.zeros((n.value, n.value))
        b = np.zeros(n.value)
        mask = np.zeros(n.value, np.int32)
        st = self.lib.oracle_second_new_system(*args, C.byref(n), _p(H), _p(b), mask.ctypes.data_as(C.POINTER(C.c_int32)))
        return st, dict(H=H, b=b, mask=mask.astype(bool))

    def preintegrate(self, samples, t0, t1, sb, params, g0, g1):
        out = np.zeros(self.abi.IMU_STRIDE)
        s = np.ascontiguousarray(samples, np.float64)
        sb = np.ascontiguousarray(sb, np.float64)
        g0 = np.ascontiguousarray(g0, np.float64)
        g1 = np.ascontiguousarray(g1, np.float64)
        st = self.lib.oracle_preintegrate(_p(s), s.shape[0], t0, t1, _p(sb), C.byref(params), _p(g0), _p(g1), _p(out))
        return st, out

    def preintegrate_cov(self, samples, t0, t1, sb, params, g0, g1):
        """Preintegration record and the 15x15 covariance it whitens."""
        out = np.zeros(self.abi.IMU_STRIDE)
        P = np.zeros((15, 15))
        s = np.ascontiguousarray(samples, np.float64)
        sb = np.ascontiguousarray(sb, np.float64)
        g0 = np.ascontiguousarray(g0, np.float64)
        g1 = np.ascontiguousarray(g1, np.float64)
        st = self.lib.oracle_preintegrate_cov(_p(s), s.shape[0], t0, t1, _p(sb), C.byref(params), _p(g0), _p(g1),
                                              _p(out), _p(P))
        return st, out, P

    def pinhole_observation(self, fx, fy, cx, cy, u, v, sigma):
        b = np.zeros(3)
        s = np.zeros(3)
        st = self.lib.oracle_pinhole_observation(fx, fy, cx, cy, u, v, sigma, _p(b), _p(s))
        return st, b, s

    def eigen(self, A):
        A = np.ascontiguousarray(A, np.float64)
        n = A.shape[0]
        ev = np.zeros(n)
        V = np.zeros((n, n))
        st = self.lib.oracle_eigen_sym(_p(A), n, _p(ev), _p(V))
        return st, ev, V

    def pose_plus(self, x, d):
        x = np.ascontiguousarray(x, np.float64)
        d = np.ascontiguousarray(d, np.float64)
        y = np.zeros(7)
        self.lib.oracle_pose_plus(_p(x), _p(d), _p(y))
        return y

    def pose_minus(self, x, ref):
        x = np.ascontiguousarray(x, np.float64)
        ref = np.ascontiguousarray(ref, np.float64)
        d = np.zeros(6)
        self.lib.oracle_pose_minus(_p(x), _p(ref), _p(d))
        return d
