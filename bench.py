#!/usr/bin/env python3
"""Schur window-step benchmark on one MI355X.

A step is one LM iteration of one window through the native HIP path: linearize every visual,
IMU and prior factor, eliminate the landmarks, factor and solve the damped reduced camera system,
back-substitute, evaluate the trial cost and take the accept/reject decision (continuous mode:
every iteration relinearizes, so each timed step carries the full work of an accepted step).

Default workload: K=20 keyframes, 20,000 landmarks, ~117k observations, 4 pinhole cameras,
rig extrinsic and time offset estimated online, FP64 (seeded synthetic window). The JSON line
carries the window-step throughput, p50/p99 step latency from device events, the roofline of the
dominant kernel (per-kernel durations from events on the library stream) and the oracle CPU
baseline timed on this host.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cgmli-slam_amd"))

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# refinement passes of every mixed-precision line: the count at which the accuracy line meets
# the metric's bound (dx relative error <= 1e-6 against plain FP64)
MIXED_REFINEMENT = 2
SERVE_SUSTAINED_RATE = 300.0  # requests/s of the serving line's second schedule (below the loop's capacity)


def algorithmic_bytes(win, precision="fp64"):
    """Compulsory HBM traffic of one window step (per-unit figures: 88 B per observation record
    and 76 B per landmark record in FP64, 48/48 B in mixed; RCS upper triangle written and read;
    prior square root read; frozen IMU records; states; step vector)."""
    lay = win.layout()
    nr = lay["nr"]
    b_obs, b_lm = (88, 76) if precision == "fp64" else (48, 48)
    m, n = (win.prior_A.shape if win.prior_A is not None else (0, 0))
    return (win.F * b_obs + win.L * b_lm + 8 * nr * (nr + 1) + 8 * m * n + 2408 * (win.K - 1) +
            8 * (16 * win.K + 8) + 16 * nr)


def kernel_bytes(name, win):
    """Algorithmic bytes of one launch of a chunk kernel over the whole window."""
    if name == "k_lin_visual":
        return win.F * 88 + win.L * 68          # observation records + landmark anchor records and rho
    if name == "k_trial_cost":
        return win.F * 88 + win.L * 68
    if name in ("k_schur", "k_backsub"):
        return win.L * 76
    return None


def kernel_flops(name, win):
    nr = win.layout()["nr"]
    if name == "k_cholesky":
        return nr ** 3 / 3.0 + 4.0 * nr * nr
    return None


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` from the committed FETCH_SIZE / WRITE_SIZE profile of this
    workload (counters scaled by a copy of known size), or None when no profile exists."""
    path = os.path.join(ROOT, "profiles", f"schur_step_{workload}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        prof = json.load(f)
    k = prof.get("kernels", {}).get(kernel)
    return None if k is None else k["read_bytes"] + k["write_bytes"]


def _batch_window(i):
    from vibpy import synth
    return synth.generate(**synth.batch_config(i))


def batch_windows(n, procs):
    """The n windows of the batched workload (sizes drawn per window from the seeded stream),
    generated in `procs` spawned worker processes (no GPU state is inherited)."""
    if procs <= 1:  # in this process (profilers that follow child processes)
        return [_batch_window(i) for i in range(n)]
    import multiprocessing as mp
    with mp.get_context("spawn").Pool(procs) as pool:
        return pool.map(_batch_window, range(n))


def batched_line(ctx, abi, wins, steps, warmup):
    """Batched throughput: all windows in one graph-captured launch sequence per LM iteration
    (FP64, then mixed precision), against the same windows solved one after another in FP64
    (one window per launch sequence)."""
    slots = list(range(1000, 1000 + len(wins)))
    for s, w in zip(slots, wins):
        ctx.upload(s, w)
    res, kern, tune = {}, {}, {}
    for name, prec, refine in (("fp64", abi.VIB_FP64, 0), ("mixed", abi.VIB_MIXED, MIXED_REFINEMENT)):
        opts = abi.options(max_iterations=1, precision=prec, refinement_steps=refine, flags=abi.VIB_SOLVE_CONTINUOUS)
        # work-list shapes of this batch from the autotuner (kept by the context for the timed runs)
        tune[name] = ctx.autotune(slots, opts, 2)
        ctx.bench(slots, opts, warmup, kernels=False)
        ctx.synchronize()
        it_ms, _ = ctx.bench(slots, opts, steps, kernels=False)
        ctx.synchronize()
        res[name] = it_ms
        # per-kernel device time of the batched launches alone (events around every launch)
        _, kern[name] = ctx.bench(slots, opts, 2, kernels=True)
        ctx.synchronize()
    opts = abi.options(max_iterations=1, flags=abi.VIB_SOLVE_CONTINUOUS)
    seq = []
    for s in slots:
        ms, _ = ctx.bench([s], opts, 3, kernels=False)
        seq.append(float(np.median(ms)))
    for s in slots:
        ctx.release(s)
    batch_ms = float(np.median(res["fp64"]))
    mixed_ms = float(np.median(res["mixed"]))
    seq_ms = float(np.sum(seq))
    return dict(windows=len(wins), factors=int(sum(w.F for w in wins)), landmarks=int(sum(w.L for w in wins)),
                value=len(wins) / (batch_ms * 1e-3), unit="window-steps/s", ms_per_batched_step=batch_ms,
                p99_ms=float(np.percentile(res["fp64"], 99)), mixed_value=len(wins) / (mixed_ms * 1e-3),
                mixed_ms_per_batched_step=mixed_ms, mixed_p99_ms=float(np.percentile(res["mixed"], 99)),
                sequential_value=len(wins) / (seq_ms * 1e-3), sequential_ms=seq_ms,
                speedup_vs_sequential=seq_ms / batch_ms, mixed_speedup_vs_sequential_fp64=seq_ms / mixed_ms,
                mixed_refinement_steps=MIXED_REFINEMENT, kernel_ms=kern["fp64"], mixed_kernel_ms=kern["mixed"],
                autotune=tune)


def step_error(ctx, abi, win, slot, lam=1e-4, refinement=2):
    """One LM step of `win` from its initial state at identical lambda in plain FP64 and in mixed
    precision with `refinement` passes: ||dx_mixed - dx_fp64|| / ||dx_fp64|| over the full
    tangent step (states and inverse ranges), and the relative differences of the
    linearization-point and trial costs."""
    lay = win.layout()
    n = lay["nr"] + lay["ne"]
    out = {}
    for name, prec, refine in (("fp64", abi.VIB_FP64, 0), ("mixed", abi.VIB_MIXED, refinement)):
        ctx.upload(slot, win)
        ctx.check(ctx.linearize(slot, prec), "vib_linearize")
        st, rep = ctx.solve_step(slot, lam, prec, refine)
        ctx.check(st, "vib_solve_step")
        out[name] = (ctx.read_step(slot, n), rep)
    ctx.release(slot)
    d64, r64 = out["fp64"]
    d32, r32 = out["mixed"]
    return dict(dx_rel_err=float(np.linalg.norm(d32 - d64) / np.linalg.norm(d64)),
                cost_rel_err=abs(r32.cost - r64.cost) / r64.cost,
                trial_cost_rel_err=abs(r32.trial_cost - r64.trial_cost) / r64.trial_cost,
                lam=lam, refinement_steps=refinement)


def _session_stream(args):
    """Session i of the serving workload: its first window and `n` keyframes to append. A request
    solves a window of K frames (K 8-20): K - 1 resident frames and the appended keyframe; the
    slide brings it back to K - 1."""
    i, n = args
    from vibpy import synth
    cfg = synth.session_config(i)
    K = cfg["K"]
    return synth.stream(K - 1, K + n, max(1, cfg["L"] // K), n_cams=4, obs_range=(2, 10), seed=cfg["seed"])


def serving_line(ctx, abi, n_sessions, rate, seconds, procs, budget_gb, precision, refinement, workers=1):
    """Serving workload: n_sessions sliding-window sessions; Poisson requests at `rate` per second
    for `seconds`; each request appends its session's next keyframe, runs 1-3 LM iterations
    batched with the other running requests of its size bucket and slides OLD back to the window
    size, the prior staying resident (the library's request server, vib_server_*)."""
    import multiprocessing as mp
    from vibpy import serve
    reqs = serve.poisson_requests(n_sessions, rate, seconds, np.random.default_rng(20261208))
    # a second schedule below the loop's capacity: latency without the backlog of the offered load
    low = serve.poisson_requests(n_sessions, SERVE_SUSTAINED_RATE, 2.0, np.random.default_rng(20261209))
    need = np.maximum(np.bincount([r["session"] for r in reqs], minlength=n_sessions),
                      np.bincount([r["session"] for r in low], minlength=n_sessions))
    with mp.get_context("spawn").Pool(procs) as pool:
        streams = pool.map(_session_stream, [(i, int(need[i])) for i in range(n_sessions)])
    wins = [w for w, _ in streams]
    kfs = [k for _, k in streams]
    def serve_once(schedule):
        srv = serve.Server(ctx, wins, int(budget_gb * (1 << 30)), precision=precision, refinement=refinement,
                           full_iterations=True, workers=workers)
        try:
            srv.preload()
            ctx.synchronize()
            t0 = time.perf_counter()
            done = srv.run(schedule, keyframes=kfs)
            wall = time.perf_counter() - t0
            return done, wall, srv.stats()
        finally:
            srv.close()

    done, wall, st = serve_once(reqs)
    ldone, lwall, _ = serve_once(low)
    out = serve.summary(done, wall)
    sustained = serve.summary(ldone, lwall)
    out["sustained"] = dict(arrival_rate=SERVE_SUSTAINED_RATE, offered_seconds=2.0, requests=sustained["requests"],
                            requests_per_s=sustained["requests_per_s"], steps_per_s=sustained["steps_per_s"],
                            p50_ms=sustained["p50_ms"], p99_ms=sustained["p99_ms"], failed=sustained["failed"])
    out.update(sessions=n_sessions, workers=workers, arrival_rate=rate, offered_seconds=seconds, hbm_budget_gb=budget_gb,
               uploads=st["uploads"], evictions=st["evictions"], launches=st["launches"], slides=st["slides"],
               slides_per_s=st["slides"] / wall, appends=st["appends"],
               phase_seconds=dict(admit=st["admit_seconds"], iterate=st["iterate_seconds"], slide=st["slide_seconds"]),
               unmatched_observations=int(sum(r["unmatched"] for r in done)),
               window_landmarks=int(sum(w.L for w in wins)), window_factors=int(sum(w.F for w in wins)),
               precision="mixed" if precision == abi.VIB_MIXED else "fp64", refinement_steps=refinement,
               workload=f"{n_sessions} sessions (own 4-camera rig, K 8-20 keyframes solved per request, 500-30000 landmarks per window, "
                        f"2-10 observations per landmark); Poisson {rate:g} requests/s for {seconds:g} s; each request "
                        "appends the session's next keyframe, runs 1-3 full LM iterations (relinearized, no early "
                        "stop) batched by size bucket, then slides OLD back to the window size (prior resident); "
                        "latency from arrival to the end of the slide")
    return out


def preintegration_line(ctx, abi, count, reps=5, host_seconds=2.0):
    """Device preintegration of `count` keyframe edges of 0.1 s (20 samples at 200 Hz each, every
    edge at its own bias linearization point) in one launch, against the native host routine
    (vib_preintegrate, one thread) on a bounded sample of the same edges."""
    from vibpy import synth
    rng = np.random.default_rng(20261210)
    rate, span = 200.0, 0.1
    n = int(round(count * span * rate)) + 2
    t = np.arange(n) / rate
    s = np.empty((n, 7))
    s[:, 0] = t
    for k in range(3):
        s[:, 1 + k] = 0.3 * np.sin(1.7 * t + k) + 0.05 * rng.standard_normal(n)
        s[:, 4 + k] = (9.81 if k == 2 else 0.0) + 0.8 * np.cos(2.3 * t + k) + 0.02 * rng.standard_normal(n)
    t0 = np.arange(count) * span + 0.5 / rate
    t1 = t0 + span
    b = np.floor(t0 * rate).astype(np.int32)
    e = (np.ceil(t1 * rate) + 1).astype(np.int32)
    sb = np.zeros((count, 9))
    sb[:, 3:6] = rng.normal(0, 1e-3, (count, 3))
    sb[:, 6:9] = rng.normal(0, 1e-2, (count, 3))
    g0 = np.array([synth.gyro_at(s, a) for a in t0])
    g1 = np.array([synth.gyro_at(s, a) for a in t1])
    p = abi.imu_params()
    ms = []
    for _ in range(reps):
        out, st, k = ctx.preintegrate_batch(s, b, e, t0, t1, sb, p, g0, g1, timed=True)
        if (st != abi.VIB_OK).any():
            raise RuntimeError("device preintegration rejected an interval")
        ms.append(k)
    kms = float(np.median(ms))
    one = []  # latency of a launch with a single edge
    for _ in range(reps):
        _, st, k = ctx.preintegrate_batch(s, b[:1], e[:1], t0[:1], t1[:1], sb[:1], p, g0[:1], g1[:1], timed=True)
        one.append(k)
    m = 0
    h0 = time.perf_counter()
    while m < count:
        ref = abi.preintegrate(s[b[m]:e[m]], t0[m], t1[m], sb[m], p, g0[m], g1[m])
        m += 1
        if time.perf_counter() - h0 >= host_seconds:
            break
    hs = time.perf_counter() - h0
    dev = float(np.max(np.abs(out[m - 1, :76] - ref[:76]) / (np.abs(ref[:76]) + 1e-12)))
    return dict(value=count / (kms * 1e-3), unit="intervals/s", kernel_ms=kms, intervals=count,
                single_interval_kernel_ms=float(np.median(one)),
                samples_per_interval=int(np.median(e - b)), host_value=m / hs, host_sample=f"{m} intervals, 1 thread",
                max_rel_dev_vs_host=dev,
                workload=f"{count} keyframe edges of 0.1 s at 200 Hz, one device launch (kernel time, inputs "
                         "resident), host vib_preintegrate one thread")


def gate_line(ctx, abi, win, slot=0, reps=10):
    """Statistical gate of the resident window (every visual factor's NIS, decisions, residual
    statistics with exact quantiles), wall time of the host call."""
    ms = []
    for _ in range(reps):
        t0 = time.perf_counter()
        st, rep, _ = ctx.nis_gate(slot, abi.VISUAL_NIS_THRESHOLD, apply=False)
        ms.append(1e3 * (time.perf_counter() - t0))
        if st != abi.VIB_OK:
            raise RuntimeError(f"gate failed: {st}")
    m = float(np.median(ms))
    ev = rep.evaluated
    return dict(value=win.F / (m * 1e-3), unit="factors/s", wall_ms=m, factors=win.F, over_gate=int(ev.over_gate),
                median_nis=ev.median_nis, p95_nis=ev.p95_nis,
                workload="the timed window at its state after the bench steps: NIS of every visual factor against "
                         "9.21, decisions and the evaluated/admitted statistics; host call wall time (descriptor "
                         "upload, allocation, three kernels, download)")


def whitening_line(ctx, abi, count, reps=5):
    """Device whitening of `count` pixel observations over 4 pinhole cameras (one sigma), kernel
    time with inputs resident, against the native host routine (vib_pinhole_observations, one
    thread) on the same pixels."""
    rng = np.random.default_rng(20261211)
    intr = np.array([[458.6, 457.3, 367.2, 248.4, 752, 480], [700.0, 700.0, 640.0, 360.0, 1280, 720],
                     [320.5, 318.0, 315.0, 242.0, 640, 480], [1000.0, 990.0, 800.0, 600.0, 1600, 1200]])
    cam = rng.integers(0, 4, count).astype(np.int32)
    px = np.stack([rng.uniform(0, intr[cam, 4]), rng.uniform(0, intr[cam, 5])], 1)
    ms = []
    for _ in range(reps):
        b, s, st, k = ctx.whiten(px, intr, cam, 1.0, timed=True)
        if (st != 0).any():
            raise RuntimeError("device whitening rejected an observation")
        ms.append(k)
    kms = float(np.median(ms))
    h0 = time.perf_counter()
    for c in range(4):
        abi.pinhole_observations(*intr[c, :4], 1.0, px[cam == c])
    hs = time.perf_counter() - h0
    per = 16 + 4 + 48 + 4  # pixel, camera index, bearing + square-root information, status
    return dict(value=count / (kms * 1e-3), unit="observations/s", kernel_ms=kms, observations=count,
                bytes_per_observation=per, achieved_gbps=count * per / (kms * 1e-3) / 1e9, peak_gbps=8000.0,
                host_value=count / hs, host_sample=f"{count} observations, 1 thread",
                workload=f"{count} pixel observations over 4 undistorted pinhole cameras, one launch (kernel time, "
                         "inputs resident); host vib_pinhole_observations one thread")


def slide_line(ctx, abi, win, reps=3):
    """Window slides of the timed window: MARGIN_OLD (slot 0 and its landmarks eliminated into the
    prior) and SECOND_NEW (the redundant second-newest slot removed, its IMU edges merged and
    re-integrated, tracks re-anchored); wall time of the host call, upload excluded."""
    out = {}
    m = win.meta
    t = m["frame_t"]
    for name, policy in (("old", abi.VIB_MARGIN_OLD), ("second_new", abi.VIB_MARGIN_SECOND_NEW)):
        ms = []
        for _ in range(reps):
            ctx.upload(0, win)
            ctx.synchronize()
            t0 = time.perf_counter()
            if policy == abi.VIB_MARGIN_OLD:
                st, rep = ctx.marginalize(0, policy)
            else:
                st, rep = ctx.marginalize(0, policy, m["samples"], float(t[win.K - 3]), float(t[win.K - 1]),
                                          m["imu_params"])
            ms.append(1e3 * (time.perf_counter() - t0))
            if st != abi.VIB_OK:
                raise RuntimeError(f"slide {name} failed: {st}")
        out[name + "_ms"] = float(np.median(ms))
        out[name + "_eliminated_landmarks"] = int(rep.eliminated_landmarks)
    out["workload"] = ("the timed window uploaded afresh, then one slide; wall time of vib_marginalize (topology "
                       "selection, marginal system and the slid window as device edits of the resident window, "
                       "the prior's eigen square root on the device; frame-level state and prior through the host)")
    return out


def two_frame_line(ctx, abi, reps=12):
    """Two-frame solve latency (optimizeLatestTwo, BackendOptimizer.cpp:476-590; budget <= 3
    iterations, <= 8 ms, :549-553) where the reference runs it: right after a new keyframe joins
    a solved window. A C2-sized stream (4 cameras, rig and time offset online, about 1000
    landmarks born per frame, 20 frames with the new one): per sample the next keyframe is
    appended, vib_solve_latest_two runs (timed: the local window built on the device, its IMU
    bridge checked, at most 3 iterations, the write-back), then the window solves 2 iterations and
    slides OLD back to 19 frames (untimed)."""
    from vibpy import synth
    win, kfs = synth.stream(19, 19 + reps + 2, 1000, n_cams=4, obs_range=(4, 8), seed=20261212, rig_free=True,
                            td_free=True)
    ctx.upload(0, win)
    ctx.check(ctx.solve(0, abi.options(max_iterations=5))[0], "vib_solve")
    ms, its = [], []
    for i, kf in enumerate(kfs[:reps + 2]):
        ctx.append(0, kf)
        ctx.synchronize()
        t0 = time.perf_counter()
        st, rep = ctx.solve_latest_two(0, abi.options(max_iterations=3))
        dt = 1e3 * (time.perf_counter() - t0)
        if st != abi.VIB_OK:
            raise RuntimeError(f"two-frame solve failed: {abi.STATUS_NAMES.get(st, st)}")
        if i >= 2:
            ms.append(dt)
            its.append(rep.solve.iterations)
        ctx.check(ctx.solve(0, abi.options(max_iterations=2))[0], "vib_solve")
        ctx.check(ctx.marginalize(0, abi.VIB_MARGIN_OLD)[0], "vib_marginalize")
    shape = ctx.shape(0)
    ctx.release(0)
    return dict(p50_ms=float(np.percentile(ms, 50)), p99_ms=float(np.percentile(ms, 99)), budget_ms=8.0,
                iterations=int(np.median(its)), support_landmarks=int(rep.support_landmarks), calls=reps,
                window=dict(frames=20, landmarks=int(shape[1]), observations=int(shape[2])),
                workload="C2-sized keyframe stream (20 frames with the new one, rig+td online); each call right "
                         "after the next keyframe's append; vib_solve_latest_two, at most 3 iterations, FP64; wall time")


def host_info():
    """Host cores usable by this process, CPU model and NUMA layout (printed beside the CPU line)."""
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    nodes = []
    base = "/sys/devices/system/node"
    if os.path.isdir(base):
        for d in sorted(os.listdir(base)):
            if d.startswith("node") and d[4:].isdigit():
                try:
                    with open(os.path.join(base, d, "cpulist")) as f:
                        nodes.append(f"{d}: cpus {f.read().strip()}")
                except OSError:
                    pass
    return dict(nproc=os.cpu_count(), usable_cpus=usable, cpu_model=model, numa=nodes)


def cpu_baseline(win, seconds=12.0, threads=None, bwins=None):
    """The oracle's block-sparse FP64 LM step (per-landmark Schur on the landmark's support
    columns, std::thread over landmark ranges) timed on this host: all usable cores and one
    thread, p50/p99 per step, on a bounded sample (about `seconds` in total). With `bwins`, also
    the batched workload: every window stepped once per round, whole windows spread over the
    threads."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py
    from vibpy import abi
    o = oracle_py.Oracle(abi)
    info = host_info()
    # one GPU's share of the host is 16 cores, whatever the affinity mask shows
    threads = threads or max(1, min(info["usable_cpus"], 16))
    print(f"[bench] cpu baseline: {info['cpu_model']}, {info['usable_cpus']} usable of {info['nproc']} cpus, "
          f"numa {info['numa']}", file=sys.stderr, flush=True)
    st, ms1 = o.time_steps(win, 1e-4, 1, 2)       # probe one step on one thread to size the samples
    if st != 0:
        raise RuntimeError(f"oracle step failed: {st}")
    per = float(np.median(ms1)) * 1e-3
    n1 = int(max(5, min(200, 0.35 * seconds / per)))
    n_all = int(max(10, min(2000, 0.35 * seconds / per * threads / 2)))
    st, ms_all = o.time_steps(win, 1e-4, threads, n_all)
    if st != 0:
        raise RuntimeError(f"oracle step failed: {st}")
    st, ms_one = o.time_steps(win, 1e-4, 1, n1)
    if st != 0:
        raise RuntimeError(f"oracle step failed: {st}")
    out = dict(value=1e3 / float(np.mean(ms_all)), unit="window-steps/s", cores=threads, kind="port",
               p50_ms=float(np.percentile(ms_all, 50)), p99_ms=float(np.percentile(ms_all, 99)),
               one_thread=dict(value=1e3 / float(np.mean(ms_one)), p50_ms=float(np.percentile(ms_one, 50)),
                               p99_ms=float(np.percentile(ms_one, 99)), steps=n1),
               host=info,
               sample=f"{n_all} LM steps (linearize + landmark Schur + Cholesky + back-substitution + trial cost) of "
                      f"the benchmark window on {threads} host threads, and {n1} on 1 thread; oracle block-sparse "
                      f"FP64 step (restatement of the reference CPU path's math; Ceres is not available)")
    if bwins:
        st, sec = o.time_windows(bwins, 1e-4, threads, 1)
        if st != 0:
            raise RuntimeError(f"oracle batched step failed: {st}")
        out["batched"] = dict(value=len(bwins) / sec, unit="window-steps/s", cores=threads, seconds=sec,
                              sample=f"one LM step of each of the {len(bwins)} batched windows, whole windows "
                                     f"spread over {threads} host threads")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--batch-windows", type=int, default=256,
                    help="windows of the batched workload line (0: skip)")
    ap.add_argument("--batch-procs", type=int, default=16)
    ap.add_argument("--serve-seconds", type=float, default=1.0,
                    help="seconds of Poisson arrivals of the serving line (0: skip)")
    ap.add_argument("--serve-sessions", type=int, default=64)
    ap.add_argument("--serve-rate", type=float, default=2000.0)
    ap.add_argument("--serve-workers", type=int, default=4,
                    help="server workers (contexts serving disjoint sessions on threads of their own)")
    ap.add_argument("--preint-intervals", type=int, default=4096,
                    help="keyframe edges of the device preintegration line (0: skip)")
    ap.add_argument("--whiten-observations", type=int, default=1 << 22,
                    help="pixel observations of the device whitening line (0: skip)")
    ap.add_argument("--no-slides", dest="slides", action="store_false", help="skip the window-slide line")
    ap.add_argument("--serving-only", action="store_true",
                    help="run only the serving line (prints its JSON object)")
    ap.add_argument("--batched-only", action="store_true",
                    help="run only the batched workload line (prints its JSON object; for per-kernel profiles)")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="after the measurement, launch one plain device copy of this many MiB "
                         "(reference dispatch for the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (--gpus {args.gpus} requested); windows and sessions are "
              "independent, so N GPUs would each run this workload on their own share.", file=sys.stderr)
        sys.exit(2)

    from vibpy import abi, synth
    if args.serving_only:
        ctx = abi.Context(0)
        serving = serving_line(ctx, abi, args.serve_sessions, args.serve_rate, args.serve_seconds, args.batch_procs,
                               64.0, abi.VIB_MIXED, MIXED_REFINEMENT, args.serve_workers)
        ctx.close()
        print(json.dumps(dict(serving=serving)))
        return
    if args.batched_only:
        bwins = batch_windows(args.batch_windows, args.batch_procs)
        ctx = abi.Context(0)
        batched = batched_line(ctx, abi, bwins, max(3, min(args.steps, 10)), max(1, args.warmup))
        ctx.close()
        print(json.dumps(dict(batched=batched)))
        return
    cfg = synth.CONFIGS[args.workload]
    win = synth.generate(**cfg)
    bwins = batch_windows(args.batch_windows, args.batch_procs) if args.batch_windows > 0 else []
    prec = abi.VIB_FP64
    ctx = abi.Context(0)
    ctx.upload(0, win)
    opts = abi.options(max_iterations=1, precision=prec, flags=abi.VIB_SOLVE_CONTINUOUS)
    ctx.bench([0], opts, args.warmup, kernels=False)
    ctx.synchronize()  # device-wide barrier on both sides of the timed region
    t0 = time.perf_counter()
    it_ms, _ = ctx.bench([0], opts, args.steps, kernels=False)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    print(f"[bench] timed {args.steps} steps in {wall:.3f} s", file=sys.stderr, flush=True)
    _, k_ms = ctx.bench([0], opts, max(5, min(args.steps, 20)), kernels=True)
    ctx.synchronize()
    gate = gate_line(ctx, abi, win)
    # the same window in mixed precision, then the mixed-vs-FP64 step error on it and on the
    # ill-conditioned window (depths 0.5-200 m, 30% low-parallax tracks, a near-static stretch)
    mopts = abi.options(max_iterations=1, precision=abi.VIB_MIXED, refinement_steps=MIXED_REFINEMENT,
                        flags=abi.VIB_SOLVE_CONTINUOUS)
    ctx.bench([0], mopts, args.warmup, kernels=False)
    ctx.synchronize()
    t0 = time.perf_counter()
    mit_ms, _ = ctx.bench([0], mopts, args.steps, kernels=False)
    ctx.synchronize()
    mwall = time.perf_counter() - t0
    _, mk_ms = ctx.bench([0], mopts, max(5, min(args.steps, 20)), kernels=True)
    ctx.synchronize()
    mixed = dict(value=args.steps / mwall, unit="window-steps/s", ms_per_step=1e3 * mwall / args.steps,
                 p50_ms=float(np.percentile(mit_ms, 50)), p99_ms=float(np.percentile(mit_ms, 99)), kernel_ms=mk_ms)
    mixed["accuracy"] = {args.workload: step_error(ctx, abi, win, 1, refinement=MIXED_REFINEMENT),
                         "c3": step_error(ctx, abi, synth.generate(**synth.CONFIGS["c3"]), 1, refinement=MIXED_REFINEMENT)}
    mixed["refinement_steps"] = MIXED_REFINEMENT
    batched = None
    if bwins:
        print(f"[bench] batched workload: {len(bwins)} windows", file=sys.stderr, flush=True)
        batched = batched_line(ctx, abi, bwins, max(3, min(args.steps, 10)), max(1, args.warmup))
    serving = None
    if args.serve_seconds > 0:
        print(f"[bench] serving: {args.serve_sessions} sessions, {args.serve_rate:g} requests/s", file=sys.stderr,
              flush=True)
        serving = serving_line(ctx, abi, args.serve_sessions, args.serve_rate, args.serve_seconds, args.batch_procs,
                               64.0, abi.VIB_MIXED, MIXED_REFINEMENT, args.serve_workers)
    preint = None
    if args.preint_intervals > 0:
        print(f"[bench] preintegration: {args.preint_intervals} intervals", file=sys.stderr, flush=True)
        preint = preintegration_line(ctx, abi, args.preint_intervals)
    whiten = None
    if args.whiten_observations > 0:
        print(f"[bench] whitening: {args.whiten_observations} observations", file=sys.stderr, flush=True)
        whiten = whitening_line(ctx, abi, args.whiten_observations)
    slide = slide_line(ctx, abi, win) if args.slides else None
    two_frame = two_frame_line(ctx, abi)
    if args.copy_probe_mb > 0:
        ctx.copy_probe(args.copy_probe_mb << 20)
    ctx.close()

    dom = max(k_ms, key=k_ms.get)
    roof = None
    kb = kernel_bytes(dom, win)
    kf = kernel_flops(dom, win)
    if kb is not None:
        ach = kb / (k_ms[dom] * 1e-3) / 1e9
        roof = dict(bound="hbm", achieved=ach, peak=HBM_PEAK_GBS, unit="GB/s", frac=ach / HBM_PEAK_GBS,
                    traffic=measured_traffic(args.workload, dom), kernel=dom, kernel_ms=k_ms[dom],
                    algorithmic_bytes=kb)
    elif kf is not None:
        ach = kf / (k_ms[dom] * 1e-3) / 1e12
        roof = dict(bound="mfma", achieved=ach, peak=FP64_PEAK_TFLOPS, unit="TFLOP/s", frac=ach / FP64_PEAK_TFLOPS,
                    traffic=measured_traffic(args.workload, dom), kernel=dom, kernel_ms=k_ms[dom],
                    algorithmic_flops=kf)
    value = args.steps / wall
    out = {
        "metric": "Schur window steps/sec + p50/p99 latency, mixed vs FP64 GPU at Δx rel err≤1e-6",
        "value": value,
        "unit": "window-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "p50_ms": float(np.percentile(it_ms, 50)),
        "p99_ms": float(np.percentile(it_ms, 99)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator, inputs resident in HBM)",
        "config": {"workload": f"{args.workload}: K={win.K} keyframes, L={win.L} landmarks, F={win.F} factors "
                               f"({win.F + win.L} observations), {win.cam_rotation.shape[0]} cameras, "
                               f"rig+td online, nr={win.layout()['nr']}",
                   "algorithmic_bytes_per_step": algorithmic_bytes(win)},
        "kernel_ms": k_ms,
        "roofline": roof,
        "mixed": mixed,
    }
    if batched is not None:
        out["batched"] = dict(batched, workload=f"{batched['windows']} independent windows, K 5-20, L 500-30000, "
                                               "2-4 cameras, 2-10 observations per landmark, 5% outliers at 20 px")
    if serving is not None:
        out["serving"] = serving
    out["gate"] = gate
    if slide is not None:
        out["slide"] = slide
    out["two_frame"] = two_frame
    if whiten is not None:
        out["whitening"] = whiten
    if preint is not None:
        out["preintegration"] = preint
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(win, args.cpu_seconds, bwins=bwins)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
