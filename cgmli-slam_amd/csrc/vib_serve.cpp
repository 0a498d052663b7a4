// Request server: the serving loop of many sliding-window sessions on one device.
//
// Each session's window and marginalization prior stay resident in a slot of the context. A
// request brings the session's next keyframe and an LM iteration budget; serving it is the
// reference's per-frame cycle (FixedLagBackend::append, then step: solve, then marginalize when
// the window is over its size, src/backend/FixedLagBackend.cpp:17-22, 405-451): the keyframe is
// appended on the device, the LM iterations run batched with the other running requests of the
// same size bucket, and the slide brings the window back to its size. Continuous batching:
// a request that arrives while others are on the device is admitted at its worker's next pass
// (its session idle, its window resident or uploaded within the HBM budget, idle sessions
// evicted least recently used first) and joins the next launch of its bucket.
//
// Workers: each is a context (worker 0: the caller's) with a thread and stream of its own,
// serving a fixed share of the sessions in a loop of passes (admit, start, one LM iteration of
// one bucket, slides). A pass is mostly host latency (edits, synchronizations) with the device
// idle, so workers that do not wait for each other overlap it. The driving thread only queues
// requests and collects results (vib_server_turn).
//
// The server is a client of the library's own C-ABI (upload, append, request begin/iterate,
// marginalize, read/release for eviction); it adds the scheduling: admission by HBM
// footprint, size buckets taken in turn, per-request accounting.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/vib.h"

namespace {

using Clock = std::chrono::steady_clock;

// An owned copy of a vib_window (host memory), viewable as a vib_window.
struct OwnedWindow {
    int K = 0, C = 0, L = 0, F = 0, E = 0, rows = 0, blocks = 0;
    uint32_t flags = 0;
    double huber = 1, min_rho = 1e-8, td_lo = -1, td_hi = 1;
    std::vector<int64_t> frame_id, lm_id, prior_id;
    std::vector<double> pose, sb, rig, td, cam_R, cam_t, lm_bearing, lm_vel, lm_tref, rho, ob_bearing, ob_vel, ob_tref,
        ob_sqrt, imu, prior_ref, prior_A, prior_r0;
    std::vector<uint8_t> pose_free, sb_width, lm_free;
    std::vector<int32_t> lm_frame, lm_cam, ob_lm, ob_frame, ob_cam, imu_from, imu_to, prior_kind;

    // storage for a window of these sizes (prior of `rows` x `cols`, at most `cols` blocks)
    void allocate(int k, int c, int l, int f, int e, int r, int cols) {
        K = k; C = c; L = l; F = f; E = e; rows = r; blocks = std::max(cols, 1);
        frame_id.assign(K, 0); pose.assign(7 * (size_t)K, 0); sb.assign(9 * (size_t)K, 0);
        pose_free.assign(K, 0); sb_width.assign(K, 0); rig.assign(7, 0); td.assign(1, 0);
        cam_R.assign(9 * (size_t)C, 0); cam_t.assign(3 * (size_t)C, 0);
        lm_frame.assign(L, 0); lm_cam.assign(L, 0); lm_bearing.assign(3 * (size_t)L, 0); lm_vel.assign(3 * (size_t)L, 0);
        lm_tref.assign(L, 0); rho.assign(L, 0); lm_free.assign(L, 0); lm_id.assign(L, 0);
        ob_lm.assign(F, 0); ob_frame.assign(F, 0); ob_cam.assign(F, 0); ob_bearing.assign(3 * (size_t)F, 0);
        ob_vel.assign(3 * (size_t)F, 0); ob_tref.assign(F, 0); ob_sqrt.assign(3 * (size_t)F, 0);
        imu_from.assign(E, 0); imu_to.assign(E, 0); imu.assign((size_t)VIB_IMU_STRIDE * E, 0);
        prior_kind.assign(blocks, 0); prior_id.assign(blocks, 0); prior_ref.assign(9 * (size_t)blocks, 0);
        prior_A.assign((size_t)std::max(rows, 1) * std::max(cols, 1), 0); prior_r0.assign(std::max(rows, 1), 0);
    }

    vib_window view() {
        vib_window w{};
        w.num_frames = K; w.num_cameras = C; w.num_landmarks = L; w.num_observations = F; w.num_imu = E;
        w.flags = flags; w.huber_delta = huber; w.min_inverse_depth = min_rho; w.td_lower = td_lo; w.td_upper = td_hi;
        w.frame_id = frame_id.data(); w.pose = pose.data(); w.speed_bias = sb.data(); w.pose_free = pose_free.data();
        w.sb_width = sb_width.data(); w.rig = rig.data(); w.td = td.data(); w.cam_rotation = cam_R.data();
        w.cam_translation = cam_t.data(); w.lm_anchor_frame = lm_frame.data(); w.lm_anchor_cam = lm_cam.data();
        w.lm_bearing = lm_bearing.data(); w.lm_velocity = lm_vel.data(); w.lm_tref = lm_tref.data();
        w.lm_rho = rho.data(); w.lm_free = lm_free.data(); w.obs_landmark = ob_lm.data(); w.obs_frame = ob_frame.data();
        w.obs_cam = ob_cam.data(); w.obs_bearing = ob_bearing.data(); w.obs_velocity = ob_vel.data();
        w.obs_tref = ob_tref.data(); w.obs_sqrt_info = ob_sqrt.data(); w.imu_from = imu_from.data();
        w.imu_to = imu_to.data(); w.imu_data = imu.data(); w.prior_rows = rows; w.prior_blocks = rows ? blocks : 0;
        w.prior_key_kind = prior_kind.data(); w.prior_key_id = prior_id.data(); w.prior_ref = prior_ref.data();
        w.prior_A = prior_A.data(); w.prior_r0 = prior_r0.data(); w.lm_id = lm_id.data();
        return w;
    }

    // a copy of a caller's window
    bool copy_from(const vib_window& w) {
        const int K_ = w.num_frames, C_ = w.num_cameras, L_ = w.num_landmarks, F_ = w.num_observations, E_ = w.num_imu;
        if (K_ < 1 || C_ < 1 || L_ < 0 || F_ < 0 || E_ < 0 || w.prior_rows < 0) return false;
        int cols = 0, refs = 0;
        for (int b = 0; w.prior_rows > 0 && b < w.prior_blocks; ++b) {
            const int k = w.prior_key_kind ? w.prior_key_kind[b] : -1;
            cols += k == VIB_KEY_POSE || k == VIB_KEY_RIG ? 6 : (k == VIB_KEY_SPEED_BIAS ? 9 : 1);
            refs += k == VIB_KEY_POSE || k == VIB_KEY_RIG ? 7 : (k == VIB_KEY_SPEED_BIAS ? 9 : 1);
        }
        allocate(K_, C_, L_, F_, E_, w.prior_rows, std::max(cols, w.prior_blocks));
        flags = w.flags; huber = w.huber_delta; min_rho = w.min_inverse_depth; td_lo = w.td_lower; td_hi = w.td_upper;
        auto cp = [](auto& dst, const auto* src, size_t n) {
            if (src && n) std::copy(src, src + n, dst.begin());
            return src != nullptr || n == 0;
        };
        bool ok = cp(pose, w.pose, 7 * (size_t)K) && cp(sb, w.speed_bias, 9 * (size_t)K) && cp(rig, w.rig, 7) &&
                  cp(td, w.td, 1) && cp(cam_R, w.cam_rotation, 9 * (size_t)C) && cp(cam_t, w.cam_translation, 3 * (size_t)C) &&
                  cp(lm_frame, w.lm_anchor_frame, L) && cp(lm_cam, w.lm_anchor_cam, L) &&
                  cp(lm_bearing, w.lm_bearing, 3 * (size_t)L) && cp(lm_vel, w.lm_velocity, 3 * (size_t)L) &&
                  cp(lm_tref, w.lm_tref, L) && cp(rho, w.lm_rho, L) && cp(ob_lm, w.obs_landmark, F) &&
                  cp(ob_frame, w.obs_frame, F) && cp(ob_cam, w.obs_cam, F) && cp(ob_bearing, w.obs_bearing, 3 * (size_t)F) &&
                  cp(ob_vel, w.obs_velocity, 3 * (size_t)F) && cp(ob_tref, w.obs_tref, F) &&
                  cp(ob_sqrt, w.obs_sqrt_info, 3 * (size_t)F) && cp(imu_from, w.imu_from, E) && cp(imu_to, w.imu_to, E) &&
                  cp(imu, w.imu_data, (size_t)VIB_IMU_STRIDE * E);
        if (!ok) return false;
        for (int k = 0; k < K; ++k) {
            frame_id[k] = w.frame_id ? w.frame_id[k] : k;
            pose_free[k] = w.pose_free ? w.pose_free[k] : 1;
            sb_width[k] = w.sb_width ? w.sb_width[k] : 9;
        }
        for (int l = 0; l < L; ++l) {
            lm_free[l] = w.lm_free ? w.lm_free[l] : 1;
            lm_id[l] = w.lm_id ? w.lm_id[l] : l;
        }
        if (rows) {
            if (!w.prior_key_kind || !w.prior_ref || !w.prior_A || !w.prior_r0) return false;
            blocks = w.prior_blocks;
            prior_kind.assign(w.prior_key_kind, w.prior_key_kind + blocks);
            prior_id.assign(blocks, 0);
            if (w.prior_key_id) prior_id.assign(w.prior_key_id, w.prior_key_id + blocks);
            prior_ref.assign(w.prior_ref, w.prior_ref + refs);
            prior_A.assign(w.prior_A, w.prior_A + (size_t)rows * cols);
            prior_r0.assign(w.prior_r0, w.prior_r0 + rows);
        }
        return true;
    }

    void clear() { *this = OwnedWindow(); }
};

// An owned copy of a keyframe.
struct OwnedKeyframe {
    int64_t frame_id = 0;
    int32_t pose_free = 1, sb_width = 9, nl = 0, nf = 0;
    bool has_imu = false;
    std::vector<double> pose, sb, imu, lm_bearing, lm_vel, lm_tref, lm_rho, ob_bearing, ob_vel, ob_tref, ob_sqrt;
    std::vector<int64_t> lm_id, ob_id;
    std::vector<int32_t> lm_cam, ob_cam;
    std::vector<uint8_t> lm_free;

    bool copy_from(const vib_keyframe& k) {
        if (!k.pose || !k.speed_bias || k.num_landmarks < 0 || k.num_observations < 0) return false;
        nl = k.num_landmarks;
        nf = k.num_observations;
        if (nl && (!k.lm_id || !k.lm_anchor_cam || !k.lm_bearing || !k.lm_velocity || !k.lm_tref || !k.lm_rho)) return false;
        if (nf && (!k.obs_landmark_id || !k.obs_cam || !k.obs_bearing || !k.obs_velocity || !k.obs_tref || !k.obs_sqrt_info))
            return false;
        frame_id = k.frame_id; pose_free = k.pose_free; sb_width = k.sb_width;
        pose.assign(k.pose, k.pose + 7);
        sb.assign(k.speed_bias, k.speed_bias + 9);
        has_imu = k.imu_data != nullptr;
        if (has_imu) imu.assign(k.imu_data, k.imu_data + VIB_IMU_STRIDE);
        const size_t L = nl, F = nf;
        lm_id.assign(k.lm_id, k.lm_id + L); lm_cam.assign(k.lm_anchor_cam, k.lm_anchor_cam + L);
        lm_bearing.assign(k.lm_bearing, k.lm_bearing + 3 * L); lm_vel.assign(k.lm_velocity, k.lm_velocity + 3 * L);
        lm_tref.assign(k.lm_tref, k.lm_tref + L); lm_rho.assign(k.lm_rho, k.lm_rho + L);
        lm_free.assign(L, 1);
        if (k.lm_free) lm_free.assign(k.lm_free, k.lm_free + L);
        ob_id.assign(k.obs_landmark_id, k.obs_landmark_id + F); ob_cam.assign(k.obs_cam, k.obs_cam + F);
        ob_bearing.assign(k.obs_bearing, k.obs_bearing + 3 * F); ob_vel.assign(k.obs_velocity, k.obs_velocity + 3 * F);
        ob_tref.assign(k.obs_tref, k.obs_tref + F); ob_sqrt.assign(k.obs_sqrt_info, k.obs_sqrt_info + 3 * F);
        return true;
    }

    vib_keyframe view() {
        vib_keyframe k{};
        k.frame_id = frame_id; k.pose = pose.data(); k.speed_bias = sb.data(); k.pose_free = pose_free;
        k.sb_width = sb_width; k.imu_data = has_imu ? imu.data() : nullptr; k.num_landmarks = nl; k.num_observations = nf;
        k.lm_id = lm_id.data(); k.lm_anchor_cam = lm_cam.data(); k.lm_bearing = lm_bearing.data();
        k.lm_velocity = lm_vel.data(); k.lm_tref = lm_tref.data(); k.lm_rho = lm_rho.data(); k.lm_free = lm_free.data();
        k.obs_landmark_id = ob_id.data(); k.obs_cam = ob_cam.data(); k.obs_bearing = ob_bearing.data();
        k.obs_velocity = ob_vel.data(); k.obs_tref = ob_tref.data(); k.obs_sqrt_info = ob_sqrt.data();
        return k;
    }
};

struct Request {
    int64_t id = 0;
    int32_t session = 0, iterations = 1, unmatched = 0;
    bool has_keyframe = false;
    bool starting = false;  // admitted, keyframe not yet appended and LM state not yet reset
    OwnedKeyframe keyframe;
    Clock::time_point submitted, admitted;
};

struct Session {
    int32_t id = 0;
    int32_t worker = 0;    // the worker (context) that serves the session
    int32_t slot = -1;     // slot of the worker's context while resident
    int32_t frames = 0;    // the window size a slide restores
    int32_t cameras = 0;
    bool held = false;     // its last result has not been handed out by a turn call yet
    OwnedWindow stored;    // the window while not resident
    std::deque<std::unique_ptr<Request>> queue;
    std::unique_ptr<Request> running;
    Clock::time_point last_used;
};

// One worker: a context of its own (worker 0: the caller's) on a thread of its own, serving its
// sessions in a loop: admit, start (append, LM state), one LM iteration of one size bucket,
// slides. The workers do not wait for each other: a worker whose slides take long (a large
// prior's eigendecomposition) does not hold back the others.
struct Worker {
    vib_context* ctx = nullptr;
    bool owned = false;
    std::vector<int32_t> sessions;  // ids, ascending
    std::deque<int32_t> free_slots;
    int32_t next_slot = 0;
    int64_t turn = 0;               // bucket rotation
    int64_t footprint = 0;          // footprints of its sessions' first windows (assignment)
    vib_server_stats stats{};       // guarded by the server's mutex
    std::thread thread;
};

}  // namespace

struct vib_server {
    vib_context* ctx = nullptr;
    vib_server_options opt{};
    std::map<int32_t, Session> sessions;  // added before the first request; the map is not modified afterwards
    std::vector<Worker> workers;
    // The mutex guards the sessions' queues and states (running, held, resident), the results,
    // the HBM accounting and the statistics; the workers' device work runs outside it.
    std::mutex mu;
    std::condition_variable work;      // a request was queued or a session was released
    std::condition_variable progress;  // a request finished
    std::deque<vib_served> done;       // finished requests not yet handed out
    std::vector<int32_t> handed;       // sessions whose results the last turn call handed out
    bool quit = false;
    int64_t turns = 0;
    double residency_seconds = 0;      // admission (uploads, evictions), summed over the workers
};

namespace {

int bucket_of(const vib_server* s, int nr) {
    for (int i = 0; i < s->opt.num_buckets; ++i)
        if (nr <= s->opt.bucket_limit[i]) return i;
    return std::max(0, s->opt.num_buckets - 1);
}

int32_t take_slot(const vib_server* s, Worker& w) {
    if (!w.free_slots.empty()) {
        const int32_t v = w.free_slots.front();
        w.free_slots.pop_front();
        return v;
    }
    return s->opt.first_slot + w.next_slot++;
}

// The session's window back into host memory, its slot freed (under the server's mutex; the
// session is idle, so no worker is using it).
int evict(Worker& w, Session& se) {
    int32_t K, L, F, E, rows, cols;
    int st = vib_window_shape(w.ctx, se.slot, &K, &L, &F, &E, &rows, &cols);
    if (st != VIB_OK) return st;
    se.stored.allocate(K, se.cameras, L, F, E, rows, cols);
    vib_window win = se.stored.view();
    st = vib_read_window(w.ctx, se.slot, &win);
    if (st != VIB_OK) return st;
    se.stored.flags = win.flags; se.stored.huber = win.huber_delta; se.stored.min_rho = win.min_inverse_depth;
    se.stored.td_lo = win.td_lower; se.stored.td_hi = win.td_upper; se.stored.blocks = win.prior_blocks;
    st = vib_window_release(w.ctx, se.slot);
    if (st != VIB_OK) return st;
    w.free_slots.push_back(se.slot);
    se.slot = -1;
    w.stats.evictions++;
    return VIB_OK;
}

// Makes the session's window resident within the server's HBM budget (idle sessions with nothing
// queued are evicted, least recently used first, whichever worker serves them). False: it does
// not fit now; with *status set: it never will, or the upload failed. Runs under the server's
// mutex.
bool make_resident(vib_server* s, Session& se, int* status) {
    *status = VIB_OK;
    if (se.slot >= 0) return true;
    Worker& w = s->workers[se.worker];
    vib_window win = se.stored.view();
    const int64_t need = vib_window_footprint(&win, s->opt.precision);
    if (need < 0) {
        *status = VIB_ERR_INVALID_ARGUMENT;
        return false;
    }
    const int64_t budget = s->opt.hbm_budget;
    if (budget > 0) {
        if (need > budget) {  // does not fit even with every other window evicted
            *status = VIB_ERR_OUT_OF_MEMORY;
            return false;
        }
        auto resident = [&]() {  // the server's own windows (other slots of the contexts do not count)
            int64_t total = 0;
            for (const auto& kv : s->sessions) {
                int64_t b = 0;
                if (kv.second.slot >= 0 && vib_slot_bytes(s->workers[kv.second.worker].ctx, kv.second.slot, &b) == VIB_OK)
                    total += b;
            }
            return total;
        };
        for (int64_t used = resident(); used + need > budget; used = resident()) {
            Session* victim = nullptr;
            for (auto& kv : s->sessions) {
                Session& o = kv.second;
                if (o.slot >= 0 && !o.running && !o.held && o.queue.empty() && (!victim || o.last_used < victim->last_used))
                    victim = &o;
            }
            if (!victim) return false;
            if ((*status = evict(s->workers[victim->worker], *victim)) != VIB_OK) return false;
        }
    }
    const int32_t slot = take_slot(s, w);
    if ((*status = vib_window_upload(w.ctx, slot, &win)) != VIB_OK) {
        w.free_slots.push_back(slot);
        return false;
    }
    se.slot = slot;
    se.stored.clear();
    w.stats.uploads++;
    return true;
}

// A finished request: its result queued for the next turn call, the session held until then
// (under the server's mutex).
void finish(vib_server* s, Worker& w, Session& se, std::unique_ptr<Request> r, int status, int iterations, double cost) {
    const auto now = Clock::now();
    vib_served out{};
    out.request = r->id;
    out.session = se.id;
    out.status = status;
    out.iterations = iterations;
    out.unmatched = r->unmatched;
    out.final_cost = cost;
    out.queued_seconds = std::chrono::duration<double>(r->admitted - r->submitted).count();
    out.latency_seconds = std::chrono::duration<double>(now - r->submitted).count();
    se.last_used = now;
    se.held = true;
    s->done.push_back(out);
    w.stats.finished++;
    if (status != VIB_OK) w.stats.failed++;
}

// Admission for one worker (under the server's mutex): the first queued request of every idle
// session of the worker whose window is (made) resident. Returns the worker's running sessions.
std::vector<Session*> admit(vib_server* s, Worker& w) {
    const auto t0 = Clock::now();
    std::vector<Session*> run;
    for (int32_t id : w.sessions) {
        Session& se = s->sessions[id];
        if (!se.running && !se.held && !se.queue.empty()) {
            int st = VIB_OK;
            if (!make_resident(s, se, &st)) {
                if (st != VIB_OK) {
                    finish(s, w, se, std::move(se.queue.front()), st, 0, 0.0);
                    se.queue.pop_front();
                }
                continue;
            }
            se.running = std::move(se.queue.front());
            se.queue.pop_front();
            se.running->admitted = Clock::now();
            se.running->starting = true;
        }
        if (se.running) run.push_back(&se);
    }
    s->residency_seconds += std::chrono::duration<double>(Clock::now() - t0).count();
    return run;
}

struct Outcome {
    Session* se;
    int status, iterations;
    double cost;
};

// One pass of a worker over its running sessions (outside the server's mutex: only the worker
// touches running sessions and their windows): the admitted requests started (keyframe appended,
// LM state reset), one LM iteration of the running requests of one size bucket, and the slides
// of the requests that finished.
std::vector<Outcome> serve_pass(vib_server* s, Worker& w, const std::vector<Session*>& run_all, vib_server_stats& st) {
    auto since = [](Clock::time_point t) { return std::chrono::duration<double>(Clock::now() - t).count(); };
    std::vector<Outcome> out;
    const auto t_admit = Clock::now();
    std::vector<Session*> live;
    for (Session* se : run_all) {
        Request& r = *se->running;
        if (r.starting) {
            r.starting = false;
            int code = VIB_OK;
            if (r.has_keyframe) {
                vib_keyframe k = r.keyframe.view();
                code = vib_window_append(w.ctx, se->slot, &k, &r.unmatched);
                st.appends++;
                r.keyframe = OwnedKeyframe();  // the window holds it now
            }
            vib_solve_options o{};
            o.max_iterations = r.iterations;
            o.min_iterations = 0;
            o.initial_lambda = 1e-4;
            o.precision = s->opt.precision;
            o.refinement_steps = s->opt.refinement_steps;
            o.flags = s->opt.full_iterations ? VIB_SOLVE_CONTINUOUS : 0u;
            if (code == VIB_OK) code = vib_request_begin(w.ctx, se->slot, &o);
            if (code != VIB_OK) {
                out.push_back(Outcome{se, code, 0, 0.0});
                continue;
            }
        }
        live.push_back(se);
    }
    st.admit_seconds += since(t_admit);
    // one LM iteration of the running requests of one size bucket (buckets in turn)
    std::map<int, std::vector<Session*>> by_bucket;
    for (Session* se : live) {
        int32_t nr = 0, nv = 0;
        vib_retained_dimension(w.ctx, se->slot, &nr, &nv);
        by_bucket[bucket_of(s, nr)].push_back(se);
    }
    if (by_bucket.empty()) return out;
    std::vector<int> keys;
    for (auto& kv : by_bucket) keys.push_back(kv.first);
    const int b = keys[w.turn++ % keys.size()];
    std::vector<Session*>& run = by_bucket[b];
    std::vector<int32_t> ids, dn(run.size());
    std::vector<vib_solve_report> reps(run.size());
    for (Session* se : run) ids.push_back(se->slot);
    const auto t_iter = Clock::now();
    const int code = vib_request_iterate(w.ctx, (int32_t)ids.size(), ids.data(), s->opt.precision,
                                         s->opt.refinement_steps, dn.data(), reps.data());
    st.iterate_seconds += since(t_iter);
    st.launches++;
    st.window_iterations += (int64_t)ids.size();
    // the finished requests' slides back to their windows' sizes (FixedLagBackend.cpp:434-451),
    // all in one batch
    std::vector<size_t> fin, slide;
    std::vector<int> status(run.size(), VIB_OK);
    for (size_t i = 0; i < run.size(); ++i) {
        if (code == VIB_OK && !dn[i]) continue;
        fin.push_back(i);
        status[i] = code != VIB_OK ? code : reps[i].status;
        int32_t K = 0, L, F, E, rows, cols;
        if (status[i] == VIB_OK && s->opt.slide_policy >= 0 &&
            vib_window_shape(w.ctx, run[i]->slot, &K, &L, &F, &E, &rows, &cols) == VIB_OK && K > run[i]->frames)
            slide.push_back(i);
    }
    if (!slide.empty()) {
        std::vector<int32_t> sl;
        for (size_t i : slide) sl.push_back(run[i]->slot);
        std::vector<vib_marginal_report> mr(slide.size());
        const auto t_slide = Clock::now();
        (void)vib_marginalize_batch(w.ctx, (int32_t)sl.size(), sl.data(), s->opt.slide_policy, mr.data());
        st.slide_seconds += since(t_slide);
        st.slides += (int64_t)slide.size();
        st.slide_launches++;
        for (size_t q = 0; q < slide.size(); ++q) status[slide[q]] = mr[q].status;
    }
    for (size_t i : fin)
        out.push_back(Outcome{run[i], status[i], code == VIB_OK ? reps[i].iterations : 0,
                              code == VIB_OK ? reps[i].final_cost : 0.0});
    return out;
}

bool has_work(const vib_server* s, const Worker& w) {
    for (int32_t id : w.sessions) {
        const Session& se = s->sessions.at(id);
        if (se.running || (!se.held && !se.queue.empty())) return true;
    }
    return false;
}

void worker_loop(vib_server* s, int index) {
    Worker& w = s->workers[index];
    for (;;) {
        std::vector<Session*> run;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->work.wait(lk, [&] { return s->quit || has_work(s, w); });
            if (s->quit) return;
            run = admit(s, w);
        }
        if (run.empty()) {  // nothing admissible now (budget): wait for a release or a new request
            std::unique_lock<std::mutex> lk(s->mu);
            s->work.wait_for(lk, std::chrono::milliseconds(1));
            continue;
        }
        vib_server_stats st{};
        std::vector<Outcome> out = serve_pass(s, w, run, st);
        std::lock_guard<std::mutex> lk(s->mu);
        w.stats.appends += st.appends; w.stats.launches += st.launches; w.stats.window_iterations += st.window_iterations;
        w.stats.slides += st.slides; w.stats.slide_launches += st.slide_launches; w.stats.admit_seconds += st.admit_seconds;
        w.stats.iterate_seconds += st.iterate_seconds; w.stats.slide_seconds += st.slide_seconds;
        for (Outcome& o : out) finish(s, w, *o.se, std::move(o.se->running), o.status, o.iterations, o.cost);
        if (!out.empty()) s->progress.notify_all();
    }
}

void stop_workers(vib_server* s) {
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->quit = true;
    }
    s->work.notify_all();
    for (Worker& w : s->workers)
        if (w.thread.joinable()) w.thread.join();
}

void release_all(vib_server* s) {
    stop_workers(s);
    for (auto& kv : s->sessions)
        if (kv.second.slot >= 0) (void)vib_window_release(s->workers[kv.second.worker].ctx, kv.second.slot);
    for (Worker& w : s->workers)
        if (w.owned && w.ctx) (void)vib_destroy(w.ctx);
}

}  // namespace

extern "C" {

int vib_server_create(vib_context* ctx, const vib_server_options* o, vib_server** out) {
    if (!ctx || !o || !out) return VIB_ERR_INVALID_ARGUMENT;
    if (o->precision != VIB_FP64 && o->precision != VIB_MIXED) return VIB_ERR_INVALID_ARGUMENT;
    if (o->num_buckets < 0 || o->num_buckets > 8 || o->first_slot < 0) return VIB_ERR_INVALID_ARGUMENT;
    if (o->slide_policy != -1 && o->slide_policy != VIB_MARGIN_OLD && o->slide_policy != VIB_MARGIN_SECOND_NEW)
        return VIB_ERR_INVALID_ARGUMENT;
    if (o->workers < 0 || o->workers > VIB_SERVER_MAX_WORKERS || o->hbm_budget < 0) return VIB_ERR_INVALID_ARGUMENT;
    int32_t device = 0;
    if (vib_context_device(ctx, &device) != VIB_OK) return VIB_ERR_INVALID_ARGUMENT;
    auto* s = new vib_server();
    s->ctx = ctx;
    s->opt = *o;
    const int n = std::max(1, o->workers);
    s->workers.resize(n);
    for (int i = 0; i < n; ++i) {
        Worker& w = s->workers[i];
        if (i == 0) {
            w.ctx = ctx;
        } else if (vib_create(&w.ctx, device) == VIB_OK) {
            w.owned = true;
        } else {
            w.ctx = nullptr;
            release_all(s);
            delete s;
            return VIB_ERR_DEVICE;
        }
    }
    for (int i = 0; i < n; ++i) s->workers[i].thread = std::thread(worker_loop, s, i);
    *out = s;
    return VIB_OK;
}

int vib_server_destroy(vib_server* s) {
    if (!s) return VIB_ERR_INVALID_ARGUMENT;
    release_all(s);
    delete s;
    return VIB_OK;
}

int vib_server_add_session(vib_server* s, int32_t session, const vib_window* w) {
    if (!s || !w) return VIB_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->sessions.count(session)) return VIB_ERR_INVALID_ARGUMENT;
    Session se;
    se.id = session;
    if (!se.stored.copy_from(*w)) return VIB_ERR_INVALID_ARGUMENT;
    se.frames = w->num_frames;
    se.cameras = w->num_cameras;
    // the worker with the smallest share of session footprints
    const int64_t fp = std::max<int64_t>(0, vib_window_footprint(w, s->opt.precision));
    int best = 0;
    for (int i = 1; i < (int)s->workers.size(); ++i)
        if (s->workers[i].footprint < s->workers[best].footprint) best = i;
    se.worker = best;
    Worker& wk = s->workers[best];
    wk.footprint += fp;
    wk.sessions.insert(std::upper_bound(wk.sessions.begin(), wk.sessions.end(), session), session);
    s->sessions.emplace(session, std::move(se));
    return VIB_OK;
}

int vib_server_submit(vib_server* s, int64_t request, int32_t session, int32_t iterations, const vib_keyframe* kf) {
    if (!s || iterations < 1) return VIB_ERR_INVALID_ARGUMENT;
    auto r = std::make_unique<Request>();
    r->id = request;
    r->session = session;
    r->iterations = iterations;
    r->has_keyframe = kf != nullptr;
    if (kf && !r->keyframe.copy_from(*kf)) return VIB_ERR_INVALID_ARGUMENT;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->sessions.find(session);
        if (it == s->sessions.end()) return VIB_ERR_NOT_FOUND;
        r->submitted = Clock::now();
        it->second.queue.push_back(std::move(r));
        s->workers[it->second.worker].stats.submitted++;
    }
    s->work.notify_all();
    return VIB_OK;
}

int vib_server_turn(vib_server* s, vib_served* finished, int32_t capacity, int32_t* count) {
    if (!s || !count || capacity < 0 || (capacity > 0 && !finished)) return VIB_ERR_INVALID_ARGUMENT;
    *count = 0;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        s->turns++;
        // the sessions handed out by the previous call may take their next request now
        for (int32_t id : s->handed) s->sessions[id].held = false;
        s->handed.clear();
        if (s->done.empty()) s->progress.wait_for(lk, std::chrono::milliseconds(1));
        while (*count < capacity && !s->done.empty()) {
            finished[(*count)++] = s->done.front();
            s->handed.push_back(s->done.front().session);
            s->done.pop_front();
        }
    }
    s->work.notify_all();
    return VIB_OK;
}

int vib_server_preload(vib_server* s) {
    if (!s) return VIB_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(s->mu);
    for (auto& kv : s->sessions) {
        int st = VIB_OK;
        if (!make_resident(s, kv.second, &st)) return st;
        kv.second.last_used = Clock::now();
    }
    return VIB_OK;
}

int vib_server_get_stats(const vib_server* cs, vib_server_stats* out) {
    if (!cs || !out) return VIB_ERR_INVALID_ARGUMENT;
    vib_server* s = const_cast<vib_server*>(cs);  // the mutex
    std::lock_guard<std::mutex> lk(s->mu);
    *out = vib_server_stats{};
    for (const Worker& w : s->workers) {  // counters and phase times summed over the workers
        const vib_server_stats& x = w.stats;
        out->submitted += x.submitted; out->finished += x.finished; out->failed += x.failed;
        out->launches += x.launches; out->window_iterations += x.window_iterations; out->uploads += x.uploads;
        out->evictions += x.evictions; out->appends += x.appends; out->slides += x.slides;
        out->slide_launches += x.slide_launches; out->admit_seconds += x.admit_seconds;
        out->iterate_seconds += x.iterate_seconds; out->slide_seconds += x.slide_seconds;
    }
    for (const auto& kv : s->sessions) {
        int64_t b = 0;
        if (kv.second.slot >= 0 && vib_slot_bytes(s->workers[kv.second.worker].ctx, kv.second.slot, &b) == VIB_OK)
            out->resident_bytes += b;
        out->queued += (int32_t)kv.second.queue.size();
        out->running += kv.second.running ? 1 : 0;
        out->resident += kv.second.slot >= 0 ? 1 : 0;
    }
    out->turns = s->turns;
    out->admit_seconds += s->residency_seconds;
    out->pending_results = (int32_t)s->done.size();
    return VIB_OK;
}

int vib_server_session_slot(const vib_server* s, int32_t session, int32_t* slot) {
    return vib_server_session_context(s, session, nullptr, slot);
}

int vib_server_session_context(const vib_server* cs, int32_t session, vib_context** ctx, int32_t* slot) {
    if (!cs || !slot) return VIB_ERR_INVALID_ARGUMENT;
    vib_server* s = const_cast<vib_server*>(cs);
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->sessions.find(session);
    if (it == s->sessions.end()) return VIB_ERR_NOT_FOUND;
    *slot = it->second.slot;
    if (ctx) *ctx = s->workers[it->second.worker].ctx;
    return VIB_OK;
}

}  // extern "C"
