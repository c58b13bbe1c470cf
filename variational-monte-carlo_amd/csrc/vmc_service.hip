// vmc_service.hip -- host-side services over the ensemble API: the
// autotuner and the request server (include/vmc_hip.h).  No device code.
//
// The reference's callers that issue many energy evaluations are its Jastrow
// optimiser (src/optimizer/jastrow_optimizer.cpp:21-171: a grid of b values,
// each a short adaptive run) and main's threads (src/main.cpp:50-68).  The
// server turns such evaluations into requests that share the GPU: each is
// admitted when its HBM footprint fits, owns a stream, and advances a few
// sweeps per scheduler turn, so requests that arrive while others run start
// at once and their kernels overlap on the device.
#include "../../include/vmc_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <set>
#include <thread>
#include <vector>

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Ensemble aggregate as Simulation::run reports it for many chains
// (src/main.cpp:79-104): mean of per-walker means (blocked when two blocks are
// complete), SE = sqrt(sum SE^2) / W over walkers with an SE, mean acceptance.
void aggregate(const std::vector<double> &st, int64_t W, double *mean, double *se, double *acc) {
    double e = 0.0, v = 0.0, a = 0.0;
    int64_t with_se = 0;
    for (int64_t w = 0; w < W; ++w) {
        const double *t = &st[static_cast<size_t>(w) * VMC_STAT_COUNT];
        double m = t[VMC_STAT_NMEAS] > 0 ? t[VMC_STAT_SUM_E] / t[VMC_STAT_NMEAS] : 0.0;
        const double nb = t[VMC_STAT_BLK_N];
        if (nb >= 2) {
            m = t[VMC_STAT_BLK_MEAN];
            const double s = std::sqrt(t[VMC_STAT_BLK_M2] / (nb - 1.0) / nb);
            v += s * s;
            ++with_se;
        }
        e += m;
        a += t[VMC_STAT_PROPOSED] > 0 ? t[VMC_STAT_ACCEPTED] / t[VMC_STAT_PROPOSED] : 0.0;
    }
    *mean = e / static_cast<double>(W);
    *se = with_se ? std::sqrt(v) / static_cast<double>(W) : -1.0;
    *acc = a / static_cast<double>(W);
}

}  // namespace

extern "C" int vmc_default_walker_threads(const vmc_system *sys) {
    // Wavefronts per walker of the generic kernel by size: one up to 63
    // electrons, two up to 127, four above (a function of the system only, so
    // a request's result does not depend on what else runs).
    const int N = sys->n_up + sys->n_dn;
    return N < 64 ? 64 : N < 128 ? 128 : 256;
}

namespace {

int walker_threads_for(const vmc_system &s) { return vmc_default_walker_threads(&s); }

struct Candidate {
    int kernel, variant, placement, threads;
};

// Device time of `sweeps` measured sweeps of a fresh ensemble with the given
// settings (after init and two adaptive sweeps), or < 0 on failure.
double time_candidate(const vmc_system &sys, int64_t W, const Candidate &c, int sweeps, hipStream_t s,
                      double *drift_max) {
    vmc_ensemble *e = nullptr;
    if (vmc_ensemble_create(&sys, W, 12345, 0.1 * sys.box_length, &e)) return -1.0;
    double ms = -1.0;
    hipEvent_t a = nullptr, b = nullptr;
    if (!vmc_ensemble_set_kernel(e, c.kernel) && !vmc_ensemble_set_rr_variant(e, c.variant) &&
        !vmc_ensemble_set_inverse_placement(e, c.placement) &&
        !vmc_ensemble_set_walker_threads(e, c.threads) && !vmc_ensemble_init_positions(e, s) &&
        !vmc_ensemble_sweep(e, 2, VMC_SWEEP_ADAPT, 1, s) && hipEventCreate(&a) == hipSuccess &&
        hipEventCreate(&b) == hipSuccess && hipEventRecord(a, s) == hipSuccess &&
        !vmc_ensemble_sweep(e, sweeps, VMC_SWEEP_MEASURE, 1, s) && hipEventRecord(b, s) == hipSuccess &&
        hipEventSynchronize(b) == hipSuccess) {
        float t = 0.0f;
        if (hipEventElapsedTime(&t, a, b) == hipSuccess) ms = t;
        if (drift_max) {
            std::vector<double> d(static_cast<size_t>(W));
            *drift_max = vmc_ensemble_read_drift(e, d.data(), s) ? INFINITY
                                                                  : *std::max_element(d.begin(), d.end());
        }
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    vmc_ensemble_destroy(e);
    return ms;
}

}  // namespace

extern "C" {

int vmc_autotune(const vmc_system *sys, int64_t n_walkers, double drift_tol, int32_t sweeps,
                 vmc_tuning *out, void *stream) {
    if (!sys || !out || n_walkers < 1 || sweeps < 1 || !(drift_tol > 0.0)) return VMC_ERR_INVALID;
    hipStream_t s = static_cast<hipStream_t>(stream);
    // Kernel families this system can use.
    std::vector<Candidate> cands;
    {
        vmc_ensemble *e = nullptr;
        int rc = vmc_ensemble_create(sys, 1, 1, 0.1 * sys->box_length, &e);
        if (rc) return rc;
        if (!vmc_ensemble_set_kernel(e, VMC_KERNEL_REGISTER)) {
            cands.push_back({VMC_KERNEL_REGISTER, 0, 0, 64});
            cands.push_back({VMC_KERNEL_REGISTER, 1, 0, 64});
        }
        if (!vmc_ensemble_set_kernel(e, VMC_KERNEL_DELAYED)) cands.push_back({VMC_KERNEL_DELAYED, 0, 0, 64});
        const bool lds_ok = !vmc_ensemble_set_inverse_placement(e, 1);
        for (int th : {64, 128, 256}) {
            if (lds_ok) cands.push_back({VMC_KERNEL_GENERIC, 0, 1, th});
            cands.push_back({VMC_KERNEL_GENERIC, 0, 2, th});
        }
        vmc_ensemble_destroy(e);
    }
    vmc_system t = *sys;
    double drift = 0.0;
    // Re-inversion interval: the longest K in {256, 64, 16} whose inverse
    // drift stays below drift_tol (mixed precision only).
    if (sys->precision == VMC_PREC_MIXED) {
        t.reinvert_interval = 16;
        for (int K : {256, 64, 16}) {
            t.reinvert_interval = K;
            double dmax = INFINITY;
            if (time_candidate(t, n_walkers, cands.front(), sweeps, s, &dmax) < 0) return VMC_ERR_DEVICE;
            drift = dmax;
            if (dmax < drift_tol) break;
        }
    }
    double best = -1.0;
    for (const Candidate &c : cands) {
        const double ms = time_candidate(t, n_walkers, c, sweeps, s, nullptr);
        if (ms <= 0.0) continue;
        const double rate = static_cast<double>(n_walkers) * (t.n_up + t.n_dn) * sweeps / (ms * 1e-3);
        if (rate > best) {
            best = rate;
            out->kernel = c.kernel;
            out->walker_threads = c.threads;
            out->rr_variant = c.variant;
            out->inv_placement = c.placement;
        }
    }
    if (best <= 0.0) return VMC_ERR_DEVICE;
    out->reinvert_interval = t.reinvert_interval;
    out->moves_per_s = best;
    out->drift = drift;
    return VMC_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ server

struct vmc_server {
    struct Pending {
        int64_t id;
        vmc_request req;
        double t_submit;
        int64_t bytes;
    };
    struct Active {
        Pending p;
        vmc_ensemble *e = nullptr;
        hipStream_t stream = nullptr;
        int warm_left = 0, meas_left = 0;
        int64_t sweeps_done = 0;  // sweeps issued without error
        int32_t status = VMC_OK;  // the first error a sweep returned
        std::deque<hipEvent_t> inflight;  // one event per issued turn
        double t_start = 0.0;
    };

    int64_t budget = 0;
    int32_t max_active = 0;
    int32_t sweeps_per_turn = 1;
    int device = 0;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Pending> pending;
    std::vector<vmc_result> done;
    int64_t next_id = 1;
    int64_t used = 0;
    int64_t in_flight = 0;  // submitted, result not yet in `done`
    std::set<int64_t> live;     // ids submitted whose result is not yet in `done`
    std::set<int64_t> cancels;  // running requests to end after their turn in flight
    bool stopping = false;
    std::vector<hipStream_t> free_streams;
    std::vector<hipEvent_t> free_events;
    std::thread worker;

    void run();
    void finish(Active &a, std::vector<vmc_result> &out);
};

void vmc_server::finish(Active &a, std::vector<vmc_result> &out) {
    vmc_result r{};
    r.id = a.p.id;
    r.walkers = a.p.req.n_walkers;
    r.status = a.status;
    std::vector<double> st(static_cast<size_t>(a.p.req.n_walkers) * VMC_STAT_COUNT);
    if (hipStreamSynchronize(a.stream) != hipSuccess && r.status == VMC_OK) r.status = VMC_ERR_DEVICE;
    if (r.status == VMC_OK) r.status = vmc_ensemble_read_stats(a.e, st.data(), a.stream);
    // (a cancelled request reports its moves and times, no energy)
    if (r.status == VMC_OK) aggregate(st, a.p.req.n_walkers, &r.mean_energy, &r.standard_error, &r.acceptance_rate);
    const int N = a.p.req.system.n_up + a.p.req.system.n_dn;
    r.moves = static_cast<double>(a.p.req.n_walkers) * N * static_cast<double>(a.sweeps_done);
    const double t = now_s();
    r.queued_s = a.t_start - a.p.t_submit;
    r.latency_s = t - a.p.t_submit;
    for (hipEvent_t ev : a.inflight) free_events.push_back(ev);
    a.inflight.clear();
    vmc_ensemble_destroy(a.e);
    a.e = nullptr;
    free_streams.push_back(a.stream);
    out.push_back(r);
}

void vmc_server::run() {
    (void)hipSetDevice(device);
    std::vector<Active> active;
    for (;;) {
        std::vector<Pending> admit;
        {
            std::unique_lock<std::mutex> lk(mu);
            if (active.empty())
                cv.wait(lk, [&] { return stopping || !pending.empty(); });
            if (stopping && pending.empty() && active.empty()) return;
            // Admission in arrival order while the HBM budget and the slot
            // count allow.
            while (!pending.empty() && static_cast<int32_t>(active.size() + admit.size()) < max_active &&
                   used + pending.front().bytes <= budget) {
                used += pending.front().bytes;
                admit.push_back(pending.front());
                pending.pop_front();
            }
        }
        std::vector<vmc_result> results;
        for (Pending &p : admit) {
            Active a;
            a.p = p;
            a.t_start = now_s();
            if (free_streams.empty()) {
                hipStream_t s = nullptr;
                (void)hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
                free_streams.push_back(s);
            }
            a.stream = free_streams.back();
            free_streams.pop_back();
            int rc = vmc_ensemble_create(&p.req.system, p.req.n_walkers, p.req.seed, p.req.step0, &a.e);
            if (!rc) rc = vmc_ensemble_set_walker_threads(a.e, walker_threads_for(p.req.system));
            if (!rc) rc = vmc_ensemble_init_positions(a.e, a.stream);
            if (rc) {
                vmc_result r{};
                r.id = p.id;
                r.status = rc;
                r.walkers = p.req.n_walkers;
                r.queued_s = a.t_start - p.t_submit;
                r.latency_s = now_s() - p.t_submit;
                if (a.e) vmc_ensemble_destroy(a.e);
                free_streams.push_back(a.stream);
                results.push_back(r);
                std::lock_guard<std::mutex> lk(mu);
                used -= p.bytes;
                continue;
            }
            a.warm_left = p.req.warmup_sweeps;
            a.meas_left = p.req.measure_sweeps;
            active.push_back(std::move(a));
        }
        // Cancelled requests issue no more turns; they report once their
        // turns in flight complete.
        {
            std::lock_guard<std::mutex> lk(mu);
            for (Active &a : active)
                if (cancels.erase(a.p.id)) {
                    if (a.status == VMC_OK) a.status = VMC_ERR_CANCELLED;
                    a.warm_left = a.meas_left = 0;
                }
        }
        // One turn per active request: a few sweeps on its own stream, at
        // most two turns in flight so every stream's queue stays short.
        bool issued = false;
        for (Active &a : active) {
            while (!a.inflight.empty() && hipEventQuery(a.inflight.front()) == hipSuccess) {
                free_events.push_back(a.inflight.front());
                a.inflight.pop_front();
            }
            if (a.warm_left + a.meas_left == 0 || a.inflight.size() >= 2) continue;
            int rc;
            int n_issued;
            if (a.warm_left > 0) {
                const int n = std::min(a.warm_left, sweeps_per_turn);
                n_issued = n;
                rc = vmc_ensemble_sweep(a.e, n, VMC_SWEEP_ADAPT, 1, a.stream);
                a.warm_left -= n;
                if (!rc && a.warm_left == 0) rc = vmc_ensemble_reset_stats(a.e, a.stream);
            } else {
                const int n = std::min(a.meas_left, sweeps_per_turn);
                n_issued = n;
                rc = vmc_ensemble_sweep(a.e, n, VMC_SWEEP_MEASURE, a.p.req.block_size, a.stream);
                a.meas_left -= n;
            }
            if (rc) {
                // The request ends here and reports the error, not a result.
                a.status = rc;
                a.warm_left = a.meas_left = 0;
                continue;
            }
            a.sweeps_done += n_issued;
            hipEvent_t ev = nullptr;
            if (!free_events.empty()) {
                ev = free_events.back();
                free_events.pop_back();
            } else {
                (void)hipEventCreateWithFlags(&ev, hipEventDisableTiming);
            }
            (void)hipEventRecord(ev, a.stream);
            a.inflight.push_back(ev);
            issued = true;
        }
        // Requests whose last turn has completed report and free their slot.
        for (size_t i = 0; i < active.size();) {
            Active &a = active[i];
            bool complete = a.warm_left + a.meas_left == 0;
            if (complete && !a.inflight.empty()) complete = hipEventQuery(a.inflight.back()) == hipSuccess;
            if (complete) {
                const int64_t bytes = a.p.bytes;
                finish(a, results);
                active.erase(active.begin() + static_cast<std::ptrdiff_t>(i));
                std::lock_guard<std::mutex> lk(mu);
                used -= bytes;
            } else {
                ++i;
            }
        }
        if (!results.empty()) {
            std::lock_guard<std::mutex> lk(mu);
            for (auto &r : results) {
                done.push_back(r);
                live.erase(r.id);
                cancels.erase(r.id);
            }
            in_flight -= static_cast<int64_t>(results.size());
            cv.notify_all();
        }
        if (!issued && !active.empty()) std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
}

extern "C" {

int vmc_server_create(int64_t hbm_budget_bytes, int32_t max_active, int32_t sweeps_per_turn,
                      vmc_server **out) {
    if (!out || hbm_budget_bytes <= 0 || max_active < 1 || sweeps_per_turn < 1) return VMC_ERR_INVALID;
    auto *s = new vmc_server();
    s->budget = hbm_budget_bytes;
    s->max_active = max_active;
    s->sweeps_per_turn = sweeps_per_turn;
    if (hipGetDevice(&s->device) != hipSuccess) {
        delete s;
        return VMC_ERR_DEVICE;
    }
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return VMC_OK;
}

int vmc_server_submit(vmc_server *s, const vmc_request *req, int64_t *id) {
    if (!s || !req || !id) return VMC_ERR_INVALID;
    int64_t bytes = 0;
    int rc = vmc_ensemble_footprint(&req->system, req->n_walkers, &bytes);
    if (rc) return rc;
    if (req->warmup_sweeps < 0 || req->measure_sweeps < 1) return VMC_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stopping) return VMC_ERR_STATE;
    if (bytes > s->budget) return VMC_ERR_MEMORY;
    *id = s->next_id++;
    ++s->in_flight;
    s->live.insert(*id);
    s->pending.push_back({*id, *req, now_s(), bytes});
    s->cv.notify_all();
    return VMC_OK;
}

int vmc_server_poll(vmc_server *s, vmc_result *out, int32_t capacity, int32_t *n_out, int32_t wait) {
    if (!s || !n_out || (capacity > 0 && !out)) return VMC_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    if (wait && s->done.empty())
        s->cv.wait(lk, [&] { return !s->done.empty() || s->in_flight == 0; });
    const int32_t n = std::min<int32_t>(capacity, static_cast<int32_t>(s->done.size()));
    for (int32_t i = 0; i < n; ++i) out[i] = s->done[static_cast<size_t>(i)];
    s->done.erase(s->done.begin(), s->done.begin() + n);
    *n_out = n;
    return VMC_OK;
}

int vmc_server_cancel(vmc_server *s, int64_t id) {
    if (!s) return VMC_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    if (!s->live.count(id)) return VMC_ERR_STATE;
    for (auto it = s->pending.begin(); it != s->pending.end(); ++it) {
        if (it->id != id) continue;
        vmc_result r{};
        r.id = id;
        r.status = VMC_ERR_CANCELLED;
        r.walkers = it->req.n_walkers;
        r.standard_error = -1.0;
        r.queued_s = r.latency_s = now_s() - it->t_submit;
        s->pending.erase(it);
        s->done.push_back(r);
        s->live.erase(id);
        --s->in_flight;
        s->cv.notify_all();
        return VMC_OK;
    }
    s->cancels.insert(id);  // admitted: the worker ends it after its turn in flight
    return VMC_OK;
}

int vmc_server_destroy(vmc_server *s) {
    if (!s) return VMC_ERR_INVALID;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
        s->cv.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    for (hipStream_t st : s->free_streams) (void)hipStreamDestroy(st);
    for (hipEvent_t ev : s->free_events) (void)hipEventDestroy(ev);
    delete s;
    return VMC_OK;
}

}  // extern "C"
