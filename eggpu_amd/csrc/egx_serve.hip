// egx_serve.hip -- device-resident PageRank / personalised PageRank serving.
//
// A server owns one worker thread and a pool of device-resident graphs (the
// caller's egx_graph handles).  Requests (global PageRank: no seeds;
// personalised: 1..n seeds) are queued by egx_server_submit from any thread
// and answered by continuous batching:
//
//   * every graph with work has one batch session (64 column slots, its own
//     HIP stream, X rows as wide as its highest occupied slot);
//   * between iteration steps the worker admits queued requests into free
//     slots -- a request that arrives while others are mid-solve joins the
//     running session at the next step boundary;
//   * a new session opens only while the HBM budget holds its footprint;
//     idle sessions are closed (least recently used first) to make room;
//   * each round the worker enqueues `step_iters` iterations on every busy
//     session before it polls any, so sessions on different graphs overlap
//     on the device; a finished column is normalised and copied into a
//     pinned staging buffer on its session's stream and its slot freed at
//     once; delivery threads wait for the copy and move the result into the
//     callers' buffers, so the worker never blocks on a result.
//
// A column's arithmetic does not depend on the other columns of its session
// (inactive and finished columns are copied through; sums use a fixed
// order), so each result is bit-identical alone and in the mix.  For the
// same reason a request identical to one already in a slot of its graph's
// session (same seed list, alpha, max_iter and tol) rides on that column:
// it receives the column's result and iteration count, exactly what a
// column of its own would have produced (request coalescing).
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "egx_graph.h"

namespace {

using Clock = std::chrono::steady_clock;

double seconds_since(Clock::time_point t0, Clock::time_point t) {
    return std::chrono::duration<double>(t - t0).count();
}

constexpr int kSlots = 64;
constexpr int kDeliveryThreads = 8;

struct Request {
    int64_t id = 0;
    int32_t graph = 0;
    std::vector<int32_t> seeds;
    double alpha = 0.85;
    int32_t max_iter = 100;
    double tol = 1e-6;
    double* out = nullptr;
    Clock::time_point t_submit, t_start, t_done;
    int32_t iterations = 0;
    int32_t width = 0;  // columns of the session when the request finished
    int status = EGX_SUCC;
    std::string msg;  // failure detail, from the thread where the failure happened
    bool done = false;
    std::vector<std::shared_ptr<Request>> riders;  // identical requests sharing this column

    bool same_solve(const Request& o) const {
        return graph == o.graph && seeds == o.seeds && alpha == o.alpha && max_iter == o.max_iter && tol == o.tol;
    }
};

struct Session {
    int32_t graph = -1;
    egx_batch* b = nullptr;
    int64_t bytes = 0;
    std::vector<std::shared_ptr<Request>> slot;  // kSlots entries
    int32_t busy = 0;
    bool in_flight = false;  // a step is queued and not yet collected
    Clock::time_point last_used;
};

int64_t session_bytes(const egx_graph_info& inf) {
    // two X buffers at the widest layout, teleport slots, output, partials
    const int64_t n = inf.n;
    return 2 * n * kSlots * 4 + n * 4 + n * 8 + (int64_t)1024 * kSlots * 4 +
           (int64_t)(inf.tiles + inf.hub_segments + 8192) * 3 * kSlots * 8 + ((int64_t)1 << 20);
}

// A finished column on its way to the callers' buffers.
struct Delivery {
    std::shared_ptr<Request> r;
    char* staged = nullptr;    // pinned packed copy of the result (null: nobody asked for it)
    hipEvent_t ev = nullptr;   // recorded after the copy into `staged`
};

}  // namespace

struct egx_server {
    std::vector<egx_graph*> pool;
    std::vector<egx_graph_info> info;
    egx_server_opts opts{};
    int device = -1;

    std::mutex mu;
    std::condition_variable cv_work;  // worker: new request or stop
    std::condition_variable cv_done;  // waiters: a request finished
    std::deque<std::shared_ptr<Request>> pending;
    std::map<int64_t, std::shared_ptr<Request>> all;  // submitted, not yet collected
    int64_t next_id = 1;
    bool stop = false;

    // worker-owned
    std::list<Session> sessions;
    int64_t used_bytes = 0;
    egx_server_stats stats{};
    std::thread worker;

    // result delivery (guarded by mu)
    std::deque<Delivery> deliveries;
    std::condition_variable cv_deliver;
    bool deliver_stop = false;
    std::vector<std::thread> deliverers;
    std::vector<std::vector<char*>> staged_free;  // per graph: pinned buffers of one packed result
    std::vector<char*> staged_all;
    std::vector<hipEvent_t> events_free, events_all;

    ~egx_server() {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv_work.notify_all();
        if (worker.joinable()) worker.join();
        {
            std::lock_guard<std::mutex> lk(mu);
            deliver_stop = true;
        }
        cv_deliver.notify_all();
        for (auto& t : deliverers)
            if (t.joinable()) t.join();  // every queued delivery is completed first
        for (Session& s : sessions)
            if (s.b) egx_batch_close(s.b);
        for (hipEvent_t e : events_all) (void)hipEventDestroy(e);
        for (char* p : staged_all) (void)hipHostFree(p);
    }

    // A pinned buffer for one result of `graph` and an event, from the pools.
    int take_staging(int32_t graph, char** buf, hipEvent_t* ev) {
        *buf = nullptr;
        *ev = nullptr;
        {
            std::lock_guard<std::mutex> lk(mu);
            auto& fl = staged_free[(size_t)graph];
            if (!fl.empty()) {
                *buf = fl.back();
                fl.pop_back();
            }
            if (!events_free.empty()) {
                *ev = events_free.back();
                events_free.pop_back();
            }
        }
        if (!*buf) {
            void* p = nullptr;
            const size_t bytes = egx::kPackedHeader + (size_t)info[(size_t)graph].n * sizeof(float);
            if (hipHostMalloc(&p, bytes) != hipSuccess) return EGX_FAILED_TO_ALLOCATE_HOST_MEM;
            *buf = static_cast<char*>(p);
            std::lock_guard<std::mutex> lk(mu);
            staged_all.push_back(*buf);
        }
        if (!*ev) {
            if (hipEventCreateWithFlags(ev, hipEventDisableTiming) != hipSuccess) {
                give_back(graph, *buf, nullptr);
                *buf = nullptr;
                return EGX_DEVICE_ERR;
            }
            std::lock_guard<std::mutex> lk(mu);
            events_all.push_back(*ev);
        }
        return EGX_SUCC;
    }

    void give_back(int32_t graph, char* buf, hipEvent_t ev) {
        std::lock_guard<std::mutex> lk(mu);
        if (buf) staged_free[(size_t)graph].push_back(buf);
        if (ev) events_free.push_back(ev);
    }

    // Delivery thread: wait for a staged (packed) result, finish its FP64
    // values into the request's buffer and copy them to those of the
    // requests riding on its column, complete them.
    void deliver_loop() {
        (void)hipSetDevice(device);
        for (;;) {
            Delivery d;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_deliver.wait(lk, [&] { return deliver_stop || !deliveries.empty(); });
                if (deliveries.empty()) return;  // stopping, and nothing left to deliver
                d = std::move(deliveries.front());
                deliveries.pop_front();
            }
            int st = EGX_SUCC;
            if (d.ev) {
                const hipError_t e = hipEventSynchronize(d.ev);
                if (e != hipSuccess) {
                    egx::set_last_error(std::string("result copy event: ") + hipGetErrorString(e));
                    st = EGX_DEVICE_ERR;
                }
            }
            if (st == EGX_SUCC && d.staged) {
                const int32_t n = info[(size_t)d.r->graph].n;
                const double* first = nullptr;
                if (d.r->out) {
                    egx::unpack_result(d.staged, n, d.r->out);
                    first = d.r->out;
                }
                for (auto& q : d.r->riders) {
                    if (!q->out) continue;
                    if (first) {
                        std::memcpy(q->out, first, (size_t)n * sizeof(double));
                    } else {
                        egx::unpack_result(d.staged, n, q->out);
                        first = q->out;
                    }
                }
            }
            give_back(d.r->graph, d.staged, d.ev);
            finish(d.r, st);
        }
    }

    // Completes r and the requests riding on its column.  A failure keeps
    // this thread's last error text, which egx_server_wait hands to the caller.
    void finish(const std::shared_ptr<Request>& r, int status) {
        const std::string why = status != EGX_SUCC ? std::string(egx_last_error()) : std::string();
        std::lock_guard<std::mutex> lk(mu);
        const Clock::time_point now = Clock::now();
        auto one = [&](Request& q) {
            q.status = status;
            q.msg = why;
            q.t_done = now;
            q.done = true;
            stats.completed += 1;
            if (status != EGX_SUCC) stats.failed += 1;
        };
        one(*r);
        for (auto& q : r->riders) {
            q->iterations = r->iterations;
            q->width = r->width;
            one(*q);
        }
        r->riders.clear();
        cv_done.notify_all();
    }

    Session* session_for(int32_t graph) {
        for (Session& s : sessions)
            if (s.graph == graph) return &s;
        return nullptr;
    }

    // Open a session for `graph` within the budget, closing idle ones first.
    int open_session(int32_t graph, Session** out) {
        *out = nullptr;
        const int64_t need = session_bytes(info[(size_t)graph]);
        if (need > opts.hbm_budget_bytes) {
            egx::set_last_error("a batch session on this graph exceeds the server's HBM budget");
            return EGX_FAILED_TO_ALLOCATE_DEVICE_MEM;
        }
        while (used_bytes + need > opts.hbm_budget_bytes ||
               (int32_t)sessions.size() >= opts.max_sessions) {
            auto victim = sessions.end();
            for (auto it = sessions.begin(); it != sessions.end(); ++it)
                if (it->busy == 0 && !it->in_flight &&
                    (victim == sessions.end() || it->last_used < victim->last_used))
                    victim = it;
            if (victim == sessions.end()) return EGX_SUCC;  // wait for a busy session to drain
            egx_batch_close(victim->b);
            used_bytes -= victim->bytes;
            sessions.erase(victim);
            std::lock_guard<std::mutex> lk(mu);
            stats.sessions_closed += 1;
        }
        Session s;
        s.graph = graph;
        s.bytes = need;
        s.slot.assign(kSlots, nullptr);
        s.last_used = Clock::now();
        int rc = egx_batch_open(pool[(size_t)graph], &s.b);
        if (rc != EGX_SUCC) return rc;
        used_bytes += need;
        {
            std::lock_guard<std::mutex> lk(mu);
            stats.sessions_opened += 1;
            stats.peak_session_bytes = std::max<int64_t>(stats.peak_session_bytes, used_bytes);
        }
        sessions.push_back(std::move(s));
        *out = &sessions.back();
        return EGX_SUCC;
    }

    // Move queued requests into free slots (arrival order per graph).
    void admit() {
        std::vector<std::shared_ptr<Request>> take;
        {
            std::lock_guard<std::mutex> lk(mu);
            take.assign(pending.begin(), pending.end());
            pending.clear();
        }
        std::vector<std::shared_ptr<Request>> keep;
        struct Picks {
            Session* s;
            std::vector<std::pair<int32_t, std::shared_ptr<Request>>> slots;
        };
        std::vector<Picks> picks;
        for (auto& r : take) {
            Session* s = session_for(r->graph);
            if (!s) {
                int rc = open_session(r->graph, &s);
                if (rc != EGX_SUCC) {
                    finish(r, rc);
                    continue;
                }
                if (!s) {
                    keep.push_back(r);
                    continue;
                }
            }
            if (!opts.coalesce_off) {
                std::shared_ptr<Request> host;
                for (int32_t k = 0; k < kSlots && !host; ++k) {
                    const auto& q = s->slot[(size_t)k];
                    if (q && q->same_solve(*r)) host = q;
                }
                if (host) {
                    r->t_start = Clock::now();
                    host->riders.push_back(r);
                    s->last_used = r->t_start;
                    std::lock_guard<std::mutex> lk(mu);
                    stats.coalesced += 1;
                    continue;
                }
            }
            int32_t j = -1;
            for (int32_t k = 0; k < kSlots && j < 0; ++k)
                if (!s->slot[(size_t)k]) j = k;
            if (j < 0) {
                keep.push_back(r);
                continue;
            }
            // the slot is taken (busy: the session cannot be closed) until the
            // session's assignments of this pass are made together below
            s->slot[(size_t)j] = r;
            s->busy += 1;
            auto pick = std::find_if(picks.begin(), picks.end(), [&](const Picks& p) { return p.s == s; });
            if (pick == picks.end()) pick = picks.insert(picks.end(), Picks{s, {}});
            pick->slots.emplace_back(j, r);
        }
        // one update per session: one copy and one launch for all its new columns
        for (Picks& p : picks) {
            std::vector<egx::BatchAssign> reqs;
            reqs.reserve(p.slots.size());
            for (const auto& [j, r] : p.slots)
                reqs.push_back(egx::BatchAssign{j, r->seeds.empty() ? nullptr : r->seeds.data(),
                                                (int32_t)r->seeds.size(), r->alpha, r->max_iter, r->tol});
            std::vector<int> st(reqs.size(), EGX_SUCC);
            const int rc = egx::batch_assign_many(p.s->b, (int32_t)reqs.size(), reqs.data(), st.data());
            const Clock::time_point now = Clock::now();
            for (size_t i = 0; i < p.slots.size(); ++i) {
                const auto& [j, r] = p.slots[i];
                const int code = (rc != EGX_SUCC && st[i] == EGX_SUCC) ? rc : st[i];
                if (code != EGX_SUCC) {
                    p.s->slot[(size_t)j] = nullptr;
                    p.s->busy -= 1;
                    finish(r, code);  // and the requests that rode on it
                    continue;
                }
                r->t_start = now;
                p.s->last_used = now;
            }
        }
        if (!picks.empty()) {
            const int64_t act = active_columns();
            std::lock_guard<std::mutex> lk(mu);
            stats.peak_active_columns = std::max<int64_t>(stats.peak_active_columns, act);
        }
        if (!keep.empty()) {
            std::lock_guard<std::mutex> lk(mu);
            pending.insert(pending.begin(), keep.begin(), keep.end());
        }
    }

    int64_t active_columns() const {
        int64_t a = 0;
        for (const Session& s : sessions) a += s.busy;
        return a;
    }

    // Collect finished columns of session s (synchronises its stream).
    void collect(Session& s) {
        int32_t active[kSlots], done[kSlots], iters[kSlots];
        int rc = egx_batch_poll(s.b, active, done, iters, nullptr);
        if (rc != EGX_SUCC) {
            for (int32_t j = 0; j < kSlots; ++j)
                if (s.slot[(size_t)j]) {
                    finish(s.slot[(size_t)j], rc);
                    s.slot[(size_t)j] = nullptr;
                    s.busy -= 1;
                }
            return;
        }
        int32_t hi = 0;
        for (int32_t j = 0; j < kSlots; ++j)
            if (s.slot[(size_t)j]) hi = j + 1;
        int32_t width = 8;
        while (width < hi) width <<= 1;
        std::vector<std::pair<int32_t, Delivery>> finished;
        for (int32_t j = 0; j < kSlots; ++j) {
            auto& r = s.slot[(size_t)j];
            if (!r || !done[j]) continue;
            r->iterations = iters[j];
            r->width = width;
            bool wanted = r->out != nullptr;
            for (auto& q : r->riders) wanted = wanted || q->out != nullptr;
            Delivery d;
            d.r = r;
            int st = EGX_SUCC;
            if (wanted) {
                st = take_staging(s.graph, &d.staged, &d.ev);
                if (st == EGX_SUCC) st = egx::batch_result_enqueue(s.b, j, d.staged, d.ev);
            }
            if (st == EGX_SUCC) {
                finished.emplace_back(j, std::move(d));
            } else {
                drop_delivery(s, d, st);
            }
            r = nullptr;
            s.busy -= 1;
            s.last_used = Clock::now();
        }
        if (finished.empty()) return;
        // one update releases every finished column (stream order: after
        // their results have been read)
        std::vector<int32_t> js;
        js.reserve(finished.size());
        for (const auto& f : finished) js.push_back(f.first);
        const int st = egx::batch_release_many(s.b, (int32_t)js.size(), js.data());
        if (st != EGX_SUCC) {
            for (auto& f : finished) drop_delivery(s, f.second, st);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            for (auto& f : finished) deliveries.push_back(std::move(f.second));
        }
        cv_deliver.notify_all();
    }

    // A result that will not be delivered: its copy lands first, its buffer
    // goes back to the pool, and its request fails with `status`.
    void drop_delivery(Session& s, Delivery& d, int status) {
        if (d.ev) (void)hipEventSynchronize(d.ev);
        give_back(s.graph, d.staged, d.ev);
        finish(d.r, status);
    }

    // Sessions progress independently: a session whose step has finished is
    // collected and stepped again at once, without waiting for sessions on
    // larger graphs; queued requests are admitted between these checks.
    void run() {
        (void)hipSetDevice(device);
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] { return stop || !pending.empty() || active_columns() > 0; });
                if (stop) break;
            }
            admit();
            bool progressed = false;
            int64_t launched = 0;
            for (Session& s : sessions) {
                if (s.in_flight) {
                    const int r = egx_batch_ready(s.b);
                    if (r == 0) continue;
                    s.in_flight = false;
                    collect(s);  // also reports a failed step
                    progressed = true;
                }
                if (s.busy > 0) {
                    if (egx_batch_step(s.b, opts.step_iters) == EGX_SUCC) {
                        s.in_flight = true;
                        ++launched;
                    } else {
                        collect(s);
                    }
                    progressed = true;
                }
            }
            if (launched) {
                std::lock_guard<std::mutex> lk(mu);
                stats.steps += launched;
            }
            if (!progressed) {
                // work in flight and nothing finished yet: wait briefly for a
                // submission or for the device
                std::unique_lock<std::mutex> lk(mu);
                if (!stop && pending.empty())
                    cv_work.wait_for(lk, std::chrono::microseconds(50), [&] { return stop || !pending.empty(); });
            }
        }
        // fail whatever is left
        std::vector<std::shared_ptr<Request>> left;
        {
            std::lock_guard<std::mutex> lk(mu);
            left.assign(pending.begin(), pending.end());
            pending.clear();
        }
        for (auto& r : left) finish(r, EGX_UNKNOW_ERROR);
        for (Session& s : sessions)
            for (auto& r : s.slot)
                if (r) {
                    finish(r, EGX_UNKNOW_ERROR);
                    r = nullptr;
                }
    }
};

extern "C" {

int egx_server_open(egx_graph* const* pool, int32_t count, const egx_server_opts* opts, egx_server** out) {
    if (!out || count < 1 || !pool) {
        egx::set_last_error("egx_server_open needs a pool of at least one graph");
        return EGX_UNSUPPORTED_GRAPH;
    }
    *out = nullptr;
    std::unique_ptr<egx_server> sv(new (std::nothrow) egx_server());
    if (!sv) return EGX_FAILED_TO_ALLOCATE_HOST_MEM;
    sv->pool.assign(pool, pool + count);
    sv->info.resize((size_t)count);
    for (int32_t i = 0; i < count; ++i) {
        if (!pool[i]) return EGX_UNSUPPORTED_GRAPH;
        EGX_TRY(egx_graph_info_get(pool[i], &sv->info[(size_t)i]));
        if (sv->info[(size_t)i].n <= 0 || sv->info[(size_t)i].m <= 0) {
            egx::set_last_error("served graphs need nodes and arcs");
            return EGX_UNSUPPORTED_GRAPH;
        }
        if (pool[i]->g.device != pool[0]->g.device) {
            egx::set_last_error("served graphs must live on one device");
            return EGX_DEVICE_ERR;
        }
    }
    sv->device = pool[0]->g.device;
    sv->staged_free.resize((size_t)count);
    if (opts) sv->opts = *opts;
    if (sv->opts.max_sessions <= 0) sv->opts.max_sessions = 64;
    if (sv->opts.step_iters <= 0) sv->opts.step_iters = 4;
    if (sv->opts.hbm_budget_bytes <= 0) {
        size_t free_b = 0, total_b = 0;
        int prev = -1;
        (void)hipGetDevice(&prev);
        (void)hipSetDevice(sv->device);
        hipError_t e = hipMemGetInfo(&free_b, &total_b);
        if (prev >= 0) (void)hipSetDevice(prev);
        if (e != hipSuccess) {
            egx::set_last_error(hipGetErrorString(e));
            return EGX_DEVICE_ERR;
        }
        sv->opts.hbm_budget_bytes = (int64_t)(0.8 * (double)free_b);
    }
    sv->stats.budget_bytes = sv->opts.hbm_budget_bytes;
    for (const egx_graph_info& inf : sv->info)
        sv->stats.largest_session_bytes = std::max<int64_t>(sv->stats.largest_session_bytes, session_bytes(inf));
    egx_server* raw = sv.get();
    try {
        for (int i = 0; i < kDeliveryThreads; ++i) sv->deliverers.emplace_back([raw] { raw->deliver_loop(); });
        sv->worker = std::thread([raw] { raw->run(); });
    } catch (...) {
        egx::set_last_error("could not start the server threads");
        return EGX_UNKNOW_ERROR;  // the destructor stops whatever started
    }
    *out = sv.release();
    return EGX_SUCC;
}

int egx_server_submit(egx_server* sv, int32_t graph, const int32_t* seeds, int32_t nseeds, double alpha,
                      int32_t max_iter, double tol, double* PR, int64_t* request_id) {
    if (!sv || graph < 0 || graph >= (int32_t)sv->pool.size() || nseeds < 0 || (nseeds > 0 && !seeds)) {
        egx::set_last_error("egx_server_submit: bad arguments");
        return EGX_UNSUPPORTED_GRAPH;
    }
    const int32_t n = sv->info[(size_t)graph].n;
    for (int32_t i = 0; i < nseeds; ++i)
        if (seeds[i] < 0 || seeds[i] >= n) {
            egx::set_last_error("seed node out of range");
            return EGX_UNSUPPORTED_GRAPH;
        }
    auto r = std::make_shared<Request>();
    r->graph = graph;
    r->seeds.assign(seeds, seeds + nseeds);
    r->alpha = alpha;
    r->max_iter = max_iter;
    r->tol = tol;
    r->out = PR;
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        if (sv->stop) return EGX_UNKNOW_ERROR;
        r->id = sv->next_id++;
        sv->all[r->id] = r;
        sv->pending.push_back(r);
        sv->stats.submitted += 1;
    }
    sv->cv_work.notify_one();
    if (request_id) *request_id = r->id;
    return EGX_SUCC;
}

int egx_server_wait(egx_server* sv, int64_t request_id, double timeout_seconds, egx_request_info* info) {
    if (!sv) return EGX_UNSUPPORTED_GRAPH;
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->all.find(request_id);
    if (it == sv->all.end()) {
        egx::set_last_error("unknown or already collected request id");
        return EGX_UNSUPPORTED_GRAPH;
    }
    std::shared_ptr<Request> r = it->second;
    auto ready = [&] { return r->done; };
    if (timeout_seconds < 0) {
        sv->cv_done.wait(lk, ready);
    } else if (!sv->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_seconds), ready)) {
        egx::set_last_error("request still running");
        return EGX_UNKNOW_ERROR;
    }
    sv->all.erase(request_id);
    if (r->status != EGX_SUCC) egx::set_last_error(r->msg.empty() ? "request failed in the server" : r->msg);
    if (info) {
        info->iterations = r->iterations;
        info->width = r->width;
        info->queue_seconds = seconds_since(r->t_submit, r->t_start);
        info->service_seconds = seconds_since(r->t_start, r->t_done);
        info->latency_seconds = seconds_since(r->t_submit, r->t_done);
    }
    return r->status;
}

int egx_server_stats_get(egx_server* sv, egx_server_stats* out) {
    if (!sv || !out) return EGX_UNSUPPORTED_GRAPH;
    std::lock_guard<std::mutex> lk(sv->mu);
    *out = sv->stats;
    return EGX_SUCC;
}

void egx_server_close(egx_server* sv) { delete sv; }

}  // extern "C"
