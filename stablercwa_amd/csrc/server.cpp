// Request server with continuous batching over the plan C-ABI (srcwa.h).
//
// Callers submit requests (one stack, a sweep of 1..n points, planar / conical / crossed) from any
// thread.  A worker thread keeps up to `slots` batches on the GPU at once, each on its own HIP stream,
// all drawing device memory from one pool of the server's HBM budget:
//   * retire: a batch whose plan has finished (srcwa_plan_poll) has its results copied out and split
//     per request, and its memory and stream slot are released;
//   * admit: while a slot is free, the oldest pending request picks the kind of the next batch, and
//     pending requests of that kind join in arrival order while the batch stays within the point cap
//     and the requests' footprints (measured at submit with the dry-run sizing entry points) fit the
//     free pool memory.  A plan that still does not fit is split or waits for a retirement.
// Requests that arrive while batches run are admitted as soon as a slot and memory free up, so work of
// different requests and kinds overlaps on the device.  Each request's results are bitwise the same as
// solving it alone: every instance is computed by its own workgroups from its own inputs.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/srcwa.h"

namespace {

using Clock = std::chrono::steady_clock;
thread_local std::string t_error;

int fail(int code, const std::string& msg) {
    t_error = msg;
    return code;
}

struct Request {
    int64_t id = 0;
    int kind = 0, points = 0, M = 0, M2 = 0, pol = 0, factorization = 0;
    std::vector<double> stack, wl, th, ph;
    double jones[4] = {0, 0, 0, 0};
    int64_t orders = 0;          // orders per point
    int64_t footprint = 0;       // device bytes of a plan holding this request alone
    // results
    std::vector<double> R, T, DEr, DEt;
    std::vector<int> status, eig_info;
    int error = SRCWA_OK;
    std::string message;
    bool done = false;
    Clock::time_point t_submit, t_start, t_done;
};

int64_t r_words(int kind) { return kind == 1 ? 2 : 8; }     // doubles of R (and T) per order
int64_t de_words(int kind) { return kind == 2 ? 4 : 1; }    // doubles of DEr (and DEt) per order

struct Batch {
    srcwa_plan* plan = nullptr;
    int slot = 0;
    std::vector<Request*> reqs;
};

}  // namespace

struct srcwa_server {
    int device = 0;
    int max_points = 4096;
    int64_t pool_bytes = 0;
    srcwa_pool* pool = nullptr;
    std::vector<srcwa_context*> slots;
    std::vector<char> slot_busy;
    std::mutex lock;
    std::condition_variable wake, finished;
    std::deque<Request*> pending;
    std::unordered_map<int64_t, std::unique_ptr<Request>> requests;
    std::vector<Batch> inflight;
    int64_t next_id = 1;
    bool stop = false;
    std::thread worker;
    // statistics
    int64_t batches = 0, max_inflight = 0, points_done = 0, requests_done = 0;

    // Plan for a list of same-kind requests on a slot (no lock held).
    int create_plan(const std::vector<Request*>& reqs, int slot, srcwa_plan** out) {
        const int kind = reqs[0]->kind;
        std::vector<double> stacks;
        std::vector<int64_t> offsets;
        std::vector<int> idx, M, M2, pol, fz;
        std::vector<double> wl, th, ph, jones;
        for (size_t r = 0; r < reqs.size(); ++r) {
            const Request* q = reqs[r];
            offsets.push_back(static_cast<int64_t>(stacks.size()));
            stacks.insert(stacks.end(), q->stack.begin(), q->stack.end());
            for (int p = 0; p < q->points; ++p) {
                idx.push_back(static_cast<int>(r));
                wl.push_back(q->wl[p]);
                th.push_back(q->th[p]);
                ph.push_back(q->ph.empty() ? 0.0 : q->ph[p]);
                M.push_back(q->M);
                M2.push_back(q->M2);
                pol.push_back(q->pol);
                fz.push_back(q->factorization);
                jones.insert(jones.end(), q->jones, q->jones + 4);
            }
        }
        const int n = static_cast<int>(idx.size());
        const int ns = static_cast<int>(reqs.size());
        const int64_t budget = int64_t(1) << 62;    // one chunk: the pool bounds the footprint
        srcwa_context* ctx = slots[slot];
        if (kind == 1)
            return srcwa_plan1d_create(ctx, ns, stacks.data(), offsets.data(), n, idx.data(), wl.data(), th.data(),
                                       M.data(), pol.data(), budget, out);
        if (kind == 2)
            return srcwa_plan_conical_create(ctx, ns, stacks.data(), offsets.data(), n, idx.data(), wl.data(),
                                             th.data(), ph.data(), M.data(), budget, out);
        return srcwa_plan2d_create(ctx, ns, stacks.data(), offsets.data(), n, idx.data(), wl.data(), th.data(),
                                   ph.data(), M.data(), M2.data(), jones.data(), fz.data(), budget, out);
    }

    // Copies a finished batch's results into its requests (no lock held).
    void collect(Batch& b) {
        int64_t orders = 0;
        int points = 0;
        for (const Request* q : b.reqs) {
            orders += q->orders * q->points;
            points += q->points;
        }
        const int kind = b.reqs[0]->kind;
        std::vector<double> R(r_words(kind) * orders), T(r_words(kind) * orders), DEr(de_words(kind) * orders),
            DEt(de_words(kind) * orders);
        std::vector<int> st(points), info(points);
        const int rc = srcwa_plan_results(b.plan, R.data(), T.data(), DEr.data(), DEt.data(), st.data(), info.data());
        const std::string msg = rc == SRCWA_OK ? std::string() : std::string(srcwa_last_error());
        int64_t o = 0;
        int p = 0;
        for (Request* q : b.reqs) {
            const int64_t no = q->orders * q->points;
            if (rc != SRCWA_OK) {
                q->error = rc;
                q->message = msg;
            } else {
                q->R.assign(R.begin() + r_words(kind) * o, R.begin() + r_words(kind) * (o + no));
                q->T.assign(T.begin() + r_words(kind) * o, T.begin() + r_words(kind) * (o + no));
                q->DEr.assign(DEr.begin() + de_words(kind) * o, DEr.begin() + de_words(kind) * (o + no));
                q->DEt.assign(DEt.begin() + de_words(kind) * o, DEt.begin() + de_words(kind) * (o + no));
                q->status.assign(st.begin() + p, st.begin() + p + q->points);
                q->eig_info.assign(info.begin() + p, info.begin() + p + q->points);
            }
            o += no;
            p += q->points;
        }
        srcwa_plan_destroy(b.plan);
        b.plan = nullptr;
    }

    void finish(Request* q, int code, const std::string& msg) {
        q->error = code;
        q->message = msg;
        q->done = true;
        q->t_done = Clock::now();
    }

    void run() {
        std::unique_lock<std::mutex> lk(lock);
        for (;;) {
            // retire finished batches
            for (size_t b = 0; b < inflight.size();) {
                const int st = srcwa_plan_poll(inflight[b].plan);
                if (st == 0) {
                    ++b;
                    continue;
                }
                Batch done = std::move(inflight[b]);
                inflight.erase(inflight.begin() + static_cast<long>(b));
                lk.unlock();
                if (st < 0) {
                    const std::string msg = srcwa_last_error();
                    for (Request* q : done.reqs) {
                        q->error = st;
                        q->message = msg;
                    }
                    srcwa_plan_destroy(done.plan);
                } else {
                    collect(done);
                }
                lk.lock();
                const auto now = Clock::now();
                for (Request* q : done.reqs) {
                    q->done = true;
                    q->t_done = now;
                    points_done += q->points;
                    ++requests_done;
                }
                slot_busy[done.slot] = 0;
                finished.notify_all();
            }
            // admit pending requests into free slots
            bool blocked = false;
            while (!blocked && !pending.empty()) {
                int slot = -1;
                for (size_t s = 0; s < slots.size(); ++s)
                    if (!slot_busy[s]) {
                        slot = static_cast<int>(s);
                        break;
                    }
                if (slot < 0) break;
                const int kind = pending.front()->kind;
                int64_t room = srcwa_pool_free_bytes(pool);
                std::vector<Request*> batch;
                int points = 0;
                for (auto it = pending.begin(); it != pending.end();) {
                    Request* q = *it;
                    if (q->kind != kind) {
                        ++it;
                        continue;
                    }
                    if (!batch.empty() && (points + q->points > max_points || q->footprint > room)) break;
                    batch.push_back(q);
                    points += q->points;
                    room -= q->footprint;
                    it = pending.erase(it);
                }
                slot_busy[slot] = 1;
                lk.unlock();
                srcwa_plan* plan = nullptr;
                int rc = SRCWA_ENOMEM;
                std::vector<Request*> back;        // requests to return to the queue
                while (!batch.empty()) {
                    rc = create_plan(batch, slot, &plan);
                    if (rc != SRCWA_ENOMEM || batch.size() == 1) break;
                    const size_t keep = batch.size() / 2;
                    back.insert(back.begin(), batch.begin() + static_cast<long>(keep), batch.end());
                    batch.resize(keep);
                }
                if (rc == SRCWA_OK) {
                    const auto now = Clock::now();
                    for (Request* q : batch) q->t_start = now;
                    rc = srcwa_plan_execute(plan);
                }
                const std::string msg = rc == SRCWA_OK ? std::string() : std::string(srcwa_last_error());
                lk.lock();
                if (rc == SRCWA_OK) {
                    inflight.push_back(Batch{plan, slot, batch});
                    ++batches;
                    max_inflight = std::max<int64_t>(max_inflight, static_cast<int64_t>(inflight.size()));
                } else {
                    if (plan) srcwa_plan_destroy(plan);
                    slot_busy[slot] = 0;
                    if (rc == SRCWA_ENOMEM && !inflight.empty()) {
                        back.insert(back.begin(), batch.begin(), batch.end());   // wait for memory
                        blocked = true;
                    } else {
                        for (Request* q : batch) finish(q, rc, msg);
                        finished.notify_all();
                    }
                }
                for (auto it = back.rbegin(); it != back.rend(); ++it) pending.push_front(*it);
            }
            if (stop && pending.empty() && inflight.empty()) break;
            if (inflight.empty() && (pending.empty() || blocked)) {
                wake.wait(lk, [&] { return stop || !pending.empty(); });
            } else {
                lk.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(20));
                lk.lock();
            }
        }
    }
};

namespace {

int validate_and_size(srcwa_server* s, Request& q) {
    const int64_t off = 0;
    const int n = q.points;
    std::vector<int> idx(n, 0), M(n, q.M), M2(n, q.M2), pol(n, q.pol), fz(n, q.factorization);
    std::vector<double> ph = q.ph.empty() ? std::vector<double>(n, 0.0) : q.ph;
    std::vector<double> jones;
    for (int p = 0; p < n; ++p) jones.insert(jones.end(), q.jones, q.jones + 4);
    const int64_t budget = int64_t(1) << 62;
    srcwa_context* ctx = s->slots[0];
    if (q.kind == 1)
        return srcwa_plan1d_bytes(ctx, 1, q.stack.data(), &off, n, idx.data(), q.wl.data(), q.th.data(), M.data(),
                                  pol.data(), budget, &q.footprint);
    if (q.kind == 2)
        return srcwa_plan_conical_bytes(ctx, 1, q.stack.data(), &off, n, idx.data(), q.wl.data(), q.th.data(),
                                        ph.data(), M.data(), budget, &q.footprint);
    return srcwa_plan2d_bytes(ctx, 1, q.stack.data(), &off, n, idx.data(), q.wl.data(), q.th.data(), ph.data(),
                              M.data(), M2.data(), jones.data(), fz.data(), budget, &q.footprint);
}

}  // namespace

extern "C" {

const char* srcwa_server_last_error(void) { return t_error.c_str(); }

int srcwa_server_create(int device, int64_t hbm_bytes, int slots, int max_batch_points, srcwa_server** out) {
    if (!out || hbm_bytes <= 0 || slots < 1 || slots > 16 || max_batch_points < 1)
        return fail(SRCWA_EINVAL, "srcwa_server_create: invalid arguments");
    auto s = std::unique_ptr<srcwa_server>(new srcwa_server);
    s->device = device;
    s->max_points = max_batch_points;
    s->pool_bytes = hbm_bytes;
    int rc = srcwa_pool_create(device, hbm_bytes, &s->pool);
    if (rc != SRCWA_OK) return fail(rc, srcwa_last_error());
    for (int i = 0; i < slots; ++i) {
        srcwa_context* ctx = nullptr;
        rc = srcwa_context_create(device, &ctx);
        if (rc == SRCWA_OK) rc = srcwa_context_set_pool(ctx, s->pool);
        if (rc != SRCWA_OK) {
            const std::string msg = srcwa_last_error();
            if (ctx) srcwa_context_destroy(ctx);
            for (srcwa_context* c : s->slots) srcwa_context_destroy(c);
            srcwa_pool_destroy(s->pool);
            return fail(rc, msg);
        }
        s->slots.push_back(ctx);
    }
    s->slot_busy.assign(s->slots.size(), 0);
    srcwa_server* raw = s.get();
    s->worker = std::thread([raw] { raw->run(); });
    *out = s.release();
    return SRCWA_OK;
}

int srcwa_server_submit(srcwa_server* s, const srcwa_request* r, int64_t* id) {
    if (!s || !r || !id) return fail(SRCWA_EINVAL, "srcwa_server_submit: null argument");
    if (r->kind < 1 || r->kind > 3) return fail(SRCWA_EINVAL, "srcwa_server_submit: unknown request kind");
    if (r->n_points < 1 || !r->wavelength || !r->theta_deg || !r->stack || r->stack_len < 6)
        return fail(SRCWA_EINVAL, "srcwa_server_submit: a request needs a stack record and at least one point");
    if (r->kind != 1 && !r->phi_deg) return fail(SRCWA_EINVAL, "srcwa_server_submit: conical and crossed requests need phi");
    if (r->kind == 3 && !r->jones) return fail(SRCWA_EINVAL, "srcwa_server_submit: crossed requests need a Jones vector");
    auto q = std::unique_ptr<Request>(new Request);
    q->kind = r->kind;
    q->points = r->n_points;
    q->M = r->M;
    q->M2 = r->M2;
    q->pol = r->pol;
    q->factorization = r->factorization;
    q->stack.assign(r->stack, r->stack + r->stack_len);
    q->wl.assign(r->wavelength, r->wavelength + r->n_points);
    q->th.assign(r->theta_deg, r->theta_deg + r->n_points);
    if (r->kind != 1) q->ph.assign(r->phi_deg, r->phi_deg + r->n_points);
    if (r->kind == 3) std::memcpy(q->jones, r->jones, sizeof(q->jones));
    const int64_t n1 = 2 * int64_t(r->M) + 1;
    q->orders = r->kind == 3 ? n1 * (2 * int64_t(r->M2) + 1) : n1;
    const int rc = validate_and_size(s, *q);
    if (rc != SRCWA_OK) return fail(rc, srcwa_last_error());
    if (q->footprint > s->pool_bytes)
        return fail(SRCWA_ENOMEM, "srcwa_server_submit: the request needs more device memory than the server holds");
    q->t_submit = Clock::now();
    std::lock_guard<std::mutex> g(s->lock);
    if (s->stop) return fail(SRCWA_EINVAL, "srcwa_server_submit: server is shutting down");
    q->id = s->next_id++;
    *id = q->id;
    s->pending.push_back(q.get());
    s->requests[q->id] = std::move(q);
    s->wake.notify_one();
    return SRCWA_OK;
}

int srcwa_server_wait(srcwa_server* s, int64_t id, double timeout_s, int* done) {
    if (!s || !done) return fail(SRCWA_EINVAL, "srcwa_server_wait: null argument");
    std::unique_lock<std::mutex> lk(s->lock);
    auto it = s->requests.find(id);
    if (it == s->requests.end()) return fail(SRCWA_EINVAL, "srcwa_server_wait: unknown request");
    Request* q = it->second.get();
    if (timeout_s < 0) {
        s->finished.wait(lk, [&] { return q->done; });
    } else {
        s->finished.wait_for(lk, std::chrono::duration<double>(timeout_s), [&] { return q->done; });
    }
    *done = q->done ? 1 : 0;
    return SRCWA_OK;
}

int srcwa_server_result(srcwa_server* s, int64_t id, double* R, double* T, double* DEr, double* DEt, int* status,
                        int* eig_info, double* times) {
    if (!s) return fail(SRCWA_EINVAL, "srcwa_server_result: null server");
    std::lock_guard<std::mutex> g(s->lock);
    auto it = s->requests.find(id);
    if (it == s->requests.end()) return fail(SRCWA_EINVAL, "srcwa_server_result: unknown request");
    const Request& q = *it->second;
    if (!q.done) return fail(SRCWA_EINVAL, "srcwa_server_result: request still running");
    if (times) {
        using sec = std::chrono::duration<double>;
        times[0] = sec(q.t_done - q.t_submit).count();                           // latency
        times[1] = q.error ? 0.0 : sec(q.t_start - q.t_submit).count();          // queueing
    }
    if (q.error != SRCWA_OK) return fail(q.error, q.message);
    auto put = [](double* dst, const std::vector<double>& v) {
        if (dst) std::memcpy(dst, v.data(), v.size() * sizeof(double));
    };
    put(R, q.R);
    put(T, q.T);
    put(DEr, q.DEr);
    put(DEt, q.DEt);
    if (status) std::memcpy(status, q.status.data(), q.status.size() * sizeof(int));
    if (eig_info) std::memcpy(eig_info, q.eig_info.data(), q.eig_info.size() * sizeof(int));
    return SRCWA_OK;
}

int srcwa_server_release(srcwa_server* s, int64_t id) {
    if (!s) return fail(SRCWA_EINVAL, "srcwa_server_release: null server");
    std::lock_guard<std::mutex> g(s->lock);
    auto it = s->requests.find(id);
    if (it == s->requests.end()) return fail(SRCWA_EINVAL, "srcwa_server_release: unknown request");
    if (!it->second->done) return fail(SRCWA_EINVAL, "srcwa_server_release: request still running");
    s->requests.erase(it);
    return SRCWA_OK;
}

int srcwa_server_stats(srcwa_server* s, int64_t* stats4) {
    if (!s || !stats4) return fail(SRCWA_EINVAL, "srcwa_server_stats: null argument");
    std::lock_guard<std::mutex> g(s->lock);
    stats4[0] = s->batches;
    stats4[1] = s->max_inflight;
    stats4[2] = s->requests_done;
    stats4[3] = s->points_done;
    return SRCWA_OK;
}

void srcwa_server_destroy(srcwa_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->lock);
        s->stop = true;
        s->wake.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    s->requests.clear();
    for (srcwa_context* c : s->slots) srcwa_context_destroy(c);
    srcwa_pool_destroy(s->pool);
    delete s;
}

}  // extern "C"
