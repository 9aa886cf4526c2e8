// Request server: continuous batching of slide requests from many streams.
//
// Every admitted stream keeps its window, grid and union-find state resident
// in HBM. Slide requests queue per stream (executed in submission order). A
// set of lanes - each a batch object with its own HIP stream and host worker
// thread - repeatedly takes the oldest ready request of up to max_batch
// streams that are not on the device, slides them together in one launch per
// phase, copies the labels back and completes the requests. Requests that
// arrive while lanes are busy join the next batch a lane forms, so work from
// different lanes overlaps on the device. Streams are admitted while the sum
// of their device bytes fits the HBM budget.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "rtdbs.h"
#include "rtdbs_dev.h"
#include "rtdbs_host.h"

using namespace rtdbs;

namespace {

using Clock = std::chrono::steady_clock;

enum ReqState { kQueued = 0, kRunning = 1, kDone = 2 };

struct Request {
    int64_t ticket = 0;
    int32_t sid = 0;
    const float *in = nullptr;
    int32_t *label = nullptr;
    int32_t *cid = nullptr;
    int state = kQueued;
    int status = RTDBS_OK;
    std::string error;  // the failing call's message when status != RTDBS_OK
    Clock::time_point t_submit, t_done;
};

struct SStream {
    int32_t id = 0;
    rtdbs_params params{};
    rtdbs_stream *s = nullptr;  // null while evicted
    size_t bytes = 0;
    std::deque<Request *> queue;
    bool busy = false;
    bool closing = false;
    uint64_t last_used = 0;  // server clock at the last completed slide (or admission)
    // checkpoint while evicted
    std::vector<float> ckpt;
    int64_t ckpt_slides = 0;
};

}  // namespace

struct rtdbs_server {
    rtdbs_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_work;   // lanes wait for ready requests
    std::condition_variable cv_done;   // waiters on tickets / idle streams
    std::condition_variable cv_budget; // admission waiters
    bool stopping = false;
    size_t bytes_in_use = 0;
    int32_t next_id = 1;
    int64_t next_ticket = 1;
    std::map<int32_t, SStream *> streams;
    std::map<int64_t, Request *> tickets;
    std::deque<Request *> arrivals;  // global FIFO of queued requests (oldest first)
    std::vector<std::thread> workers;
    std::vector<rtdbs_batch *> lanes;
    int device = 0;
    uint64_t clock = 0;  // advanced per completed batch and admission (recency for swapping)
    // stats
    int64_t completed = 0, failed = 0, batches = 0, batched_streams = 0, swap_ins = 0, swap_outs = 0;
    std::vector<double> lat_ms;
};

namespace {

// Picks up to max_batch ready streams in arrival order (caller holds mu).
std::vector<Request *> take_ready(rtdbs_server *srv) {
    std::vector<Request *> out;
    const size_t cap = (size_t)std::max(1, srv->cfg.max_batch);
    for (auto it = srv->arrivals.begin(); it != srv->arrivals.end() && out.size() < cap;) {
        Request *r = *it;
        auto sit = srv->streams.find(r->sid);
        if (sit == srv->streams.end()) {
            it = srv->arrivals.erase(it);
            continue;
        }
        SStream *ss = sit->second;
        if (ss->busy || !ss->s || ss->queue.empty() || ss->queue.front() != r) {
            ++it;
            continue;
        }
        ss->busy = true;
        r->state = kRunning;
        out.push_back(r);
        it = srv->arrivals.erase(it);
    }
    return out;
}

// A queued request whose stream is ready to slide but not on the device
// (caller holds mu).
SStream *ready_evicted(rtdbs_server *srv) {
    for (Request *r : srv->arrivals) {
        auto it = srv->streams.find(r->sid);
        if (it == srv->streams.end()) continue;
        SStream *ss = it->second;
        if (!ss->s && !ss->busy && !ss->closing && !ss->queue.empty() && ss->queue.front() == r) return ss;
    }
    return nullptr;
}

// Swapping (cfg.swap): brings the stream of the oldest ready request back on
// the device, checkpointing the least recently used idle streams to host
// memory until it fits the budget. Victims are destroyed before the stream
// is rebuilt, so device memory never exceeds the budget; the accounting is
// reserved up front under the lock, so lanes swapping at the same time never
// over-admit. Returns true if it swapped a stream in. Caller holds mu via lk.
bool swap_in_one(rtdbs_server *srv, std::unique_lock<std::mutex> &lk) {
    SStream *ss = ready_evicted(srv);
    if (!ss) return false;
    std::vector<SStream *> victims;
    size_t freed = 0;
    while (srv->bytes_in_use - freed + ss->bytes > srv->cfg.hbm_budget_bytes) {
        SStream *v = nullptr;
        for (auto &kv : srv->streams) {
            SStream *c = kv.second;
            if (!c->s || c->busy || c->closing || !c->queue.empty()) continue;
            if (!v || c->last_used < v->last_used) v = c;
        }
        if (!v) break;
        v->busy = true;
        victims.push_back(v);
        freed += v->bytes;
    }
    if (srv->bytes_in_use - freed + ss->bytes > srv->cfg.hbm_budget_bytes) {
        for (SStream *v : victims) v->busy = false;  // not now: wait for work to finish
        return false;
    }
    ss->busy = true;
    srv->bytes_in_use = srv->bytes_in_use - freed + ss->bytes;
    lk.unlock();
    size_t kept = 0;  // victims whose checkpoint failed stay on the device
    std::string why;  // the first failed checkpoint's message
    for (SStream *v : victims) {
        std::vector<float> win((size_t)v->s->p.window * v->s->p.dim);
        const int rc = rtdbs_stream_get_window(v->s, win.data());
        if (rc != RTDBS_OK && why.empty()) why = rtdbs_last_error();
        const int64_t slides = v->s->slides;
        if (rc == RTDBS_OK) rtdbs_stream_destroy(v->s);
        std::lock_guard<std::mutex> g(srv->mu);
        if (rc == RTDBS_OK) {
            v->ckpt.swap(win);
            v->ckpt_slides = slides;
            v->s = nullptr;
            srv->swap_outs++;
        } else {
            kept += v->bytes;
        }
        v->busy = false;
    }
    rtdbs_stream *st = nullptr;
    int rc = kept ? fail(RTDBS_ENOMEM, "a stream could not be checkpointed to make room (%s)", why.c_str()) : RTDBS_OK;
    if (!rc) rc = rtdbs_stream_create(&ss->params, &st);
    if (!rc) rc = rtdbs_stream_restore(st, ss->ckpt.data(), ss->ckpt_slides);
    lk.lock();
    srv->bytes_in_use += kept;
    if (rc == RTDBS_OK) {
        ss->s = st;
        ss->ckpt.clear();
        ss->ckpt.shrink_to_fit();
        ss->last_used = ++srv->clock;
        srv->swap_ins++;
    } else {
        if (st) rtdbs_stream_destroy(st);
        srv->bytes_in_use -= ss->bytes;
        // the request that asked for the stream fails; later ones retry
        Request *r = ss->queue.front();
        ss->queue.pop_front();
        srv->arrivals.erase(std::remove(srv->arrivals.begin(), srv->arrivals.end(), r), srv->arrivals.end());
        r->state = kDone;
        r->status = rc;
        r->error = rtdbs_last_error();  // this thread's message of the failing call
        r->t_done = Clock::now();
        srv->failed++;
    }
    ss->busy = false;
    srv->cv_done.notify_all();
    srv->cv_budget.notify_all();
    return true;
}

void lane_main(rtdbs_server *srv, int lane) {
    (void)hipSetDevice(srv->device);
    rtdbs_batch *b = srv->lanes[lane];
    std::vector<rtdbs_stream *> members;
    std::vector<Request *> reqs;
    std::vector<const float *> ins;
    std::vector<int32_t *> labels, cids;
    while (true) {
        {
            std::unique_lock<std::mutex> lk(srv->mu);
            srv->cv_work.wait(lk, [&] {
                if (srv->stopping) return true;
                for (Request *r : srv->arrivals) {
                    auto it = srv->streams.find(r->sid);
                    if (it != srv->streams.end() && !it->second->busy && (it->second->s || srv->cfg.swap) &&
                        !it->second->queue.empty() && it->second->queue.front() == r)
                        return true;
                }
                return false;
            });
            if (srv->stopping) return;
            bool swapped = false;
            if (srv->cfg.swap) swapped = swap_in_one(srv, lk);
            if (srv->stopping) return;
            reqs = take_ready(srv);
            if (reqs.empty()) {
                // an evicted stream waits for room that only running work frees
                if (!swapped) srv->cv_work.wait_for(lk, std::chrono::milliseconds(1));
                continue;
            }
            members.clear();
            for (Request *r : reqs) members.push_back(srv->streams[r->sid]->s);
        }
        int rc = batch_set_members(b, members.data(), (int)members.size());
        if (rc == RTDBS_OK) {
            ins.clear();
            labels.clear();
            cids.clear();
            for (Request *r : reqs) {
                ins.push_back(r->in);
                labels.push_back(r->label);
                cids.push_back(r->cid);
            }
            // strides up, one launch per phase, labels down (merged copies;
            // pinned staging unless the caller's buffers are engine-pinned)
            rc = batch_slide_host(b, ins.data(), labels.data(), cids.data());
        }
        const auto now = Clock::now();
        const std::string err = rc == RTDBS_OK ? std::string() : std::string(rtdbs_last_error());
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            srv->batches++;
            srv->batched_streams += (int64_t)reqs.size();
            const uint64_t now_clock = ++srv->clock;
            for (Request *r : reqs) {
                SStream *ss = srv->streams[r->sid];
                ss->queue.pop_front();
                ss->busy = false;
                ss->last_used = now_clock;
                r->state = kDone;
                r->status = rc;
                r->error = err;
                r->t_done = now;
                if (rc == RTDBS_OK) {
                    srv->completed++;
                    srv->lat_ms.push_back(std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count());
                } else {
                    srv->failed++;
                }
            }
        }
        srv->cv_done.notify_all();
        srv->cv_work.notify_all();
    }
}

int admit_bytes(rtdbs_server *srv, std::unique_lock<std::mutex> &lk, size_t bytes) {
    if (bytes > srv->cfg.hbm_budget_bytes)
        return fail(RTDBS_EBUDGET, "stream needs %zu bytes, budget is %zu", bytes, srv->cfg.hbm_budget_bytes);
    const auto deadline = Clock::now() + std::chrono::milliseconds(std::max(0, srv->cfg.admit_wait_ms));
    while (srv->bytes_in_use + bytes > srv->cfg.hbm_budget_bytes) {
        if (srv->cfg.admit_wait_ms <= 0 ||
            srv->cv_budget.wait_until(lk, deadline) == std::cv_status::timeout) {
            if (srv->bytes_in_use + bytes <= srv->cfg.hbm_budget_bytes) break;
            return fail(RTDBS_EBUDGET, "HBM budget exhausted: %zu in use + %zu requested > %zu", srv->bytes_in_use,
                        bytes, srv->cfg.hbm_budget_bytes);
        }
    }
    srv->bytes_in_use += bytes;
    return RTDBS_OK;
}

}  // namespace

extern "C" int rtdbs_server_create(const rtdbs_server_config *cfg, rtdbs_server **out) {
    if (!cfg || !out) return fail(RTDBS_EINVAL, "NULL argument");
    *out = nullptr;
    if (cfg->max_batch < 1 || cfg->max_batch > kMaxBatch)
        return fail(RTDBS_EINVAL, "max_batch must be in [1, %d]", kMaxBatch);
    if (cfg->lanes < 1 || cfg->lanes > 8) return fail(RTDBS_EINVAL, "lanes must be in [1, 8]");
    rtdbs_server *srv = new (std::nothrow) rtdbs_server();
    if (!srv) return fail(RTDBS_ENOMEM, "host allocation failed");
    srv->cfg = *cfg;
    if (hipGetDevice(&srv->device) != hipSuccess) srv->device = 0;
    for (int i = 0; i < cfg->lanes; i++) {
        int rc = 0;
        rtdbs_batch *b = batch_new(cfg->max_batch, cfg->use_graph != 0, &rc);
        if (!b) {
            for (auto *x : srv->lanes) batch_free(x);
            delete srv;
            return rc;
        }
        srv->lanes.push_back(b);
    }
    for (int i = 0; i < cfg->lanes; i++) srv->workers.emplace_back(lane_main, srv, i);
    *out = srv;
    return RTDBS_OK;
}

extern "C" int rtdbs_server_open_stream(rtdbs_server *srv, const rtdbs_params *params, const float *first_window,
                                        int32_t *stream_id) {
    if (!srv || !params || !first_window || !stream_id) return fail(RTDBS_EINVAL, "NULL argument");
    int rc = validate_params(params);
    if (rc) return rc;
    const size_t bytes = stream_device_bytes(*params);
    if (srv->cfg.swap) {
        if (bytes > srv->cfg.hbm_budget_bytes)
            return fail(RTDBS_EBUDGET, "stream needs %zu bytes, budget is %zu", bytes, srv->cfg.hbm_budget_bytes);
        std::unique_lock<std::mutex> lk(srv->mu);
        if (srv->bytes_in_use + bytes > srv->cfg.hbm_budget_bytes) {
            // no room now: a checkpoint of the first window, swapped in on demand
            SStream *ss = new SStream();
            ss->id = srv->next_id++;
            ss->params = *params;
            ss->bytes = bytes;
            ss->ckpt.assign(first_window, first_window + (size_t)params->window * params->dim);
            ss->ckpt_slides = 0;
            srv->streams[ss->id] = ss;
            *stream_id = ss->id;
            return RTDBS_OK;
        }
        srv->bytes_in_use += bytes;
    } else {
        std::unique_lock<std::mutex> lk(srv->mu);
        rc = admit_bytes(srv, lk, bytes);
        if (rc) return rc;
    }
    rtdbs_stream *s = nullptr;
    rc = rtdbs_stream_create(params, &s);
    if (!rc) rc = rtdbs_stream_init(s, first_window);
    if (rc) {
        std::string msg = rtdbs_last_error();
        if (s) rtdbs_stream_destroy(s);
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->bytes_in_use -= bytes;
        srv->cv_budget.notify_all();
        return fail(rc, "%s", msg.c_str());
    }
    std::lock_guard<std::mutex> lk(srv->mu);
    SStream *ss = new SStream();
    ss->id = srv->next_id++;
    ss->params = *params;
    ss->s = s;
    ss->bytes = bytes;
    ss->last_used = ++srv->clock;
    srv->streams[ss->id] = ss;
    *stream_id = ss->id;
    return RTDBS_OK;
}

static void cancel_queued(rtdbs_server *srv, SStream *ss) {
    for (auto it = ss->queue.begin(); it != ss->queue.end();) {
        Request *r = *it;
        if (r->state == kQueued) {
            r->state = kDone;
            r->status = RTDBS_ECANCELLED;
            r->t_done = Clock::now();
            srv->arrivals.erase(std::remove(srv->arrivals.begin(), srv->arrivals.end(), r), srv->arrivals.end());
            it = ss->queue.erase(it);
        } else {
            ++it;
        }
    }
}

extern "C" int rtdbs_server_close_stream(rtdbs_server *srv, int32_t stream_id) {
    if (!srv) return fail(RTDBS_EINVAL, "server is NULL");
    SStream *ss = nullptr;
    {
        std::unique_lock<std::mutex> lk(srv->mu);
        auto it = srv->streams.find(stream_id);
        if (it == srv->streams.end()) return fail(RTDBS_ENOTFOUND, "unknown stream %d", stream_id);
        ss = it->second;
        if (ss->closing) return fail(RTDBS_ENOTFOUND, "stream %d is being closed", stream_id);
        ss->closing = true;
        cancel_queued(srv, ss);
        // a lane, an evict or a readmit may hold the stream
        srv->cv_done.wait(lk, [&] { return !ss->busy; });
        srv->streams.erase(stream_id);
        if (ss->s) srv->bytes_in_use -= ss->bytes;
    }
    srv->cv_done.notify_all();
    if (ss->s) rtdbs_stream_destroy(ss->s);
    delete ss;
    srv->cv_budget.notify_all();
    return RTDBS_OK;
}

extern "C" int rtdbs_server_evict_stream(rtdbs_server *srv, int32_t stream_id) {
    if (!srv) return fail(RTDBS_EINVAL, "server is NULL");
    rtdbs_stream *s = nullptr;
    SStream *ss = nullptr;
    {
        std::unique_lock<std::mutex> lk(srv->mu);
        auto it = srv->streams.find(stream_id);
        if (it == srv->streams.end()) return fail(RTDBS_ENOTFOUND, "unknown stream %d", stream_id);
        ss = it->second;
        if (ss->closing) return fail(RTDBS_ENOTFOUND, "stream %d is being closed", stream_id);
        srv->cv_done.wait(lk, [&] { return !ss->busy; });
        if (ss->closing) return fail(RTDBS_ENOTFOUND, "stream %d is being closed", stream_id);
        if (!ss->s) return RTDBS_OK;  // already evicted
        ss->busy = true;  // keeps lanes, close, evict and readmit away while checkpointing
        s = ss->s;
    }
    std::vector<float> win((size_t)s->p.window * s->p.dim);
    int rc = rtdbs_stream_get_window(s, win.data());
    const int64_t slides = s->slides;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        ss->busy = false;
        if (rc == RTDBS_OK) {
            ss->ckpt.swap(win);
            ss->ckpt_slides = slides;
            ss->s = nullptr;
            srv->bytes_in_use -= ss->bytes;
            srv->swap_outs++;
        }
    }
    if (rc == RTDBS_OK) rtdbs_stream_destroy(s);
    srv->cv_done.notify_all();  // close / evict / readmit wait for !busy
    srv->cv_budget.notify_all();
    srv->cv_work.notify_all();
    return rc;
}

extern "C" int rtdbs_server_readmit_stream(rtdbs_server *srv, int32_t stream_id) {
    if (!srv) return fail(RTDBS_EINVAL, "server is NULL");
    SStream *ss = nullptr;
    {
        std::unique_lock<std::mutex> lk(srv->mu);
        auto it = srv->streams.find(stream_id);
        if (it == srv->streams.end()) return fail(RTDBS_ENOTFOUND, "unknown stream %d", stream_id);
        ss = it->second;
        if (ss->closing) return fail(RTDBS_ENOTFOUND, "stream %d is being closed", stream_id);
        srv->cv_done.wait(lk, [&] { return !ss->busy; });
        if (ss->closing) return fail(RTDBS_ENOTFOUND, "stream %d is being closed", stream_id);
        if (ss->s) return RTDBS_OK;
        // held from here on: a close waits for the rebuild, a concurrent
        // readmit or evict waits and then sees the rebuilt stream
        ss->busy = true;
        int rc = admit_bytes(srv, lk, ss->bytes);
        if (rc) {
            ss->busy = false;
            lk.unlock();
            srv->cv_done.notify_all();
            return rc;
        }
    }
    rtdbs_stream *s = nullptr;
    int rc = rtdbs_stream_create(&ss->params, &s);
    if (!rc) rc = rtdbs_stream_restore(s, ss->ckpt.data(), ss->ckpt_slides);
    if (rc) {
        std::string msg = rtdbs_last_error();
        if (s) rtdbs_stream_destroy(s);
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            srv->bytes_in_use -= ss->bytes;
            ss->busy = false;
        }
        srv->cv_done.notify_all();
        srv->cv_budget.notify_all();
        return fail(rc, "%s", msg.c_str());
    }
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        ss->s = s;
        ss->busy = false;
        ss->last_used = ++srv->clock;
        ss->ckpt.clear();
        ss->ckpt.shrink_to_fit();
        srv->swap_ins++;
    }
    srv->cv_done.notify_all();
    srv->cv_work.notify_all();
    return RTDBS_OK;
}

extern "C" int rtdbs_server_submit(rtdbs_server *srv, int32_t stream_id, const float *stride_points, int32_t *label,
                                   int32_t *cid, int64_t *ticket) {
    if (!srv || !stride_points || !label || !cid || !ticket) return fail(RTDBS_EINVAL, "NULL argument");
    Request *r = new (std::nothrow) Request();
    if (!r) return fail(RTDBS_ENOMEM, "host allocation failed");
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        auto it = srv->streams.find(stream_id);
        if (it == srv->streams.end() || it->second->closing) {
            delete r;
            return fail(RTDBS_ENOTFOUND, "unknown stream %d", stream_id);
        }
        r->ticket = srv->next_ticket++;
        r->sid = stream_id;
        r->in = stride_points;
        r->label = label;
        r->cid = cid;
        r->t_submit = Clock::now();
        it->second->queue.push_back(r);
        srv->arrivals.push_back(r);
        srv->tickets[r->ticket] = r;
        *ticket = r->ticket;
    }
    srv->cv_work.notify_one();
    return RTDBS_OK;
}

extern "C" int rtdbs_server_wait(rtdbs_server *srv, int64_t ticket, int32_t timeout_ms) {
    if (!srv) return fail(RTDBS_EINVAL, "server is NULL");
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->tickets.find(ticket);
    if (it == srv->tickets.end()) return fail(RTDBS_ENOTFOUND, "unknown ticket %lld", (long long)ticket);
    Request *r = it->second;
    auto done = [&] { return r->state == kDone; };
    if (timeout_ms < 0) {
        srv->cv_done.wait(lk, done);
    } else if (!srv->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), done)) {
        return fail(RTDBS_ETIMEOUT, "ticket %lld not done after %d ms", (long long)ticket, timeout_ms);
    }
    const int st = r->status;
    const std::string err = r->error;
    srv->tickets.erase(it);
    delete r;
    if (st) return fail(st, "request failed with status %d%s%s", st, err.empty() ? "" : ": ", err.c_str());
    return RTDBS_OK;
}

extern "C" int rtdbs_server_poll(rtdbs_server *srv, int64_t ticket) {
    if (!srv) return fail(RTDBS_EINVAL, "server is NULL");
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->tickets.find(ticket);
    if (it == srv->tickets.end()) return fail(RTDBS_ENOTFOUND, "unknown ticket %lld", (long long)ticket);
    Request *r = it->second;
    if (r->state != kDone) return 0;
    const int st = r->status;
    const std::string err = r->error;
    srv->tickets.erase(it);
    delete r;
    if (st) return fail(st, "request failed with status %d%s%s", st, err.empty() ? "" : ": ", err.c_str());
    return 1;
}

extern "C" int rtdbs_server_cancel(rtdbs_server *srv, int64_t ticket) {
    if (!srv) return fail(RTDBS_EINVAL, "server is NULL");
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        auto it = srv->tickets.find(ticket);
        if (it == srv->tickets.end()) return fail(RTDBS_ENOTFOUND, "unknown ticket %lld", (long long)ticket);
        Request *r = it->second;
        if (r->state != kQueued) return fail(RTDBS_ESTATE, "ticket %lld is already on the device or done", (long long)ticket);
        auto sit = srv->streams.find(r->sid);
        if (sit != srv->streams.end()) {
            auto &q = sit->second->queue;
            q.erase(std::remove(q.begin(), q.end(), r), q.end());
        }
        srv->arrivals.erase(std::remove(srv->arrivals.begin(), srv->arrivals.end(), r), srv->arrivals.end());
        r->state = kDone;
        r->status = RTDBS_ECANCELLED;
        r->t_done = Clock::now();
    }
    srv->cv_done.notify_all();
    srv->cv_work.notify_all();
    return RTDBS_OK;
}

extern "C" int rtdbs_server_get_stats(rtdbs_server *srv, rtdbs_server_stats *st) {
    if (!srv || !st) return fail(RTDBS_EINVAL, "NULL argument");
    std::lock_guard<std::mutex> lk(srv->mu);
    std::memset(st, 0, sizeof(*st));
    st->completed = srv->completed;
    st->failed = srv->failed;
    st->batches = srv->batches;
    st->mean_batch = srv->batches ? (double)srv->batched_streams / (double)srv->batches : 0.0;
    st->bytes_in_use = srv->bytes_in_use;
    st->streams = (int32_t)srv->streams.size();
    for (auto &kv : srv->streams) st->resident += kv.second->s ? 1 : 0;
    st->swap_ins = srv->swap_ins;
    st->swap_outs = srv->swap_outs;
    if (!srv->lat_ms.empty()) {
        std::vector<double> v = srv->lat_ms;
        std::sort(v.begin(), v.end());
        auto pct = [&](double q) {
            size_t i = (size_t)std::min<double>((double)v.size() - 1, std::ceil(q * (double)v.size()) - 1);
            return v[i];
        };
        st->p50_ms = pct(0.50);
        st->p99_ms = pct(0.99);
        st->max_ms = v.back();
    }
    return RTDBS_OK;
}

extern "C" int rtdbs_server_reset_stats(rtdbs_server *srv) {
    if (!srv) return fail(RTDBS_EINVAL, "server is NULL");
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->completed = srv->failed = srv->batches = srv->batched_streams = srv->swap_ins = srv->swap_outs = 0;
    srv->lat_ms.clear();
    return RTDBS_OK;
}

extern "C" void rtdbs_server_destroy(rtdbs_server *srv) {
    if (!srv) return;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stopping = true;
        for (auto &kv : srv->streams) cancel_queued(srv, kv.second);
    }
    srv->cv_work.notify_all();
    srv->cv_done.notify_all();
    for (auto &t : srv->workers) t.join();
    for (auto *b : srv->lanes) batch_free(b);
    for (auto &kv : srv->streams) {
        if (kv.second->s) rtdbs_stream_destroy(kv.second->s);
        delete kv.second;
    }
    for (auto &kv : srv->tickets) delete kv.second;
    delete srv;
}
