/*
 * sa_serve.cpp -- request loop with continuous batching on one device.
 *
 * Clients submit single reads (inputs are copied); a scheduler thread forms
 * "waves" from the pending requests -- same model, parameters, precision and length
 * class, in arrival order, the class set by the longest request that arrived within
 * the batching delay of the oldest -- admitting each read only while the band-matrix
 * bytes of all waves in flight stay within the HBM budget.  Each wave runs on one of several
 * lanes (a context with one stream and a batch whose device buffers are reused),
 * so waves overlap on the device: a request that arrives while others run is
 * admitted as soon as a lane and memory free up.  Model tables stay resident.
 *
 * A read's result does not depend on the wave it ran in: every sub-alignment's
 * arithmetic is fixed by its own inputs (kernel shapes only change which lanes of
 * a wavefront hold which cells, and the cell arithmetic and its order are the same
 * in every shape).
 */
#include "signalalign_amd.h"
#include "sa_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
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

namespace {

struct Request {
    int64_t ticket = 0;
    const sa_model *model = nullptr;
    sa_params params{};
    sa_precision precision = SA_FP32;
    /* owned inputs */
    std::string ref;
    std::vector<double> events;
    int64_t n_events = 0, stride = 4;
    std::vector<int64_t> anchors;
    double scale = 1, shift = 0, var = 1;
    int32_t rl = 1, rr = 1;
    std::string options; /* degenerate base options of an 'X' reference, else empty */
    int64_t est_bytes = 0;
    int64_t n_diag = 0; /* lX + lY: the length of the forward sweep's dependency chain */
    /* outputs: the wave's result, shared by its requests */
    int status = SA_PENDING;
    std::shared_ptr<sa_result> wres;
    int64_t widx = 0;
    sa_request_timing timing{};
};

struct Lane {
    sa_context *ctx = nullptr;
    sa_batch *batch = nullptr;
    std::thread th;
    std::vector<Request *> wave;
    int64_t wave_bytes = 0;
    int64_t wave_id = 0;
    bool busy = false;
};

bool same_params(const sa_params &a, const sa_params &b) {
    return a.threshold == b.threshold && a.min_diags_between_traceback == b.min_diags_between_traceback &&
           a.traceback_diagonals == b.traceback_diagonals && a.diagonal_expansion == b.diagonal_expansion &&
           a.split_matrix_bigger_than_this == b.split_matrix_bigger_than_this;
}

} // namespace

struct sa_server {
    sa_server_config cfg{};
    int device = 0;
    std::chrono::steady_clock::time_point t0;
    std::vector<std::unique_ptr<Lane>> lanes;
    std::mutex mu;
    std::condition_variable cv_sched, cv_lane, cv_done;
    std::deque<Request *> pending;
    std::unordered_map<int64_t, Request *> reqs;
    int64_t next_ticket = 1;
    int64_t inflight_bytes = 0;
    int64_t next_wave = 1;
    bool stop = false;
    std::thread scheduler;
    sa_server_stats stats{};

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    void schedule_loop();
    void lane_loop(Lane *L);
};

void sa_server::schedule_loop() {
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        Lane *idle = nullptr;
        Request *first = nullptr;
        double due = -1.0;
        auto fits_now = [&](const Request *q) {
            return inflight_bytes == 0 || inflight_bytes + q->est_bytes <= cfg.hbm_budget_bytes;
        };
        auto ready = [&] {
            due = -1.0;
            if (stop) return true;
            if (pending.empty()) return false;
            idle = nullptr;
            bool others = false;
            for (auto &l : lanes) {
                if (!l->busy) { if (!idle) idle = l.get(); }
                else others = true;
            }
            if (!idle) return false;
            /* the wave's first request sets its model and length class: the longest of
               the requests that arrived within the batching delay of the oldest (a burst
               starts its long waves first and fills in behind them with short ones; the
               oldest waits at most one wave per length class among them) */
            first = pending.front();
            const double win = first->timing.submit_s + 1e-3 * std::max(cfg.batch_delay_ms, 0.0);
            Request *lp = first;
            for (Request *q : pending) {
                if (q->timing.submit_s > win) break;
                if (q->n_diag > lp->n_diag) lp = q;
            }
            /* memory: the first request must fit beside the waves in flight (alone, it is
               admitted once nothing else is in flight) */
            if (fits_now(lp)) first = lp;
            if (!fits_now(first)) return false;
            /* continuous batching: an idle device takes the wave at once; a busy one lets
               it gather requests until the oldest has waited batch_delay_ms */
            if (others && cfg.batch_delay_ms > 0.0 && (int64_t)pending.size() < cfg.max_wave_reads) {
                const double d = pending.front()->timing.submit_s + 1e-3 * cfg.batch_delay_ms;
                if (now() < d) {
                    due = d;
                    return false;
                }
            }
            return true;
        };
        while (!ready()) {
            if (due > 0.0)
                cv_sched.wait_until(lk, t0 + std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                                                 std::chrono::duration<double>(due)));
            else
                cv_sched.wait(lk);
        }
        if (stop) return;
        std::vector<Request *> wave;
        int64_t bytes = 0;
        for (auto it = pending.begin(); it != pending.end() && (int64_t)wave.size() < cfg.max_wave_reads;) {
            Request *q = *it;
            const bool fits = inflight_bytes + bytes + q->est_bytes <= cfg.hbm_budget_bytes || (q == first && wave.empty());
            /* a wave finishes with its longest read: keep reads of one length class
               (within length_ratio of the first request) together */
            const double lr = cfg.length_ratio;
            const bool same_class = lr < 0.0 || ((double)q->n_diag * lr >= (double)first->n_diag &&
                                                 (double)q->n_diag <= lr * (double)first->n_diag);
            if (q->model == first->model && q->precision == first->precision && q->options == first->options &&
                same_params(q->params, first->params) && same_class && fits) {
                wave.push_back(q);
                bytes += q->est_bytes;
                it = pending.erase(it);
            } else {
                ++it;
            }
        }
        const double t = now();
        const int64_t id = next_wave++;
        for (Request *q : wave) {
            q->timing.admit_s = t;
            q->timing.wave = id;
            q->timing.wave_reads = (int32_t)wave.size();
        }
        idle->wave = std::move(wave);
        idle->wave_bytes = bytes;
        idle->wave_id = id;
        idle->busy = true;
        inflight_bytes += bytes;
        cv_lane.notify_all();
    }
}

void sa_server::lane_loop(Lane *L) {
    (void)hipSetDevice(device);
    for (;;) {
        std::vector<Request *> wave;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_lane.wait(lk, [&] { return stop || (L->busy && !L->wave.empty()); });
            if (stop) return;
            wave = L->wave;
        }
        std::vector<sa_read> rd(wave.size());
        for (size_t i = 0; i < wave.size(); i++) {
            const Request *q = wave[i];
            rd[i].ref = q->ref.data();
            rd[i].ref_length = (int64_t)q->ref.size();
            rd[i].events = q->events.data();
            rd[i].n_events = q->n_events;
            rd[i].event_stride = q->stride;
            rd[i].anchors = q->anchors.data();
            rd[i].n_anchors = (int64_t)q->anchors.size() / 2;
            rd[i].scale = q->scale;
            rd[i].shift = q->shift;
            rd[i].var = q->var;
            rd[i].ragged_left = q->rl;
            rd[i].ragged_right = q->rr;
            rd[i].degenerate_options = q->options.empty() ? nullptr : q->options.c_str();
        }
        const Request *f = wave.front();
        int64_t cells = 0;
        double dev_ms = 0.0;
        /* runs reads [i0, i1) of the wave as one batch */
        double fill_ms = 0.0, run_ms = 0.0;
        auto run = [&](size_t i0, size_t i1) -> int {
            const int64_t n = (int64_t)(i1 - i0);
            const auto t0 = std::chrono::steady_clock::now();
            int rc = L->batch ? sa_batch_refill(L->batch, f->model, rd.data() + i0, n, &f->params)
                              : sa_batch_create(L->ctx, f->model, rd.data() + i0, n, &f->params, &L->batch);
            const auto t1 = std::chrono::steady_clock::now();
            sa_result *res = nullptr;
            if (rc == SA_OK) rc = sa_batch_run(L->batch, f->precision, &res);
            fill_ms += std::chrono::duration<double, std::milli>(t1 - t0).count();
            run_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count();
            if (rc != SA_OK) return rc;
            std::shared_ptr<sa_result> sp(res, sa_result_free);
            for (size_t i = i0; i < i1; i++) {
                Request *q = wave[i];
                sa_read_result rr;
                sa_result_read(res, (int64_t)(i - i0), &rr);
                q->status = SA_OK;
                q->wres = sp;
                q->widx = (int64_t)(i - i0);
                cells += rr.cells;
            }
            sa_timing tm;
            sa_result_timing(res, &tm);
            dev_ms += tm.total_ms;
            return SA_OK;
        };
        const int rc = run(0, wave.size());
        if (rc != SA_OK) {
            /* one bad read fails its batch: run the reads one by one so that each gets
               its own status */
            if (wave.size() == 1) {
                wave[0]->status = rc;
            } else {
                for (size_t i = 0; i < wave.size(); i++) {
                    const int r1 = run(i, i + 1);
                    if (r1 != SA_OK) wave[i]->status = r1;
                }
            }
        }
        for (Request *q : wave) { /* inputs are no longer needed */
            std::string().swap(q->ref);
            std::vector<double>().swap(q->events);
            std::vector<int64_t>().swap(q->anchors);
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            const double t = now();
            for (Request *q : wave) q->timing.done_s = t;
            stats.waves++;
            stats.reads += (int64_t)wave.size();
            stats.cells += cells;
            stats.device_ms += dev_ms;
            stats.fill_ms += fill_ms;
            stats.run_ms += run_ms;
            if ((int64_t)wave.size() > stats.max_wave_reads) stats.max_wave_reads = (int64_t)wave.size();
            inflight_bytes -= L->wave_bytes;
            L->wave.clear();
            L->wave_bytes = 0;
            L->busy = false;
        }
        cv_done.notify_all();
        cv_sched.notify_all();
    }
}

namespace {

/* HBM a read occupies while it runs: band cells (forward match plane, gap planes on
   refresh diagonals) plus per-diagonal, per-event and output arrays.  Band cells
   are bounded by the rectangles between consecutive anchors widened by the band
   expansion (band_construct keeps each diagonal within the anchors around it):
   O(anchors), cheap enough for the submitting thread. */
int64_t estimate_bytes(const sa_read *r, int kmer_length, const sa_params *p, sa_precision prec,
                       double paths_per_column, bool emission_plane) {
    const int64_t lX = r->ref_length >= kmer_length ? r->ref_length - kmer_length + 1 : 0;
    const int64_t lY = r->n_events;
    const int64_t e = p->diagonal_expansion;
    int64_t cells = 0, px = 0, py = 0;
    for (int64_t i = 0; i <= r->n_anchors; i++) {
        const int64_t x = i < r->n_anchors ? r->anchors[2 * i] + 1 : lX;
        const int64_t y = i < r->n_anchors ? r->anchors[2 * i + 1] + 1 : lY;
        cells += (std::min(lX, std::max<int64_t>(x - px, 0) + e) + 1) * (std::min(lY, std::max<int64_t>(y - py, 0) + e) + 1);
        px = x;
        py = y;
    }
    const int64_t w = prec == SA_FP32 ? 4 : 8;
    /* a degenerate reference keeps three FP64 states of every (cell, path) entry and
       per-entry and per-diagonal path tables */
    /* (an HDP model's register sweeps add one emission per cell) */
    const int64_t cell_bytes = paths_per_column > 0.0 ? (int64_t)(paths_per_column * 48.0) + 1
                             : w * 3 / 2 + (emission_plane ? w : 0);
    const int64_t diag_bytes = paths_per_column > 0.0 ? 64 : 32;
    /* a band too wide for a register window runs on the LDS sweeps, whose rings move to
       per-workgroup slices of HBM when they outgrow LDS: one forward slice for the read
       and one backward slice per traceback window */
    int64_t wmax = 1;
    px = 0;
    py = 0;
    for (int64_t i = 0; i <= r->n_anchors; i++) {
        const int64_t x = i < r->n_anchors ? r->anchors[2 * i] + 1 : lX;
        const int64_t y = i < r->n_anchors ? r->anchors[2 * i + 1] + 1 : lY;
        wmax = std::max(wmax, std::min(std::min(lX, std::max<int64_t>(x - px, 0) + e),
                                       std::min(lY, std::max<int64_t>(y - py, 0) + e)) + 1);
        px = x;
        py = y;
    }
    int64_t ring_bytes = 0;
    const double ppc = paths_per_column > 0.0 ? paths_per_column : 1.0;
    const int64_t ents = (int64_t)((double)wmax * ppc) + 1;
    if (paths_per_column > 0.0 || wmax > 512) {
        int64_t R = 1;
        while (R < ents) R <<= 1;
        const int64_t fwd = paths_per_column > 0.0 ? 72 * R : 9 * R * w + 16 * w;
        const int64_t bwd = paths_per_column > 0.0 ? 88 * R + 16 : 16 * R + 16 + 9 * R * w;
        const int64_t step = std::max<int64_t>(p->min_diags_between_traceback - p->traceback_diagonals - 1, 1);
        const int64_t nwin = (lX + lY) / step + 2;
        if (fwd > 65536) ring_bytes += fwd;
        if (bwd > 65536) ring_bytes += nwin * bwd;
    }
    return cells * cell_bytes + (lX + lY + 1) * diag_bytes + lY * (4 * w + 8 * r->event_stride) + lX * 8 + lY * 36 +
           ring_bytes + 4096;
}

/* the mean paths per column of a reference with 'X' (options per site raised to the
   'X's in each k-mer) */
double mean_paths(const sa_read *r, int k, int64_t n_opt) {
    const int64_t lX = r->ref_length >= k ? r->ref_length - k + 1 : 0;
    if (lX == 0) return 1.0;
    double sum = 0.0;
    int nx = 0;
    for (int64_t i = 0; i < r->ref_length; i++) {
        nx += r->ref[i] == 'X';
        if (i >= k) nx -= r->ref[i - k] == 'X';
        if (i >= k - 1) sum += std::pow((double)n_opt, nx);
    }
    return sum / (double)lX;
}

} // namespace

extern "C" {

int sa_server_create(int device, const sa_server_config *cfg, sa_server **out) {
    *out = nullptr;
    sa_server_config c{};
    if (cfg) c = *cfg;
    if (c.hbm_budget_bytes <= 0) c.hbm_budget_bytes = (int64_t)32 << 30;
    if (c.lanes <= 0) c.lanes = 4;
    if (c.max_wave_reads <= 0) c.max_wave_reads = 2048;
    if (c.length_ratio == 0.0) c.length_ratio = 2.0;
    if (c.batch_delay_ms == 0.0) c.batch_delay_ms = 30.0;
    if (c.length_ratio > 0.0 && c.length_ratio < 1.0) return SA_ERR_INVALID_ARGUMENT;
    if (c.lanes > 4) return SA_ERR_INVALID_ARGUMENT;
    auto *s = new sa_server();
    s->cfg = c;
    s->device = device;
    s->t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < c.lanes; i++) {
        auto L = std::make_unique<Lane>();
        const int rc = sa_context_create_ex(device, 1, &L->ctx);
        if (rc != SA_OK) {
            for (auto &l : s->lanes) sa_context_destroy(l->ctx);
            delete s;
            return rc;
        }
        s->lanes.push_back(std::move(L));
    }
    for (auto &l : s->lanes) {
        Lane *p = l.get();
        l->th = std::thread([s, p] { s->lane_loop(p); });
    }
    s->scheduler = std::thread([s] { s->schedule_loop(); });
    *out = s;
    return SA_OK;
}

namespace {

/* A request's own copy of a read (the sweeps read an event's mean and noise only: those
   two are kept), or nullptr with *rc set for a read the server refuses. */
Request *make_request(const sa_model *model, const sa_read *read, const sa_params *params, sa_precision precision,
                      int *rc) {
    if (read->ref_length < 0 || read->n_events < 0 || read->n_anchors < 0 || read->event_stride < 2) {
        *rc = SA_ERR_INVALID_ARGUMENT;
        return nullptr;
    }
    /* degenerate references (FP64 only) run in waves of their own options */
    const bool degen = read->degenerate_options && read->degenerate_options[0] && read->ref_length > 0 &&
                       memchr(read->ref, 'X', (size_t)read->ref_length);
    if (degen && precision != SA_FP64) {
        *rc = SA_ERR_UNSUPPORTED;
        return nullptr;
    }
    auto *q = new Request();
    q->model = model;
    if (params) q->params = *params; else sa_params_default(&q->params);
    q->precision = precision;
    q->ref.assign(read->ref, (size_t)read->ref_length);
    q->n_events = read->n_events;
    q->stride = 2;
    q->events.resize((size_t)(2 * read->n_events));
    for (int64_t j = 0; j < read->n_events; j++) {
        q->events[2 * j] = read->events[j * read->event_stride];
        q->events[2 * j + 1] = read->events[j * read->event_stride + 1];
    }
    q->anchors.assign(read->anchors, read->anchors + 2 * read->n_anchors);
    q->scale = read->scale;
    q->shift = read->shift;
    q->var = read->var;
    q->rl = read->ragged_left;
    q->rr = read->ragged_right;
    if (degen) q->options = read->degenerate_options;
    const int k = sa_model_kmer_length(model);
    q->est_bytes = estimate_bytes(read, k, &q->params, precision,
                                  degen ? mean_paths(read, k, (int64_t)q->options.size()) : 0.0, model->ek == 2);
    q->n_diag = std::max<int64_t>(read->ref_length - k + 1, 0) + read->n_events;
    *rc = SA_OK;
    return q;
}

} // namespace

int sa_server_submit(sa_server *s, const sa_model *model, const sa_read *read, const sa_params *params,
                     sa_precision precision, int64_t *ticket) {
    if (!s || !model || !read || !ticket || (precision != SA_FP32 && precision != SA_FP64))
        return SA_ERR_INVALID_ARGUMENT;
    int rc;
    Request *q = make_request(model, read, params, precision, &rc);
    if (!q) return rc;
    std::lock_guard<std::mutex> lk(s->mu);
    q->ticket = s->next_ticket++;
    q->timing.submit_s = s->now();
    s->reqs[q->ticket] = q;
    s->pending.push_back(q);
    *ticket = q->ticket;
    s->cv_sched.notify_all();
    return SA_OK;
}

int sa_server_submit_batch(sa_server *s, const sa_model *model, const sa_read *reads, int64_t n,
                           const sa_params *params, sa_precision precision, int64_t *tickets) {
    if (!s || !model || (!reads && n > 0) || (!tickets && n > 0) || n < 0 ||
        (precision != SA_FP32 && precision != SA_FP64))
        return SA_ERR_INVALID_ARGUMENT;
    std::vector<Request *> qs((size_t)n, nullptr);
    std::vector<int> rcs((size_t)n, SA_OK);
    /* the copies on up to 8 threads, whole reads each */
    const int nth = (int)std::max<int64_t>(1, std::min<int64_t>(8, n / 64));
    auto work = [&](int64_t i0, int64_t i1) {
        for (int64_t i = i0; i < i1; i++) qs[i] = make_request(model, &reads[i], params, precision, &rcs[i]);
    };
    {
        std::vector<std::thread> th;
        for (int i = 1; i < nth; i++) th.emplace_back(work, n * i / nth, n * (i + 1) / nth);
        work(0, n / nth);
        for (auto &x : th) x.join();
    }
    for (int64_t i = 0; i < n; i++)
        if (!qs[i]) {
            for (Request *q : qs) delete q;
            return rcs[i];
        }
    std::lock_guard<std::mutex> lk(s->mu);
    const double t = s->now();
    for (int64_t i = 0; i < n; i++) {
        Request *q = qs[i];
        q->ticket = s->next_ticket++;
        q->timing.submit_s = t;
        s->reqs[q->ticket] = q;
        s->pending.push_back(q);
        tickets[i] = q->ticket;
    }
    s->cv_sched.notify_all();
    return SA_OK;
}

int sa_server_wait(sa_server *s, int64_t ticket, double timeout_s, sa_read_result *out, sa_request_timing *timing) {
    if (!s) return SA_ERR_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return SA_ERR_INVALID_ARGUMENT;
    Request *q = it->second;
    auto done = [&] { return q->timing.done_s > 0.0; };
    if (!done()) {
        if (timeout_s <= 0.0) return SA_PENDING;
        if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done)) return SA_PENDING;
    }
    if (timing) *timing = q->timing;
    if (q->status != SA_OK) return q->status;
    if (out) sa_result_read(q->wres.get(), q->widx, out);
    return SA_OK;
}

int sa_server_release(sa_server *s, int64_t ticket) {
    if (!s) return SA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end() || it->second->timing.done_s <= 0.0) return SA_ERR_INVALID_ARGUMENT;
    delete it->second;
    s->reqs.erase(it);
    return SA_OK;
}

int sa_server_get_stats(sa_server *s, sa_server_stats *out) {
    if (!s || !out) return SA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    out->now_s = s->now();
    out->pending = (int64_t)s->pending.size();
    out->inflight_bytes = s->inflight_bytes;
    return SA_OK;
}

void sa_server_destroy(sa_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_sched.notify_all();
    s->cv_lane.notify_all();
    s->cv_done.notify_all();
    if (s->scheduler.joinable()) s->scheduler.join();
    for (auto &l : s->lanes) {
        if (l->th.joinable()) l->th.join();
        if (l->batch) sa_batch_destroy(l->batch);
        sa_context_destroy(l->ctx);
    }
    for (auto &kv : s->reqs) delete kv.second; /* every request, pending or done, until released */
    delete s;
}

} // extern "C"
