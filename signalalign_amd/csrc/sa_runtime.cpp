/*
 * sa_runtime.cpp -- host runtime of libsignalalign_amd: contexts, resident model
 * tables, batches, split points, launch planning and result assembly behind the
 * C-ABI of include/signalalign_amd.h.
 *
 * The host side does only O(reads + anchors) integer work: the split points of
 * getSplitPoints (impl/pairwiseAligner.c:1706-1757), the sub-alignment anchor lists
 * of getPosteriorProbsWithBandingSplittingAlignmentsByLargeGaps (:1773-1837) and
 * the launch plan.  Everything proportional to diagonals or cells runs on the GPU.
 */
#include "sa_internal.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

using sa::TaskDev;
using sa::WinDev;
using sa_rt::AllocStream;
using sa_rt::correct_length;
using sa_rt::device_model;
using sa_rt::fail;
using sa_rt::split_points;
using sa_rt::tables;

struct ReadOut {
    int64_t first_pair = 0, n_pairs = 0;
    double log_likelihood = 0.0;
    int64_t cells = 0, n_windows = 0, n_sub = 0;
};

struct ReadPlan {
    int64_t first_task = 0;
    int64_t n_tasks = 0;
};

struct TaskHost {
    int32_t read;
    int32_t x1, y1;
};

struct sa_batch {
    sa_context *ctx = nullptr;
    const sa_model *model = nullptr;
    sa_params params;
    int64_t n_reads = 0;
    std::vector<ReadPlan> reads;
    std::vector<TaskDev> tasks;
    std::vector<TaskHost> thost;
    int64_t total_pos = 0, total_ev = 0, total_diag = 0, total_wslots = 0, total_anc = 0;
    int64_t ref_bytes = 0, ev_records = 0;
    int64_t stride = 4;
    /* raw inputs resident in HBM */
    DBuf d_ref, d_events, d_ancx, d_ancy, d_tasks, d_x1, d_y1;
    DBuf d_prep_order; /* sub-alignments longest first: the order prep's workgroups start in */
    /* derived */
    DBuf d_kid, d_evm, d_evn, d_evl, d_evi, d_xlo, d_xhi, d_doff, d_foff, d_dflag, d_dcls, d_doffxy;
    DBuf d_task_cells, d_task_xycells, d_task_maxw, d_task_span, d_task_hspan, d_task_nwin, d_win_slots, d_err;
    DBuf d_fM, d_fX, d_fY;
    DBuf d_wins, d_task_order, d_win_order, d_out, d_win_count, d_win_total, d_win_dst, d_packed;
    DBuf d_pq, d_first_win; /* persistent sweeps: work queue, first window per sub-alignment */
    DBuf d_ring_g; /* LDS sweeps' rings in HBM when a band is too wide for LDS */
    /* HDP emission plane, its sub-alignments and chunk numbering; the column pass's
       per-position column ranges and offsets, its (task, column) entries grouped by
       density row and the groups */
    DBuf d_em, d_em_tasks, d_em_chunk0, d_col_a, d_col_b, d_col_off, d_em_ent, d_em_grp;
    DBuf d_em_pt, d_pk, d_pv, d_runs, d_sort_tmp; /* the device grouping: tasks, keys, values, runs, sort scratch */
    /* degenerate references ('X' with options): multi-path sweeps, FP64 */
    int32_t paths = 0;
    std::string options;      /* the bases an 'X' stands for */
    std::string href;         /* the packed reference bytes (host copy, path tables) */
    DBuf d_pcol, d_pent, d_pplane, d_pp, d_psa, d_psb, d_pecol, d_pekid, d_pdoff, d_pdiag, d_pentr, d_pmaxent;
    HBuf h_pentr, h_pdiag; /* the two largest path tables, built in pinned memory */
    /* EM expectations run (sa_expect): FP64, register forward sweeps and LDS backward sweeps with the accumulators */
    int32_t expect = 0;
    double ex_rec_threshold = 1.0;
    DBuf d_ex_trans, d_ex_s0, d_ex_s1, d_ex_obs, d_ex_rec, d_ex_count;
    sa_tuning tune{1, 0, 2.5, 0.5, 0, 0};
    int32_t nruns[2] = {0, 0}; /* runs since the batch or its tuning last changed, per precision */
    /* a recorded run per precision (0 FP64, 1 FP32), replayed while the batch is unchanged */
    struct GraphCache {
        hipGraphExec_t exec = nullptr;
        std::vector<ReadOut> reads;
        std::vector<int32_t> last_win;  /* per sub-alignment: its final window, or -1 */
        std::vector<sa_launch> launches; /* shapes; durations filled from the replay's events */
        int64_t nw = 0, total_pairs = 0, cells = 0;
        int32_t ek = 0;
        int64_t em_c[5] = {0, 0, 0, 0, 0}; /* emission plane chunks per stream lane (prefix) */
        void reset() {
            if (exec) (void)hipGraphExecDestroy(exec);
            exec = nullptr;
            reads.clear();
            last_win.clear();
            launches.clear();
        }
    } gc[2];
    ~sa_batch() {
        for (auto &g : gc) g.reset();
        DBuf *all[] = {&d_ref, &d_events, &d_ancx, &d_ancy, &d_tasks, &d_x1, &d_y1, &d_prep_order, &d_kid, &d_evm,
                       &d_evn,
                       &d_evl, &d_evi, &d_xlo, &d_xhi, &d_doff, &d_foff, &d_dflag, &d_dcls, &d_doffxy, &d_task_cells,
                       &d_task_xycells, &d_task_maxw, &d_task_span, &d_task_hspan,
                       &d_task_nwin, &d_win_slots, &d_err, &d_fM, &d_fX, &d_fY, &d_wins, &d_task_order,
                       &d_win_order, &d_out, &d_win_count, &d_win_total, &d_win_dst, &d_packed, &d_ring_g,
                       &d_pq, &d_first_win,
                       &d_em, &d_em_tasks, &d_em_chunk0, &d_col_a, &d_col_b, &d_col_off, &d_em_ent, &d_em_grp,
                       &d_em_pt, &d_pk, &d_pv, &d_runs, &d_sort_tmp, &d_ex_trans, &d_ex_s0, &d_ex_s1, &d_ex_obs, &d_ex_rec, &d_ex_count,
                       &d_pcol, &d_pent, &d_pplane, &d_pp, &d_psa, &d_psb, &d_pecol, &d_pekid, &d_pdoff,
                       &d_pdiag, &d_pentr, &d_pmaxent};
        for (DBuf *b : all) b->release();
        h_pentr.release();
        h_pdiag.release();
    }
};


/* EM expectations of one read (summed over its sub-alignments) */
struct ReadExpect {
    double trans[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double likelihood = 0.0;
    std::vector<double> s0, s1;
    std::vector<int8_t> obs;
    std::vector<sa_match_record> recs;
};

struct sa_result {
    std::vector<sa_pair> pairs; /* small lists and degenerate batches */
    /* a large list stays in the pinned buffer it was copied into (pool-owned) */
    HBuf pinned;
    std::shared_ptr<sa_rt::PinnedPool> pool;
    std::vector<int32_t> paths; /* per pair, degenerate batches only */
    std::vector<ReadOut> reads;
    sa_timing timing;
    std::vector<sa_launch> launches;
    std::vector<ReadExpect> expect; /* expectations runs only */
    const sa_pair *pair_data() const { return pinned.p ? pinned.as<sa_pair>() : pairs.data(); }
    sa_result() = default;
    sa_result(const sa_result &) = delete; /* the pinned buffer has one owner */
    sa_result &operator=(const sa_result &) = delete;
    ~sa_result() {
        if (pool) pool->give(pinned);
        else pinned.release();
    }
};

/* Lists of at least this many pairs are handed over in pinned memory rather than
   copied into the result (3 MiB; below it the copy costs less than a pinned
   buffer held per live result). */
constexpr int64_t kPinnedPairs = 1 << 18;

/* Queues the device-to-host copy of n packed pairs on st into the buffer the
   result will hold: a pinned buffer from the context's pool for large lists, else
   the context's staging buffer (copied into res->pairs by finish_pairs after the
   stream synchronises). */
static int fetch_pairs(sa_context *ctx, sa_result *res, const void *d_packed, int64_t n, hipStream_t st) {
    const size_t bytes = sizeof(sa_pair) * (size_t)std::max<int64_t>(n, 1);
    HBuf *dst = &ctx->stage;
    if (n >= kPinnedPairs) {
        res->pool = ctx->pair_pool;
        res->pinned = res->pool->take();
        dst = &res->pinned;
    }
    SA_HIP(ctx, dst->ensure(bytes));
    if (n) SA_HIP(ctx, hipMemcpyAsync(dst->p, d_packed, sizeof(sa_pair) * (size_t)n, hipMemcpyDeviceToHost, st));
    return SA_OK;
}

/* The emission plane's span on the device: the longest of the stream lanes' plane
   launches (they overlap one another and the sweeps) */
static double plane_ms(sa_context *ctx, const int64_t *em_c) {
    double best = 0.0;
    for (int lane = 0; lane < 4; lane++) {
        float ms = 0.f;
        if (em_c[lane + 1] > em_c[lane] &&
            hipEventElapsedTime(&ms, ctx->ev[16 + 2 * lane], ctx->ev[17 + 2 * lane]) == hipSuccess)
            best = std::max(best, (double)ms);
    }
    (void)hipGetLastError();
    return best;
}

static void finish_pairs(sa_context *ctx, sa_result *res, int64_t n) {
    if (!res->pinned.p) res->pairs.assign(ctx->stage.as<sa_pair>(), ctx->stage.as<sa_pair>() + n);
}

struct sa_expect_result {
    std::vector<ReadExpect> reads;
    sa_timing timing;
};

namespace {


/* n elements from host memory (pinned: one DMA at the PCIe rate) */
template <typename T>
int upload_raw(sa_context *ctx, DBuf &b, const T *v, int64_t n) {
    SA_HIP(ctx, b.ensure(sizeof(T) * (size_t)std::max<int64_t>(n, 1)));
    if (n > 0) {
        SA_HIP(ctx, hipMemcpyAsync(b.p, v, sizeof(T) * (size_t)n, hipMemcpyHostToDevice, ctx->stream));
        SA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return SA_OK;
}

template <typename T>
int upload(sa_context *ctx, DBuf &b, const std::vector<T> &v) {
    SA_HIP(ctx, b.ensure(sizeof(T) * std::max<size_t>(v.size(), 1)));
    /* stream-ordered: a blocking copy on the null stream would wait for the work of
       every other context on the device */
    if (!v.empty()) {
        SA_HIP(ctx, hipMemcpyAsync(b.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, ctx->stream));
        SA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return SA_OK;
}

template <typename T>
int upload_tables(sa_context *ctx, sa_model *m, DevTables<T> &t, const sa_model_desc *d) {
    const int n = m->n_kmers;
    if (m->kind != SA_EMIT_HDP) {
        std::vector<T> mu(n), isd(n), ca(n), nmu(n), nb(n);
        auto fill = [&](const double *tab) {
            for (int i = 0; i < n; i++) {
                const double lm = tab[5 * i + 0], sd = tab[5 * i + 1], nm = tab[5 * i + 2], lam = tab[5 * i + 4];
                mu[i] = (T)lm;
                isd[i] = sd != 0.0 ? (T)(1.0 / sd) : (T)0;
                ca[i] = sd != 0.0 ? (T)((-0.91893853320467267 - std::log(sd)) +
                                        0.5 * (std::log(lam) - 1.8378770664093453))
                                  : (T)-INFINITY;
                nmu[i] = (T)nm;
                nb[i] = (T)(0.5 * lam / (nm * nm));
            }
        };
        fill(d->match);
        int rc;
        if ((rc = upload(ctx, t.mmu, mu)) || (rc = upload(ctx, t.misd, isd)) || (rc = upload(ctx, t.mca, ca)) ||
            (rc = upload(ctx, t.mnmu, nmu)) || (rc = upload(ctx, t.mnb, nb)))
            return rc;
        fill(d->gapy ? d->gapy : d->match);
        if (!d->gapy) /* level_sd x 1.75 as stateMachine3_loadFromFile does */
            for (int i = 0; i < n; i++) {
                const double sd = d->match[5 * i + 1] * 1.75, lam = d->match[5 * i + 4];
                isd[i] = sd != 0.0 ? (T)(1.0 / sd) : (T)0;
                ca[i] = sd != 0.0 ? (T)((-0.91893853320467267 - std::log(sd)) +
                                        0.5 * (std::log(lam) - 1.8378770664093453))
                                  : (T)-INFINITY;
            }
        if ((rc = upload(ctx, t.ymu, mu)) || (rc = upload(ctx, t.yisd, isd)) || (rc = upload(ctx, t.yca, ca)) ||
            (rc = upload(ctx, t.ynmu, nmu)) || (rc = upload(ctx, t.ynb, nb)))
            return rc;
        std::vector<T> gx(n);
        for (int i = 0; i < n; i++) gx[i] = (T)(d->gapx ? d->gapx[i] : -2.3025850929940455);
        if ((rc = upload(ctx, t.gx, gx))) return rc;
    } else {
        std::vector<T> hmu(n);
        std::vector<int32_t> hrow(n);
        for (int i = 0; i < n; i++) {
            hmu[i] = (T)d->level_mean[i];
            hrow[i] = (int32_t)d->kmer_row[i];
        }
        const int64_t G = d->grid_length;
        std::vector<T> gx(G), ys((size_t)2 * G * d->n_rows);
        for (int64_t i = 0; i < G; i++) gx[i] = (T)d->grid[i];
        for (int64_t r = 0; r < d->n_rows; r++)
            for (int64_t i = 0; i < G; i++) {
                ys[(size_t)(2 * (r * G + i))] = (T)d->density[r * G + i];
                ys[(size_t)(2 * (r * G + i) + 1)] = (T)d->slopes[r * G + i];
            }
        int rc;
        if ((rc = upload(ctx, t.hmu, hmu)) || (rc = upload(ctx, t.hrow, hrow)) || (rc = upload(ctx, t.gridx, gx)) ||
            (rc = upload(ctx, t.ys, ys)))
            return rc;
    }
    return SA_OK;
}

std::mutex g_anchor_mu;
hipEvent_t g_anchor[64] = {};

int next_pow2(int v) {
    int r = 1;
    while (r < v) r <<= 1;
    return r;
}

} // namespace

namespace sa_rt {

hipEvent_t timeline_anchor(int device) { return device >= 0 && device < 64 ? g_anchor[device] : nullptr; }

void timeline_init(int device, hipStream_t st) {
    if (device < 0 || device >= 64) return;
    std::lock_guard<std::mutex> lk(g_anchor_mu);
    if (g_anchor[device]) return;
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return;
    if (hipEventRecord(e, st) != hipSuccess || hipEventSynchronize(e) != hipSuccess) {
        (void)hipEventDestroy(e);
        (void)hipGetLastError();
        return;
    }
    g_anchor[device] = e;
}

double timeline_ms(int device, hipEvent_t ev) {
    const hipEvent_t a = timeline_anchor(device);
    float ms = 0.f;
    if (!a || hipEventElapsedTime(&ms, a, ev) != hipSuccess) {
        (void)hipGetLastError();
        return 0.0;
    }
    return ms;
}

} // namespace sa_rt

extern "C" {

void sa_params_default(sa_params *p) {
    p->threshold = 0.01;
    p->min_diags_between_traceback = 1000;
    p->traceback_diagonals = 40;
    p->diagonal_expansion = 20;
    p->split_matrix_bigger_than_this = (int64_t)3000 * 3000;
}

int sa_device_count(int *count) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) { *count = 0; return SA_ERR_NO_DEVICE; }
    *count = n;
    return SA_OK;
}

int sa_context_create_ex(int device, int n_streams, sa_context **out) {
    *out = nullptr;
    if (n_streams != 1 && n_streams != 4) return SA_ERR_INVALID_ARGUMENT;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) return SA_ERR_NO_DEVICE;
    sa_context *c = new sa_context();
    c->device = device;
    c->n_streams = n_streams;
    /* non-blocking streams at the default priority: no implicit waits on the null stream */
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return SA_ERR_HIP;
    }
    for (auto &e : c->ev) (void)hipEventCreate(&e);
    for (auto &e : c->gev) (void)hipEventCreate(&e);
    for (auto &q : c->aux) {
        if (n_streams == 1) q = c->stream;
        else (void)hipStreamCreateWithFlags(&q, hipStreamNonBlocking);
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->n_cu = prop.multiProcessorCount;
    sa_rt::timeline_init(device, c->stream);
    /* the no-allocation entry point's scratch array on this device (once per device) */
    hipError_t se = hipSuccess;
    c->scratch = sa_rt::scratch_setup(device, &se);
    if (se != hipSuccess) {
        sa_context_destroy(c);
        return SA_ERR_HIP;
    }
    *out = c;
    return SA_OK;
}

int sa_context_create(int device, sa_context **out) { return sa_context_create_ex(device, 4, out); }

void sa_context_destroy(sa_context *ctx) {
    if (!ctx) return;
    if (ctx->batch) sa_batch_destroy(ctx->batch);
    ctx->batch = nullptr;
    sa_rt::enter(ctx);
    for (auto &e : ctx->ev) if (e) (void)hipEventDestroy(e);
    for (auto &e : ctx->gev) if (e) (void)hipEventDestroy(e);
    for (auto &q : ctx->aux) if (q && q != ctx->stream) (void)hipStreamDestroy(q);
    ctx->stage.release();
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *sa_last_error(const sa_context *ctx) { return ctx ? ctx->err : ""; }

int sa_model_create(sa_context *ctx, const sa_model_desc *d, sa_model **out) {
    *out = nullptr;
    if (!ctx || !d || !d->alphabet) return SA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(ctx->mu);
    sa_rt::enter(ctx);
    if (d->kind < 0 || d->kind > 2) return fail(ctx, SA_ERR_UNSUPPORTED, "unknown emission kind");
    std::string alpha(d->alphabet);
    std::sort(alpha.begin(), alpha.end()); /* sequence_prepareAlphabet (impl/pairwiseAligner.c:294-323) */
    if (alpha.empty() || std::adjacent_find(alpha.begin(), alpha.end()) != alpha.end())
        return fail(ctx, SA_ERR_INVALID_ARGUMENT, "alphabet characters must be distinct");
    if (d->kmer_length < 1 || d->kmer_length > sa::kMaxKmer)
        return fail(ctx, SA_ERR_INVALID_ARGUMENT, "kmer_length out of range");
    double nk = std::pow((double)alpha.size(), (double)d->kmer_length);
    if (nk > 2.0e8) return fail(ctx, SA_ERR_UNSUPPORTED, "too many k-mers");
    if (d->kind != SA_EMIT_HDP && !d->match) return fail(ctx, SA_ERR_INVALID_ARGUMENT, "match table missing");
    if (d->kind == SA_EMIT_HDP &&
        (!d->grid || d->grid_length < 2 || !d->density || !d->slopes || !d->kmer_row || !d->level_mean ||
         d->n_rows < 1))
        return fail(ctx, SA_ERR_INVALID_ARGUMENT, "HDP grid incomplete");
    sa_model *m = new sa_model();
    m->ctx = ctx;
    m->device = ctx->device;
    m->kind = d->kind;
    m->k = d->kmer_length;
    m->A = (int32_t)alpha.size();
    m->n_kmers = (int32_t)nk;
    m->alphabet = alpha;
    std::memset(m->lut, -1, sizeof(m->lut));
    for (size_t i = 0; i < alpha.size(); i++) m->lut[(uint8_t)alpha[i]] = (int8_t)i;
    for (int i = 0; i < 9; i++) m->trans[i] = d->transitions[i];
    /* stateMachine3_endStateProb / raggedEndStateProb (impl/stateMachine.c:1033-1061) */
    m->endv[0] = d->transitions[SA_T_MATCH_CONTINUE];
    m->endv[1] = d->transitions[SA_T_MATCH_FROM_GAP_X];
    m->endv[2] = d->transitions[SA_T_MATCH_FROM_GAP_Y];
    m->rendv[0] = (d->transitions[SA_T_GAP_OPEN_X] + d->transitions[SA_T_GAP_OPEN_Y]) / 2.0;
    m->rendv[1] = d->transitions[SA_T_GAP_EXTEND_X];
    m->rendv[2] = d->transitions[SA_T_GAP_EXTEND_Y];
    if (d->kind == SA_EMIT_HDP) {
        m->grid_length = (int32_t)d->grid_length;
        m->inv_dx = 1.0 / (d->grid[1] - d->grid[0]);
        m->grid0 = d->grid[0];
        m->grid1 = d->grid[1];
        m->gridn = d->grid[d->grid_length - 1];
        for (int64_t i = 0; i < m->n_kmers; i++)
            if (d->kmer_row[i] < 0 || d->kmer_row[i] >= d->n_rows) {
                delete m;
                return fail(ctx, SA_ERR_INVALID_ARGUMENT, "kmer_row out of range");
            }
    }
    if (d->kind == SA_EMIT_HDP) {
        m->ek = 2;
    } else {
        /* tables as the .model loader builds them: extra-event = match with sd x 1.75
           (impl/stateMachine.c:1410-1412), one skip emission (:1474-1476), sd > 0 */
        bool std_tables = true;
        for (int64_t i = 0; i < m->n_kmers && std_tables; i++) {
            const double *mt = d->match + 5 * i;
            if (!(mt[1] > 0.0)) std_tables = false;
            if (d->gapy) {
                const double *gt = d->gapy + 5 * i;
                if (gt[0] != mt[0] || gt[1] != mt[1] * 1.75 || gt[2] != mt[2] || gt[4] != mt[4]) std_tables = false;
            }
            if (d->gapx && d->gapx[i] != d->gapx[0]) std_tables = false;
        }
        m->gx_const = d->gapx ? d->gapx[0] : -2.3025850929940455;
        m->ek = std_tables ? 0 : 1;
    }
    if (d->kind == 2) {
        m->n_rows = (int32_t)d->n_rows;
    }
    std::vector<int8_t> lut(m->lut, m->lut + 256);
    int rc = upload(ctx, m->d_lut, lut);
    if (!rc) rc = upload_tables<double>(ctx, m, m->t64, d);
    if (!rc) rc = upload_tables<float>(ctx, m, m->t32, d);
    if (rc) {
        m->t64.release(); m->t32.release(); m->d_lut.release();
        delete m;
        return rc;
    }
    *out = m;
    return SA_OK;
}

int sa_model_kmer_length(const sa_model *m) { return m ? m->k : 0; }

void sa_model_destroy(sa_model *m) {
    if (!m) return;
    (void)hipSetDevice(m->device); /* (its context may be gone already) */
    m->t64.release();
    m->t32.release();
    m->d_lut.release();
    delete m;
}

int sa_batch_create(sa_context *ctx, const sa_model *model, const sa_read *reads, int64_t n_reads,
                    const sa_params *params, sa_batch **out);
int sa_batch_run(sa_batch *b, sa_precision precision, sa_result **out);
void sa_batch_destroy(sa_batch *b);

} // extern "C"

namespace {

/* Splits the reads into sub-alignments and uploads the raw inputs into the batch's
   device buffers (which only grow, so a refilled batch reuses them). */
int batch_fill(sa_context *ctx, sa_batch *b, const sa_model *model, const sa_read *reads, int64_t n_reads,
               const sa_params *params) {
    sa_params p;
    if (params) p = *params; else sa_params_default(&p);
    /* getPosteriorProbsWithBanding prerequisites (impl/pairwiseAligner.c:1291-1295) */
    if (!(p.traceback_diagonals >= 1) || !(p.diagonal_expansion >= 0) || p.diagonal_expansion % 2 != 0 ||
        !(p.min_diags_between_traceback >= 2) || !(p.traceback_diagonals + 1 < p.min_diags_between_traceback) ||
        !(p.threshold >= 0.0 && p.threshold <= 1.0))
        return fail(ctx, SA_ERR_INVALID_ARGUMENT, "invalid banding parameters");
    for (auto &g : b->gc) g.reset();
    b->nruns[0] = b->nruns[1] = 0;
    b->ctx = ctx;
    b->model = model;
    b->params = p;
    b->n_reads = n_reads;
    b->reads.assign(n_reads, ReadPlan());
    b->tasks.clear();
    b->thost.clear();
    b->total_pos = b->total_ev = b->total_diag = b->total_wslots = b->total_anc = 0;
    b->ref_bytes = b->ev_records = 0;
    const int k = model->k;
    std::vector<int64_t> splits;
    std::vector<int32_t> ancx, ancy;
    int64_t ref_off = 0, ev_off = 0;
    int64_t stride = -1;
    b->paths = 0;
    b->options.clear();
    b->href.clear();
    for (int64_t r = 0; r < n_reads; r++) {
        const sa_read &rd = reads[r];
        if (rd.ref_length > 0 && rd.ref && memchr(rd.ref, 'X', (size_t)rd.ref_length)) {
            if (!rd.degenerate_options || !rd.degenerate_options[0])
                return fail(ctx, SA_ERR_BAD_KMER, "read " + std::to_string(r) + ": 'X' with no degenerate options");
            if (b->paths && b->options != rd.degenerate_options)
                return fail(ctx, SA_ERR_UNSUPPORTED, "reads with 'X' in one batch take the same degenerate options");
            b->options = rd.degenerate_options;
            b->paths = 1;
        }
    }
    for (int64_t r = 0; r < n_reads; r++) {
        const sa_read &rd = reads[r];
        if (rd.event_stride < 2 || (stride >= 0 && rd.event_stride != stride) || rd.n_events < 0 ||
            rd.ref_length < 0 || rd.n_anchors < 0 || !(rd.var != 0.0)) {
            return fail(ctx, SA_ERR_INVALID_ARGUMENT, "read " + std::to_string(r) + ": bad sizes");
        }
        stride = rd.event_stride;
        const int64_t lX = correct_length(rd.ref_length, k), lY = rd.n_events;
        if (!split_points(rd.anchors, rd.n_anchors, lX, lY, p.split_matrix_bigger_than_this, rd.ragged_left,
                          rd.ragged_right, splits)) {
            return fail(ctx, SA_ERR_ANCHORS, "read " + std::to_string(r) + ": anchors out of order or range");
        }
        const int64_t nsplit = (int64_t)splits.size() / 4;
        b->reads[r].first_task = (int64_t)b->tasks.size();
        b->reads[r].n_tasks = nsplit;
        int64_t j = 0;
        for (int64_t i = 0; i < nsplit; i++) {
            const int64_t x1 = splits[4 * i], y1 = splits[4 * i + 1], x2 = splits[4 * i + 2], y2 = splits[4 * i + 3];
            TaskDev t;
            std::memset(&t, 0, sizeof(t));
            t.lX = (int32_t)(x2 - x1);
            t.lY = (int32_t)(y2 - y1);
            if (x2 - x1 > INT32_MAX / 4 || y2 - y1 > INT32_MAX / 4) {
                return fail(ctx, SA_ERR_UNSUPPORTED, "sub-alignment too long");
            }
            t.ref_off = ref_off + x1;
            t.ev_off = ev_off + y1;
            t.pos_off = b->total_pos;
            t.evc_off = b->total_ev;
            t.diag_off = b->total_diag;
            t.anc_off = (int64_t)ancx.size();
            int32_t na = 0;
            int64_t px = -1, py = -1;
            while (j < rd.n_anchors) {
                const int64_t x = rd.anchors[2 * j], y = rd.anchors[2 * j + 1];
                if (x + y >= x2 + y2) break;
                if (!(x >= x1 && x < x2 && y >= y1 && y < y2) || x - x1 <= px || y - y1 <= py) {
                    return fail(ctx, SA_ERR_ANCHORS, "read " + std::to_string(r) + ": anchors out of order");
                }
                px = x - x1;
                py = y - y1;
                ancx.push_back((int32_t)(x - x1 + 1));
                ancy.push_back((int32_t)(y - y1 + 1));
                na++;
                j++;
            }
            t.n_anc = na;
            t.ragged_left = rd.ragged_left || i > 0;
            t.ragged_right = rd.ragged_right || i < nsplit - 1;
            t.read = (int32_t)r;
            const int64_t N = (int64_t)t.lX + t.lY;
            t.win_off = (int32_t)b->total_wslots;
            t.win_cap = sa::window_cap(N, p.min_diags_between_traceback, p.traceback_diagonals);
            t.scale = rd.scale;
            t.shift = rd.shift;
            t.var = rd.var;
            b->total_pos += t.lX;
            b->total_ev += t.lY;
            /* per-diagonal arrays of a sub-alignment start on 128-byte boundaries (32
               slots): the persistent sweeps read one sub-alignment's offsets through the
               scalar cache while another's are being written */
            b->total_diag += (N + 1 + 31) / 32 * 32;
            b->total_wslots += t.win_cap;
            if (N > 0) {
                b->tasks.push_back(t);
                b->thost.push_back({(int32_t)r, (int32_t)x1, (int32_t)y1});
            } else {
                b->reads[r].n_tasks--;
            }
        }
        if (j != rd.n_anchors) {
            return fail(ctx, SA_ERR_ANCHORS, "read " + std::to_string(r) + ": anchors not consumed");
        }
        ref_off += rd.ref_length;
        ev_off += rd.n_events;
    }
    b->ref_bytes = ref_off;
    b->ev_records = ev_off;
    b->total_anc = (int64_t)ancx.size();
    if (stride < 0) stride = 4;
    b->stride = stride;
    /* pack the raw inputs straight into pinned staging, then one DMA per array */
    const size_t ref_bytes = (size_t)std::max<int64_t>(ref_off, 1);
    const size_t ref_pad = (ref_bytes + 255) & ~(size_t)255;
    const size_t ev_bytes = sizeof(double) * (size_t)std::max<int64_t>(ev_off * stride, 1);
    SA_HIP(ctx, ctx->stage.ensure(ref_pad + ev_bytes));
    char *h_ref = ctx->stage.as<char>();
    double *h_ev = reinterpret_cast<double *>(h_ref + ref_pad);
    {
        /* a large batch (a serving wave, a bench batch) is packed by several threads,
           whole reads each: one thread copies a few GB/s into pinned memory */
        std::vector<int64_t> ro(n_reads + 1, 0), eo(n_reads + 1, 0);
        for (int64_t r = 0; r < n_reads; r++) {
            ro[r + 1] = ro[r] + reads[r].ref_length;
            eo[r + 1] = eo[r] + reads[r].n_events;
        }
        auto pack = [&](int64_t r0, int64_t r1) {
            for (int64_t r = r0; r < r1; r++) {
                if (reads[r].ref_length) std::memcpy(h_ref + ro[r], reads[r].ref, (size_t)reads[r].ref_length);
                if (reads[r].n_events)
                    std::memcpy(h_ev + eo[r] * stride, reads[r].events,
                                sizeof(double) * (size_t)(reads[r].n_events * stride));
            }
        };
        const int64_t bytes = ro[n_reads] + (int64_t)sizeof(double) * eo[n_reads] * stride;
        const int nth = (int)std::min<int64_t>(std::min<int64_t>(8, n_reads), bytes >> 23); /* 8 MB a thread */
        if (nth <= 1) {
            pack(0, n_reads);
        } else {
            std::vector<std::thread> th;
            for (int i = 1; i < nth; i++) th.emplace_back(pack, n_reads * i / nth, n_reads * (i + 1) / nth);
            pack(0, n_reads / nth);
            for (auto &x : th) x.join();
        }
    }
    if (b->paths) b->href.assign(h_ref, (size_t)ref_off);
    SA_HIP(ctx, b->d_ref.ensure(ref_bytes));
    SA_HIP(ctx, b->d_events.ensure(ev_bytes));
    SA_HIP(ctx, hipMemcpyAsync(b->d_ref.p, h_ref, ref_bytes, hipMemcpyHostToDevice, ctx->stream));
    SA_HIP(ctx, hipMemcpyAsync(b->d_events.p, h_ev, ev_bytes, hipMemcpyHostToDevice, ctx->stream));
    SA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<int32_t> x1(b->tasks.size()), y1(b->tasks.size());
    for (size_t i = 0; i < b->tasks.size(); i++) { x1[i] = b->thost[i].x1; y1[i] = b->thost[i].y1; }
    /* prep's workgroups take the sub-alignments longest first: a long one's
       workgroup then starts in the first round instead of holding the kernel's tail */
    std::vector<int32_t> porder(b->tasks.size());
    for (size_t i = 0; i < porder.size(); i++) porder[i] = (int32_t)i;
    std::stable_sort(porder.begin(), porder.end(), [&](int32_t u, int32_t v) {
        return (int64_t)b->tasks[u].lX + b->tasks[u].lY > (int64_t)b->tasks[v].lX + b->tasks[v].lY;
    });
    int rc;
    if ((rc = upload(ctx, b->d_ancx, ancx)) || (rc = upload(ctx, b->d_ancy, ancy)) ||
        (rc = upload(ctx, b->d_tasks, b->tasks)) || (rc = upload(ctx, b->d_x1, x1)) ||
        (rc = upload(ctx, b->d_y1, y1)) || (rc = upload(ctx, b->d_prep_order, porder)))
        return rc;
    return SA_OK;

}

void batch_clear(sa_batch *b) {
    b->n_reads = 0;
    b->reads.clear();
    b->tasks.clear();
    b->thost.clear();
    b->total_pos = b->total_ev = b->total_diag = b->total_wslots = b->total_anc = 0;
}

} // namespace

extern "C" {

int sa_batch_create(sa_context *ctx, const sa_model *model, const sa_read *reads, int64_t n_reads,
                    const sa_params *params, sa_batch **out) {
    *out = nullptr;
    if (!ctx || !model || (!reads && n_reads > 0) || n_reads < 0) return SA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(ctx->mu);
    sa_rt::enter(ctx);
    AllocStream as(ctx->stream);
    sa_batch *b = new sa_batch();
    const int rc = batch_fill(ctx, b, model, reads, n_reads, params);
    if (rc != SA_OK) {
        delete b;
        return rc;
    }
    *out = b;
    return SA_OK;
}

int sa_batch_refill(sa_batch *b, const sa_model *model, const sa_read *reads, int64_t n_reads,
                    const sa_params *params) {
    if (!b || !model || (!reads && n_reads > 0) || n_reads < 0) return SA_ERR_INVALID_ARGUMENT;
    sa_context *ctx = b->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    sa_rt::enter(ctx);
    AllocStream as(ctx->stream);
    const int rc = batch_fill(ctx, b, model, reads, n_reads, params);
    if (rc != SA_OK) batch_clear(b);
    return rc;
}

void sa_tuning_default(sa_tuning *t) {
    t->use_graph = 1;
    t->waves_per_simd = 0;
    t->long_factor = 2.5;
    t->stream_split = 0.5;
    t->persist_waves = 0;
    t->reserved = 0;
}

int sa_batch_set_tuning(sa_batch *b, const sa_tuning *t) {
    if (!b || !t || t->waves_per_simd < 0 || t->waves_per_simd > 8 || !(t->stream_split >= 0.0) ||
        !(t->stream_split < 1.0) || t->persist_waves < -1 || t->persist_waves > 8)
        return SA_ERR_INVALID_ARGUMENT;
    /* the persistent launch hung on the device with its cause not found: refused */
    if (t->persist_waves > 0) return SA_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lk(b->ctx->mu);
    (void)hipSetDevice(b->ctx->device);
    for (auto &g : b->gc) g.reset();
    b->nruns[0] = b->nruns[1] = 0;
    b->tune = *t;
    return SA_OK;
}

int sa_batch_get_tuning(const sa_batch *b, sa_tuning *t) {
    if (!b || !t) return SA_ERR_INVALID_ARGUMENT;
    *t = b->tune;
    return SA_OK;
}

int sa_batch_autotune(sa_batch *b, sa_precision precision, sa_tuning *best, double *best_ms) {
    if (!b) return SA_ERR_INVALID_ARGUMENT;
    sa_tuning cur = b->tune;
    const int32_t keep_graph = cur.use_graph;
    cur.use_graph = 0;
    /* the sweeps' device time (what the knobs change), best of two runs */
    auto time_it = [&](const sa_tuning &t, double *ms) -> int {
        int rc = sa_batch_set_tuning(b, &t);
        if (rc) return rc;
        *ms = 1e30;
        for (int k = 0; k < 2; k++) {
            sa_result *r = nullptr;
            if ((rc = sa_batch_run(b, precision, &r))) return rc;
            *ms = std::min(*ms, (double)r->timing.fb_ms);
            sa_result_free(r);
        }
        return SA_OK;
    };
    double cur_ms = 0.0;
    int rc = time_it(cur, &cur_ms);
    if (rc) return rc;
    const double lf[] = {2.5, 2.0, 3.0, 1.5, 4.0, -1.0};
    const double sp[] = {0.5, 0.55, 0.6, 0.65, 0.4, 0.7, 0.3};
    const int wp[] = {0, 2, 3, 4};
    for (int knob = 0; knob < 3; knob++) {
        const int n = knob == 0 ? 6 : knob == 1 ? 7 : 4;
        for (int i = 0; i < n; i++) {
            sa_tuning t = cur;
            if (knob == 0) t.long_factor = lf[i];
            else if (knob == 1) t.stream_split = sp[i];
            else t.waves_per_simd = wp[i];
            if (t.long_factor == cur.long_factor && t.stream_split == cur.stream_split &&
                t.waves_per_simd == cur.waves_per_simd && t.persist_waves == cur.persist_waves)
                continue;
            double ms = 0.0;
            if ((rc = time_it(t, &ms))) return rc;
            if (ms < 0.99 * cur_ms) { cur = t; cur_ms = ms; } /* a clear gain, not noise */
        }
    }
    cur.use_graph = keep_graph;
    if ((rc = sa_batch_set_tuning(b, &cur))) return rc;
    if (best) *best = cur;
    if (best_ms) *best_ms = cur_ms;
    return SA_OK;
}

void sa_batch_destroy(sa_batch *b) {
    if (!b) return;
    sa_context *ctx = b->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    sa_rt::enter(ctx);
    AllocStream as(ctx->stream); /* buffers go back to the pool on the context's stream */
    delete b;
}

} // extern "C"

namespace {

constexpr int kRetryLds = 1000;

template <typename T>
int run_batch(sa_batch *b, sa_result *res, bool force_lds) {
    sa_context *ctx = b->ctx;
    const sa_model *model = b->model;
    const bool expect = b->expect != 0;
    if (expect && sizeof(T) != 8) return fail(ctx, SA_ERR_UNSUPPORTED, "expectations run in FP64");
    /* an expectations run takes the register forward sweeps (every diagonal's three
       states stored) and the LDS backward sweeps, which carry the accumulators */
    const bool lds_backward = force_lds || expect;
    hipStream_t st = ctx->stream;
    const int64_t nt = (int64_t)b->tasks.size();
    sa_timing &tm = res->timing;
    std::memset(&tm, 0, sizeof(tm));
    res->reads.assign(b->n_reads, ReadOut());
    if (nt == 0) {
        for (int64_t r = 0; r < b->n_reads; r++) res->reads[r].n_sub = 0;
        return SA_OK;
    }
    /* derived buffers */
    /* +64 slots: a block with no k-mers (or no events) still reads slot 0 of its range */
    SA_HIP(ctx, b->d_kid.ensure(sizeof(int32_t) * (b->total_pos + 64)));
    SA_HIP(ctx, b->d_evm.ensure(sizeof(T) * (b->total_ev + 64)));
    SA_HIP(ctx, b->d_evn.ensure(sizeof(T) * (b->total_ev + 64)));
    SA_HIP(ctx, b->d_evl.ensure(sizeof(T) * (b->total_ev + 64)));
    SA_HIP(ctx, b->d_evi.ensure(sizeof(T) * (b->total_ev + 64)));
    SA_HIP(ctx, hipMemsetAsync(b->d_kid.p, 0, sizeof(int32_t) * (b->total_pos + 64), st));
    SA_HIP(ctx, hipMemsetAsync(b->d_evm.p, 0, sizeof(T) * (b->total_ev + 64), st));
    SA_HIP(ctx, hipMemsetAsync(b->d_evn.p, 0, sizeof(T) * (b->total_ev + 64), st));
    SA_HIP(ctx, hipMemsetAsync(b->d_evl.p, 0, sizeof(T) * (b->total_ev + 64), st));
    SA_HIP(ctx, hipMemsetAsync(b->d_evi.p, 0, sizeof(T) * (b->total_ev + 64), st));
    SA_HIP(ctx, b->d_xlo.ensure(sizeof(int32_t) * b->total_diag));
    SA_HIP(ctx, b->d_xhi.ensure(sizeof(int32_t) * b->total_diag));
    SA_HIP(ctx, b->d_doff.ensure(sizeof(int32_t) * b->total_diag));
    SA_HIP(ctx, b->d_foff.ensure(sizeof(double) * b->total_diag));
    SA_HIP(ctx, b->d_dflag.ensure(sizeof(int8_t) * b->total_diag + 4)); /* + padding for dword reads */
    SA_HIP(ctx, b->d_dcls.ensure(sizeof(int8_t) * b->total_diag + 4));
    SA_HIP(ctx, b->d_doffxy.ensure(sizeof(int32_t) * b->total_diag));
    SA_HIP(ctx, b->d_task_cells.ensure(sizeof(int64_t) * nt));
    SA_HIP(ctx, b->d_task_xycells.ensure(sizeof(int64_t) * nt));
    SA_HIP(ctx, b->d_task_maxw.ensure(sizeof(int32_t) * nt));
    SA_HIP(ctx, b->d_task_span.ensure(sizeof(int32_t) * nt));
    SA_HIP(ctx, b->d_task_hspan.ensure(sizeof(int32_t) * nt));
    SA_HIP(ctx, b->d_task_nwin.ensure(sizeof(int32_t) * nt));
    SA_HIP(ctx, b->d_win_slots.ensure(sizeof(WinDev) * std::max<int64_t>(b->total_wslots, 1)));
    SA_HIP(ctx, b->d_err.ensure(sizeof(int32_t)));
    SA_HIP(ctx, hipMemsetAsync(b->d_err.p, 0, sizeof(int32_t), st));

    const DevTables<T> &tab = tables<T>(model);
    sa::Args<T> a;
    std::memset(&a, 0, sizeof(a));
    a.m = device_model<T>(model, tab);
    a.tasks = b->d_tasks.as<TaskDev>();
    a.task_order = b->d_prep_order.as<int32_t>(); /* (the sweeps' launches set their own) */
    a.ref = b->d_ref.as<char>();
    a.events = b->d_events.as<double>();
    a.event_stride = b->stride;
    a.anc_x = b->d_ancx.as<int32_t>();
    a.anc_y = b->d_ancy.as<int32_t>();
    a.lut = model->d_lut.as<int8_t>();
    a.kid = b->d_kid.as<int32_t>();
    a.evm = b->d_evm.as<T>(); a.evn = b->d_evn.as<T>(); a.evl = b->d_evl.as<T>(); a.evi = b->d_evi.as<T>();
    a.xlo = b->d_xlo.as<int32_t>(); a.xhi = b->d_xhi.as<int32_t>(); a.doff = b->d_doff.as<int32_t>();
    a.foff = b->d_foff.as<double>();
    a.dflag = b->d_dflag.as<int8_t>();
    a.dcls = b->d_dcls.as<int8_t>();
    a.doffxy = b->d_doffxy.as<int32_t>();
    a.task_cells = b->d_task_cells.as<int64_t>();
    a.task_xycells = b->d_task_xycells.as<int64_t>();
    a.task_maxw = b->d_task_maxw.as<int32_t>();
    a.task_span = b->d_task_span.as<int32_t>();
    a.task_hspan = b->d_task_hspan.as<int32_t>();
    a.task_nwin = b->d_task_nwin.as<int32_t>();
    a.win_slots = b->d_win_slots.as<WinDev>();
    a.err = b->d_err.as<int32_t>();
    a.threshold = b->params.threshold;
    a.min_diags = b->params.min_diags_between_traceback;
    a.traceback = b->params.traceback_diagonals;
    a.expansion = b->params.diagonal_expansion;
    a.expect = expect ? 1 : 0;
    a.ex_rec_threshold = b->ex_rec_threshold;
    a.paths = b->paths;
    /* HDP: the register sweeps read each cell's density emission from a plane filled
       before them (SA_HDP_PLANE=0: they gather the densities themselves) */
    {
        const char *hp = getenv("SA_HDP_PLANE");
        a.em_plane = model->ek == 2 && !force_lds && !(hp && std::strcmp(hp, "0") == 0) ? 1 : 0;
    }
    if (a.em_plane) {
        /* per position slot pos_off + task index + x, x = 0..lX */
        const size_t ncs = sizeof(int32_t) * (size_t)(b->total_pos + nt + 64);
        SA_HIP(ctx, b->d_col_a.ensure(ncs));
        SA_HIP(ctx, b->d_col_b.ensure(ncs));
        SA_HIP(ctx, b->d_col_off.ensure(ncs));
        a.col_a = b->d_col_a.as<int32_t>();
        a.col_b = b->d_col_b.as<int32_t>();
        a.col_off = b->d_col_off.as<int32_t>();
    }

    SA_HIP(ctx, hipEventRecord(ctx->ev[0], st));
    SA_HIP(ctx, sa::launch_prep<T>(a, (int)nt, st));
    SA_HIP(ctx, hipEventRecord(ctx->ev[1], st));
    std::vector<int64_t> cells(nt), xycells(nt);
    std::vector<int32_t> maxw(nt), nwin(nt), span(nt), hspan(nt);
    std::vector<WinDev> slots(b->total_wslots);
    int32_t err = 0;
    SA_HIP(ctx, hipMemcpyAsync(cells.data(), b->d_task_cells.p, sizeof(int64_t) * nt, hipMemcpyDeviceToHost, st));
    SA_HIP(ctx, hipMemcpyAsync(xycells.data(), b->d_task_xycells.p, sizeof(int64_t) * nt, hipMemcpyDeviceToHost, st));
    SA_HIP(ctx, hipMemcpyAsync(maxw.data(), b->d_task_maxw.p, sizeof(int32_t) * nt, hipMemcpyDeviceToHost, st));
    SA_HIP(ctx, hipMemcpyAsync(span.data(), b->d_task_span.p, sizeof(int32_t) * nt, hipMemcpyDeviceToHost, st));
    SA_HIP(ctx, hipMemcpyAsync(hspan.data(), b->d_task_hspan.p, sizeof(int32_t) * nt, hipMemcpyDeviceToHost, st));
    SA_HIP(ctx, hipMemcpyAsync(nwin.data(), b->d_task_nwin.p, sizeof(int32_t) * nt, hipMemcpyDeviceToHost, st));
    SA_HIP(ctx, hipMemcpyAsync(slots.data(), b->d_win_slots.p, sizeof(WinDev) * b->total_wslots,
                               hipMemcpyDeviceToHost, st));
    SA_HIP(ctx, hipMemcpyAsync(&err, b->d_err.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    SA_HIP(ctx, hipStreamSynchronize(st));
    if (err & sa::ERR_BAD_KMER) return fail(ctx, SA_ERR_BAD_KMER, "reference k-mer with a character outside the alphabet");
    if (err & sa::ERR_BAND) return fail(ctx, SA_ERR_ANCHORS, "invalid band (anchors outside the matrix or out of order)");
    if (err & sa::ERR_WINDOWS) return fail(ctx, SA_ERR_HIP, "window schedule overflow");

    /* plan: forward planes, windows, outputs */
    int64_t total_cells = 0, total_xy = 0;
    std::vector<TaskDev> tasks = b->tasks;
    for (int64_t i = 0; i < nt; i++) {
        tasks[i].cell_off = total_cells;
        tasks[i].xy_off = total_xy;
        total_cells += cells[i];
        total_xy += xycells[i];
    }
    std::vector<WinDev> wins;
    std::vector<int32_t> task_first_win(nt);
    int64_t out_total = 0, rec_total = 0;
    for (int64_t i = 0; i < nt; i++) {
        task_first_win[i] = (int32_t)wins.size();
        for (int32_t w = 0; w < nwin[i]; w++) {
            WinDev W = slots[tasks[i].win_off + w];
            W.task = (int32_t)i;
            W.out_off = out_total;
            W.out_cap = 2 * (W.TF - W.T) + 64;
            out_total += W.out_cap;
            /* every posterior cell may hold a record at threshold 0; above it a few
               per diagonal (a window that overflows runs again with its count) */
            W.rec_off = rec_total;
            W.rec_cap = !expect ? 0
                      : b->ex_rec_threshold > 0.0 ? (int32_t)std::min<int64_t>(W.cells, 2 * (W.TF - W.T) + 64)
                                                  : (int32_t)W.cells;
            rec_total += W.rec_cap;
            wins.push_back(W);
        }
    }
    const int64_t nw = (int64_t)wins.size();

    /* Kernels per sub-alignment.  Forward: the register sweep with the smallest lane
       window covering its band span, one wavefront per sub-alignment; long
       sub-alignments (and every sub-alignment of a small call) get four wavefronts
       per sub-alignment to shorten each anti-diagonal step.  Backward: one wavefront
       per traceback window.  Bands no register window covers use the LDS sweeps. */
    static const int kCpt[] = {1, 2, 3, 4, 6, 8};
    auto cpt1 = [&](int sp) {
        for (int k = 0; k < 6; k++)
            if (64 * kCpt[k] >= sp + kCpt[k] - 1) return kCpt[k];
        return 0;
    };
    /* multi-wavefront forward shapes, narrowest first: (waves, cpt) */
    static const int kMW[4][2] = {{2, 1}, {3, 1}, {4, 1}, {4, 2}};
    auto mw_shape = [&](int sp) {
        for (int k = 0; k < 4; k++)
            if (64 * kMW[k][0] * kMW[k][1] >= sp + kMW[k][1] - 1) return k;
        return -1;
    };
    std::vector<int64_t> nd(nt);
    for (int64_t i = 0; i < nt; i++) nd[i] = (int64_t)b->tasks[i].lX + b->tasks[i].lY;
    int64_t long_n = INT64_MAX;
    {
        std::vector<int64_t> srt(nd);
        std::sort(srt.begin(), srt.end());
        const int64_t med = srt[srt.size() / 2];
        const double lf = b->tune.long_factor == 0.0 ? 2.5 : b->tune.long_factor;
        long_n = nt <= 2 * (int64_t)std::max(ctx->n_cu, 1) ? 0
               : lf < 0.0 ? INT64_MAX
                          : std::max<int64_t>(20000, (int64_t)(lf * (double)med));
        const char *fw = getenv("SA_FWD_WAVES"); /* "1": one wavefront per forward sweep; "4": several, always */
        if (fw && std::strcmp(fw, "1") == 0) long_n = INT64_MAX;
        if (fw && std::strcmp(fw, "4") == 0) long_n = 0;
    }
    /* forward group key: 0 = LDS; 1 + k = one wavefront, kCpt[k]; 7 + k = kMW[k] (the
       narrowest several-wavefront window that holds the band); 11 + k = the halo sweep
       with k + 2 wavefronts (preferred where its window holds the band at every
       synchronisation) */
    constexpr int kFG = 15, kBG = 7, kPG = 14; /* kPG: the persistent launch */
    const char *halo_env = getenv("SA_HALO"); /* "0": the several-wavefront sweeps synchronise every diagonal */
    const bool use_halo = !(halo_env && std::strcmp(halo_env, "0") == 0);
    auto halo_shape = [&](int hsp) {
        for (int w = 2; w <= 4; w++)
            if (sa::halo_window(w) >= hsp) return w;
        return 0;
    };
    std::vector<int> fgrp(nt), bgrp(nt);
    for (int64_t i = 0; i < nt; i++) {
        int f = 0, bb = 0;
        /* the register sweeps address a sub-alignment's planes with 32-bit byte offsets:
           one whose planes reach 4 GB runs on the LDS sweeps */
        if (!force_lds && (uint64_t)cells[i] * sizeof(T) < ((uint64_t)1 << 32)) {
            const int c1 = cpt1(span[i]), mw = mw_shape(span[i]);
            if (c1) {
                bb = 1 + (int)(std::find(kCpt, kCpt + 6, c1) - kCpt);
                f = bb;
            }
            if (mw >= 0 && nd[i] >= long_n) f = 7 + mw;
            const int hw = use_halo ? halo_shape(hspan[i]) : 0;
            if (hw && nd[i] >= long_n) f = 11 + hw - 2;
        }
        fgrp[i] = f;
        bgrp[i] = lds_backward ? 0 : bb;
    }
    /* backward group per traceback window: the narrowest register window that holds
       the band over the window's own diagonals (LDS sub-alignments stay on LDS) */
    std::vector<int> wgrp(nw);
    for (int64_t i = 0; i < nt; i++)
        for (int32_t w = 0; w < nwin[i]; w++) {
            const int32_t wi = task_first_win[i] + w;
            const int c = bgrp[i] == 0 ? 0 : cpt1(wins[wi].span);
            wgrp[wi] = c == 0 ? bgrp[i] : std::min(bgrp[i], 1 + (int)(std::find(kCpt, kCpt + 6, c) - kCpt));
        }
    /* Persistent sweeps: the one-wavefront sub-alignments whose register windows hold at
       most three positions per lane in both directions run as one persistent launch
       (their forward sweeps, then the windows of those done) on stream 3, enqueued
       after the other streams' forward launches; their windows leave the backward
       launches. */
    /* Off: sa_batch_set_tuning refuses persist_waves 1-8 (the persistent launch hung on
       the GPU with its cause not found). */
    const bool persist = !expect && b->tune.persist_waves > 0;
    int64_t n_pers = 0, n_pers_win = 0;
    if (persist)
        for (int64_t i = 0; i < nt; i++)
            if (fgrp[i] >= 1 && fgrp[i] <= 3 && bgrp[i] >= 1 && bgrp[i] <= 3) {
                fgrp[i] = kPG;
                n_pers++;
                n_pers_win += nwin[i];
                for (int32_t w = 0; w < nwin[i]; w++) wgrp[task_first_win[i] + w] = kBG;
            }
    /* Streams: sub-alignments whose forward sweep uses several wavefronts (the long
       ones) run on stream 1, LDS sweeps on stream 2, one-wavefront forward sweeps
       on streams 0 and 3 (the sub-alignments holding the first half of their cells,
       longest first, on 0).  Each stream runs its forward launches and then the
       backward launches of its own sub-alignments' windows, so backward passes of
       short sub-alignments overlap the forward passes of long ones. */
    constexpr int kLanes = 4;
    std::vector<int> tlane(nt);
    {
        std::vector<int32_t> nw1;
        int64_t nw1_cells = 0;
        for (int64_t i = 0; i < nt; i++) {
            tlane[i] = fgrp[i] == 0 ? 2 : fgrp[i] == kPG ? 3 : fgrp[i] >= 7 ? 1 : 0;
            if (tlane[i] == 0) { nw1.push_back((int32_t)i); nw1_cells += cells[i]; }
        }
        std::stable_sort(nw1.begin(), nw1.end(), [&](int32_t x, int32_t y) { return cells[x] > cells[y]; });
        /* one stream (a serving lane): the lanes run one after another, so the
           one-wavefront sweeps stay in one launch and the several-wavefront ones too */
        const bool streams4 = ctx->n_streams >= 4;
        int64_t acc = 0;
        const double split = b->tune.stream_split > 0.0 && b->tune.stream_split < 1.0 ? b->tune.stream_split : 0.5;
        for (int32_t i : nw1) {
            if (streams4 && (double)acc >= split * (double)nw1_cells) tlane[i] = 3;
            acc += cells[i];
        }
        /* With no LDS sweeps, stream 2 takes the longest several-wavefront forward
           sweeps (within 85% of the longest): the backward passes of the other long
           sub-alignments then start when those finish, not when the longest does. */
        bool any_lds_task = false;
        int64_t mw_max = 0;
        for (int64_t i = 0; i < nt; i++) {
            any_lds_task |= fgrp[i] == 0;
            if (fgrp[i] >= 7 && fgrp[i] != kPG) mw_max = std::max(mw_max, nd[i]);
        }
        if (streams4 && !any_lds_task && mw_max > 0) {
            int n_head = 0, n_mw = 0;
            for (int64_t i = 0; i < nt; i++) n_mw += fgrp[i] >= 7 && fgrp[i] != kPG;
            for (int64_t i = 0; i < nt; i++)
                if (fgrp[i] >= 7 && fgrp[i] != kPG && (double)nd[i] >= 0.85 * (double)mw_max && n_head < n_mw - 1) {
                    tlane[i] = 2;
                    n_head++;
                }
        }
    }
    struct Launch { int lane, group; int64_t off, n, cells; };
    std::vector<Launch> fl, bl;
    std::vector<int32_t> task_order, win_order;
    int32_t lds_maxw = 1;
    for (int lane = 0; lane < kLanes; lane++) {
        for (int g = 0; g < kFG; g++) {
            Launch L{lane, g, (int64_t)task_order.size(), 0, 0};
            for (int64_t i = 0; i < nt; i++)
                if (tlane[i] == lane && fgrp[i] == g) { task_order.push_back((int32_t)i); L.cells += cells[i]; }
            L.n = (int64_t)task_order.size() - L.off;
            std::stable_sort(task_order.begin() + L.off, task_order.end(),
                             [&](int32_t x, int32_t y) { return cells[x] > cells[y]; });
            if (L.n) fl.push_back(L);
        }
        for (int g = 0; g < kBG; g++) {
            Launch L{lane, g, (int64_t)win_order.size(), 0, 0};
            for (int64_t i = 0; i < nt; i++) {
                if (tlane[i] != lane) continue;
                if (g == 0 && bgrp[i] == 0) lds_maxw = std::max(lds_maxw, maxw[i]);
                for (int32_t w = 0; w < nwin[i]; w++)
                    if (wgrp[task_first_win[i] + w] == g) {
                        win_order.push_back(task_first_win[i] + w);
                        L.cells += wins[task_first_win[i] + w].cells;
                    }
            }
            L.n = (int64_t)win_order.size() - L.off;
            std::stable_sort(win_order.begin() + L.off, win_order.end(), [&](int32_t x, int32_t y) {
                return (int64_t)(wins[x].D - wins[x].T) * maxw[wins[x].task] >
                       (int64_t)(wins[y].D - wins[y].T) * maxw[wins[y].task];
            });
            if (L.n) bl.push_back(L);
        }
    }
    for (int64_t i = 0; i < nt; i++)
        if (fgrp[i] == 0) lds_maxw = std::max(lds_maxw, maxw[i]);
    const int ring = next_pow2(lds_maxw);
    bool any_lds = false;
    int64_t lds_nf = 0, lds_nb = 0; /* workgroups of the largest LDS-sweep launch per direction */
    for (const Launch &L : fl)
        if (L.group == 0) { any_lds = true; lds_nf = std::max(lds_nf, L.n); }
    for (const Launch &L : bl)
        if (L.group == 0) { any_lds = true; lds_nb = std::max(lds_nb, L.n); }
    if (any_lds && ring > (1 << 20)) return fail(ctx, SA_ERR_UNSUPPORTED, "band diagonal wider than 2^20");
    /* LDS sweeps keep their rings in LDS up to 64 KiB per workgroup, else in a slice of
       HBM per workgroup (all LDS launches run on one stream, one after another, so one
       buffer serves them) */
    constexpr size_t kLdsMax = 65536;
    const size_t f_bytes = sa::forward_block_bytes<T>(ring), b_bytes = sa::backward_block_bytes<T>(ring);
    const bool f_hbm = any_lds && f_bytes > kLdsMax, b_hbm = any_lds && b_bytes > kLdsMax;
    const size_t f_stride = (f_bytes + 255) / 256 * 256, b_stride = (b_bytes + 255) / 256 * 256;
    if (f_hbm || b_hbm)
        SA_HIP(ctx, b->d_ring_g.ensure(std::max(f_hbm ? (size_t)lds_nf * f_stride : 0,
                                                b_hbm ? (size_t)lds_nb * b_stride : 0)));
    SA_HIP(ctx, b->d_fM.ensure(sizeof(T) * total_cells));
    SA_HIP(ctx, b->d_fX.ensure(sizeof(T) * std::max<int64_t>(total_xy, 1)));
    SA_HIP(ctx, b->d_fY.ensure(sizeof(T) * std::max<int64_t>(total_xy, 1)));
    int rc;
    /* HDP emission plane over the sub-alignments of the register sweeps: their columns
       grouped by density row for the column pass (sorted on the device by the rows of
       the k-mer ids prep wrote),
       their diagonals numbered in chunks of kPlaneDiags for the transpose.  The column
       plane lives in the forward match plane, which the sweeps fill only after it. */
    /* Per stream lane (each lane's stream builds the plane of its own sub-alignments
       before its sweeps, so the long sweeps start at once and the plane of the others
       overlaps them): column groups and chunks of each lane, in lane order. */
    std::vector<int32_t> em_ctask, em_chunk0(nt, 0), em_grp;
    int64_t em_g[kLanes + 1] = {0, 0, 0, 0, 0}, em_c[kLanes + 1] = {0, 0, 0, 0, 0};
    int n_em = 0;
    if (a.em_plane) {
        for (int64_t i = 0; i < nt; i++) n_em += fgrp[i] != 0;
    }
    if (n_em) {
        /* the columns of each lane's sub-alignments, in lane order: chunks for the
           transpose here, the grouping by density row on the device */
        std::vector<sa::PlaneTask> pts;
        int64_t n_cols = 0;
        for (int lane = 0; lane < kLanes; lane++) {
            for (int64_t i = 0; i < nt; i++) {
                if (fgrp[i] == 0 || tlane[i] != lane) continue;
                em_chunk0[i] = (int32_t)em_ctask.size();
                const int64_t nc = (nd[i] + 1 + sa::kPlaneDiags - 1) / sa::kPlaneDiags;
                if ((int64_t)em_ctask.size() + nc > INT32_MAX)
                    return fail(ctx, SA_ERR_UNSUPPORTED, "emission plane: too many diagonals");
                em_ctask.insert(em_ctask.end(), (size_t)nc, (int32_t)i);
                sa::PlaneTask q{};
                q.pos_off = b->tasks[i].pos_off;
                q.col0 = n_cols;
                q.task = (int32_t)i;
                q.lane = lane;
                q.lX = b->tasks[i].lX;
                pts.push_back(q);
                n_cols += (int64_t)b->tasks[i].lX + 1;
            }
            em_c[lane + 1] = (int64_t)em_ctask.size();
        }
        if (2 * n_cols > INT32_MAX) return fail(ctx, SA_ERR_UNSUPPORTED, "emission plane: too many columns");
        const int32_t nrow = model->n_rows;
        const size_t nkey = (size_t)sa::kPlaneLanes * (size_t)(nrow + 1);
        const size_t tmp = sa::plane_sort_temp_bytes(n_cols, nrow);
        if (!tmp) return fail(ctx, SA_ERR_HIP, "emission plane: column sort");
        SA_HIP(ctx, b->d_pk.ensure(sizeof(uint32_t) * 2 * (size_t)n_cols));
        SA_HIP(ctx, b->d_pv.ensure(sizeof(uint64_t) * (size_t)n_cols));
        SA_HIP(ctx, b->d_em_ent.ensure(sizeof(uint64_t) * (size_t)n_cols));
        SA_HIP(ctx, b->d_runs.ensure(sizeof(int32_t) * (2 * nkey + 1)));
        SA_HIP(ctx, b->d_sort_tmp.ensure(tmp));
        if ((rc = upload(ctx, b->d_em_pt, pts))) return rc;
        uint32_t *keys = b->d_pk.as<uint32_t>();
        int32_t *runs = b->d_runs.as<int32_t>();
        SA_HIP(ctx, sa::launch_plane_sort(b->d_em_pt.as<sa::PlaneTask>(), (int)pts.size(), n_cols, b->d_kid.as<int32_t>(),
                                          a.m.hrow, nrow, keys, keys + n_cols,
                                          b->d_pv.as<uint64_t>(), b->d_em_ent.as<uint64_t>(), (uint32_t *)runs,
                                          runs + nkey, runs + 2 * nkey, b->d_sort_tmp.p, tmp, st));
        int32_t nr = 0;
        SA_HIP(ctx, hipMemcpyAsync(&nr, runs + 2 * nkey, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        SA_HIP(ctx, hipStreamSynchronize(st));
        if (nr < 0 || (size_t)nr > nkey) return fail(ctx, SA_ERR_HIP, "emission plane: column sort");
        std::vector<int32_t> rk((size_t)nr), rn((size_t)nr);
        if (nr) {
            SA_HIP(ctx, hipMemcpyAsync(rk.data(), runs, sizeof(int32_t) * nr, hipMemcpyDeviceToHost, st));
            SA_HIP(ctx, hipMemcpyAsync(rn.data(), runs + nkey, sizeof(int32_t) * nr, hipMemcpyDeviceToHost, st));
            SA_HIP(ctx, hipStreamSynchronize(st));
        }
        /* groups of at most kColumnGroup columns of one (lane, row) key; keys ascend,
           so each lane's groups are contiguous and in row order */
        int64_t e0 = 0;
        int lane_done = 0;
        for (int32_t k = 0; k < nr; k++) {
            const int lane = (int)((uint32_t)rk[k] / (uint32_t)(nrow + 1));
            const int32_t r = (int32_t)((uint32_t)rk[k] % (uint32_t)(nrow + 1)) - 1;
            if (lane >= kLanes || rn[k] <= 0) return fail(ctx, SA_ERR_HIP, "emission plane: column sort");
            for (; lane_done < lane; lane_done++) em_g[lane_done + 1] = (int64_t)(em_grp.size() / 3);
            for (int64_t e = 0; e < rn[k]; e += sa::kColumnGroup) {
                em_grp.push_back(r);
                em_grp.push_back((int32_t)(e0 + e));
                em_grp.push_back((int32_t)std::min<int64_t>(sa::kColumnGroup, rn[k] - e));
            }
            e0 += rn[k];
        }
        for (; lane_done < kLanes; lane_done++) em_g[lane_done + 1] = (int64_t)(em_grp.size() / 3);
        if (e0 != n_cols) return fail(ctx, SA_ERR_HIP, "emission plane: column sort");
        SA_HIP(ctx, b->d_em.ensure(sizeof(T) * std::max<int64_t>(total_cells, 1)));
        if ((rc = upload(ctx, b->d_em_tasks, em_ctask)) || (rc = upload(ctx, b->d_em_chunk0, em_chunk0)) ||
            (rc = upload(ctx, b->d_em_grp, em_grp)))
            return rc;
        a.em = b->d_em.as<T>();
        a.emc = b->d_fM.as<T>();
    }
    if ((rc = upload(ctx, b->d_tasks, tasks)) || (rc = upload(ctx, b->d_wins, wins)) ||
        (rc = upload(ctx, b->d_task_order, task_order)) || (rc = upload(ctx, b->d_win_order, win_order)))
        return rc;
    SA_HIP(ctx, b->d_out.ensure(sizeof(int32_t) * 3 * std::max<int64_t>(out_total, 1)));
    SA_HIP(ctx, b->d_win_count.ensure(sizeof(int32_t) * std::max<int64_t>(nw, 1)));
    SA_HIP(ctx, b->d_win_total.ensure(sizeof(double) * std::max<int64_t>(nw, 1)));
    a.tasks = b->d_tasks.as<TaskDev>();
    a.wins = b->d_wins.as<WinDev>();
    a.fM = b->d_fM.as<T>(); a.fX = b->d_fX.as<T>(); a.fY = b->d_fY.as<T>();
    a.out = b->d_out.as<int32_t>();
    a.win_count = b->d_win_count.as<int32_t>();
    a.win_total = b->d_win_total.as<double>();
    a.ring = ring;
    const int64_t nk = model->n_kmers;
    auto zero_expect = [&]() -> int {
        SA_HIP(ctx, hipMemsetAsync(b->d_ex_trans.p, 0, sizeof(double) * 10 * std::max<int64_t>(nw, 1), st));
        SA_HIP(ctx, hipMemsetAsync(b->d_ex_s0.p, 0, sizeof(double) * nt * nk, st));
        SA_HIP(ctx, hipMemsetAsync(b->d_ex_s1.p, 0, sizeof(double) * nt * nk, st));
        SA_HIP(ctx, hipMemsetAsync(b->d_ex_obs.p, 0, sizeof(int8_t) * nt * nk, st));
        SA_HIP(ctx, hipMemsetAsync(b->d_ex_count.p, 0, sizeof(int32_t) * std::max<int64_t>(nw, 1), st));
        return SA_OK;
    };
    if (expect) {
        SA_HIP(ctx, b->d_ex_trans.ensure(sizeof(double) * 10 * std::max<int64_t>(nw, 1)));
        SA_HIP(ctx, b->d_ex_s0.ensure(sizeof(double) * nt * nk));
        SA_HIP(ctx, b->d_ex_s1.ensure(sizeof(double) * nt * nk));
        SA_HIP(ctx, b->d_ex_obs.ensure(sizeof(int8_t) * nt * nk));
        SA_HIP(ctx, b->d_ex_rec.ensure(sizeof(sa::ExRec) * std::max<int64_t>(rec_total, 1)));
        SA_HIP(ctx, b->d_ex_count.ensure(sizeof(int32_t) * std::max<int64_t>(nw, 1)));
        if ((rc = zero_expect())) return rc;
        a.ex_trans = b->d_ex_trans.as<double>();
        a.ex_s0 = b->d_ex_s0.as<double>();
        a.ex_s1 = b->d_ex_s1.as<double>();
        a.ex_obs = b->d_ex_obs.as<int8_t>();
        a.ex_rec = b->d_ex_rec.as<sa::ExRec>();
        a.ex_count = b->d_ex_count.as<int32_t>();
    }
    const int32_t *d_to = b->d_task_order.as<int32_t>();
    const int32_t *d_wo = b->d_win_order.as<int32_t>();
    int64_t n_lds = 0;
    for (const Launch &L : fl) if (L.group == 0) n_lds += L.n;
    const int lds_fnt = (n_lds >= 4 * std::max(ctx->n_cu, 1) || lds_maxw <= 64) && lds_maxw <= 256 ? 64 : 256;
    const int ek = n_em ? 3 : model->ek;
    if ((int64_t)(fl.size() + bl.size()) * 2 > (int64_t)(sizeof(ctx->gev) / sizeof(ctx->gev[0])))
        return fail(ctx, SA_ERR_HIP, "too many launches for the timing events");

    hipStream_t gst[kLanes] = {st, ctx->aux[0], ctx->aux[1], ctx->aux[2]};
    /* occupancy cap: dynamic LDS per one-wavefront workgroup so that at most
       waves_per_simd of them share a SIMD (160 KiB of LDS per CU, four SIMDs) */
    const int wps = b->tune.waves_per_simd;
    const size_t shmem1 = wps > 0 && wps < 8 ? (size_t)std::max(0, 163840 / (4 * wps) - 4096) : 0;
    if (n_pers) {
        SA_HIP(ctx, b->d_pq.ensure(sizeof(int32_t) * (size_t)(3 + n_pers + nt)));
        if ((rc = upload(ctx, b->d_first_win, task_first_win))) return rc;
    }
    auto launch_fwd = [&](const Launch &L, hipStream_t sg) -> hipError_t {
        if (L.group == kPG) {
            /* queue counters zeroed by a kernel (graph-safe), then the workers: enough
               for every item, at most persist_waves per SIMD */
            int32_t *pq = b->d_pq.as<int32_t>();
            hipError_t e = sa::launch_zero(pq, (int)(3 + L.n + nt), sg);
            if (e != hipSuccess) return e;
            sa::PersistQ q;
            q.ctr = pq;
            q.done = pq + 3;
            q.wclaim = pq + 3 + L.n;
            q.first_win = b->d_first_win.as<int32_t>();
            q.order = d_to + L.off;
            q.n_fwd = (int32_t)L.n;
            q.zero = 0;
            q.wait_ticks = 200000000; /* 2 s; a step takes milliseconds */
            const int pw = b->tune.persist_waves;
            const int64_t workers = std::min<int64_t>(L.n + n_pers_win, (int64_t)4 * std::max(ctx->n_cu, 1) * pw);
            return sa::launch_fb_persistent<T>(a, q, (int)workers, ek, sg);
        }
        if (L.group == 0) {
            sa::Args<T> al = a;
            al.task_order = d_to + L.off;
            al.ring_g = f_hbm ? b->d_ring_g.as<unsigned char>() : nullptr;
            al.ring_g_stride = (int64_t)f_stride;
            return sa::launch_forward<T>(al, (int)L.n, lds_fnt, sg);
        }
        /* wave priorities follow the dependency chains: the longest several-wavefront
           sweeps (stream 2 without LDS work) 3, the other several-wavefront ones 2,
           the longer one-wavefront half (stream 0) 1, the rest 0 */
        sa::Args<T> ap = a;
        ap.prio = L.group >= 7 ? (L.lane == 2 ? 3 : 2) : L.lane == 0 ? 1 : 0;
        if (L.group < 7)
            return sa::launch_forward_reg<T>(ap, d_to + L.off, (int)L.n, kCpt[L.group - 1], 1, ek, shmem1, sg);
        if (L.group >= 11) return sa::launch_forward_halo<T>(ap, d_to + L.off, (int)L.n, L.group - 11 + 2, ek, sg);
        return sa::launch_forward_reg<T>(ap, d_to + L.off, (int)L.n, kMW[L.group - 7][1], kMW[L.group - 7][0], ek,
                                         0, sg);
    };
    auto launch_bwd = [&](const Launch &L, hipStream_t sg) -> hipError_t {
        if (L.group == 0) {
            sa::Args<T> al = a;
            al.win_order = d_wo + L.off;
            al.ring_g = b_hbm ? b->d_ring_g.as<unsigned char>() : nullptr;
            al.ring_g_stride = (int64_t)b_stride;
            return sa::launch_backward<T>(al, (int)L.n, sg);
        }
        return sa::launch_backward_reg<T>(a, d_wo + L.off, (int)L.n, kCpt[L.group - 1], ek, shmem1, sg);
    };
    /* fork on ev[2], per-launch events gev[2k], gev[2k+1] (forward launches first),
       join on ev[4] */
    auto run_sweeps = [&](bool forward) -> int {
        SA_HIP(ctx, hipEventRecord(ctx->ev[2], st));
        for (int q = 1; q < kLanes; q++)
            if (gst[q] != st) SA_HIP(ctx, hipStreamWaitEvent(gst[q], ctx->ev[2], 0));
        for (int lane = 0; lane < kLanes; lane++) {
            hipStream_t sg = gst[lane];
            if (forward && em_c[lane + 1] > em_c[lane]) {
                SA_HIP(ctx, hipEventRecord(ctx->ev[16 + 2 * lane], sg));
                SA_HIP(ctx, sa::launch_hdp_plane<T>(a, b->d_em_ent.as<int32_t>(),
                                                    b->d_em_grp.as<int32_t>() + 3 * em_g[lane],
                                                    (int)(em_g[lane + 1] - em_g[lane]), b->d_em_tasks.as<int32_t>(),
                                                    b->d_em_chunk0.as<int32_t>(), (int)em_c[lane],
                                                    (int)em_c[lane + 1], ctx->n_cu, sg));
                SA_HIP(ctx, hipEventRecord(ctx->ev[17 + 2 * lane], sg));
            }
            if (forward)
                for (size_t k = 0; k < fl.size(); k++) {
                    if (fl[k].lane != lane) continue;
                    SA_HIP(ctx, hipEventRecord(ctx->gev[2 * k], sg));
                    SA_HIP(ctx, launch_fwd(fl[k], sg));
                    SA_HIP(ctx, hipEventRecord(ctx->gev[2 * k + 1], sg));
                }
            for (size_t k = 0; k < bl.size(); k++) {
                if (bl[k].lane != lane) continue;
                const size_t e = 2 * (fl.size() + k);
                SA_HIP(ctx, hipEventRecord(ctx->gev[e], sg));
                SA_HIP(ctx, launch_bwd(bl[k], sg));
                SA_HIP(ctx, hipEventRecord(ctx->gev[e + 1], sg));
            }
        }
        for (int q = 1; q < kLanes; q++) {
            if (gst[q] == st) continue;
            SA_HIP(ctx, hipEventRecord(ctx->ev[8 + q], gst[q]));
            SA_HIP(ctx, hipStreamWaitEvent(st, ctx->ev[8 + q], 0));
        }
        SA_HIP(ctx, hipEventRecord(ctx->ev[4], st));
        return SA_OK;
    };
    if ((rc = run_sweeps(true))) return rc;
    tm.forward_launches = (int64_t)fl.size();
    tm.backward_launches = (int64_t)bl.size();
    std::vector<int32_t> wcount(nw);
    SA_HIP(ctx, hipMemcpyAsync(wcount.data(), b->d_win_count.p, sizeof(int32_t) * nw, hipMemcpyDeviceToHost, st));
    SA_HIP(ctx, hipMemcpyAsync(&err, b->d_err.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    SA_HIP(ctx, hipStreamSynchronize(st));
    if (err & sa::ERR_WINDOWS) return kRetryLds;
    if (err & sa::ERR_QUEUE) return fail(ctx, SA_ERR_HIP, "persistent sweeps: a worker waited too long for a forward sweep");
    {
        float fb = 0.f;
        (void)hipEventElapsedTime(&fb, ctx->ev[2], ctx->ev[4]);
        tm.fb_ms = fb;
        double fwd_end = 0.0, bwd_start = fb;
        auto record = [&](const Launch &L, int dir, size_t e) {
            float t0 = 0.f, t1 = 0.f;
            (void)hipEventElapsedTime(&t0, ctx->ev[2], ctx->gev[e]);
            (void)hipEventElapsedTime(&t1, ctx->ev[2], ctx->gev[e + 1]);
            sa_launch li;
            std::memset(&li, 0, sizeof(li));
            li.direction = dir;
            const bool pl = dir == 0 && L.group == kPG;
            li.cpt = L.group == 0 ? 0 : pl ? 3 : dir == 0 && L.group >= 11 ? 1 : dir == 0 && L.group >= 7 ? kMW[L.group - 7][1]
                                                                                                        : kCpt[L.group - 1];
            li.waves = pl ? 1 : dir == 0 && L.group >= 11 ? L.group - 11 + 2 : dir == 0 && L.group >= 7 ? kMW[L.group - 7][0] : 1;
            li.persistent = pl ? 1 : 0;
            li.stream = L.lane;
            li.sync_every = pl ? 0 : dir == 0 && L.group >= 11 ? sa::kHaloK : dir == 0 && L.group >= 7 ? 1 : 0;
            li.items = L.n;
            li.cells = L.cells;
            li.start_ms = t0;
            li.ms = t1 - t0;
            res->launches.push_back(li);
            if (dir == 0) fwd_end = std::max(fwd_end, (double)t1);
            else bwd_start = std::min(bwd_start, (double)t0);
        };
        for (size_t k = 0; k < fl.size(); k++) record(fl[k], 0, 2 * k);
        for (size_t k = 0; k < bl.size(); k++) record(bl[k], 1, 2 * (fl.size() + k));
        tm.forward_ms = fwd_end;
        tm.backward_ms = fb - bwd_start;
    }
    /* windows whose pairs (or expectation records) overflowed their slot run again
       with the exact count */
    std::vector<int32_t> redo;
    for (int64_t w = 0; w < nw; w++)
        if (wcount[w] > wins[w].out_cap) redo.push_back((int32_t)w);
    bool ex_redo = false;
    if (expect) {
        std::vector<int32_t> ecnt(nw);
        SA_HIP(ctx, hipMemcpy(ecnt.data(), b->d_ex_count.p, sizeof(int32_t) * nw, hipMemcpyDeviceToHost));
        rec_total = 0;
        for (int64_t w = 0; w < nw; w++) {
            if (ecnt[w] > wins[w].rec_cap) {
                wins[w].rec_cap = ecnt[w];
                ex_redo = true;
            }
            wins[w].rec_off = rec_total;
            rec_total += wins[w].rec_cap;
        }
        if (ex_redo) {
            SA_HIP(ctx, b->d_ex_rec.ensure(sizeof(sa::ExRec) * std::max<int64_t>(rec_total, 1)));
            a.ex_rec = b->d_ex_rec.as<sa::ExRec>();
        }
    }
    if (!redo.empty() || ex_redo) {
        int64_t extra = 0;
        for (int32_t w : redo) {
            wins[w].out_off = out_total + extra;
            wins[w].out_cap = wcount[w];
            extra += wcount[w];
        }
        /* ensure() may reallocate and drop the first-pass pairs, so every window runs
           again into the resized buffer */
        SA_HIP(ctx, b->d_out.ensure(sizeof(int32_t) * 3 * (out_total + extra)));
        if ((rc = upload(ctx, b->d_wins, wins))) return rc;
        a.out = b->d_out.as<int32_t>();
        a.wins = b->d_wins.as<WinDev>();
        if (expect && (rc = zero_expect())) return rc; /* the sums start again */
        if ((rc = run_sweeps(false))) return rc;
        tm.backward_launches += (int64_t)bl.size();
        SA_HIP(ctx, hipMemcpyAsync(wcount.data(), b->d_win_count.p, sizeof(int32_t) * nw, hipMemcpyDeviceToHost, st));
        SA_HIP(ctx, hipStreamSynchronize(st));
    }
    tm.emission_path = ek;
    tm.emission_ms = plane_ms(ctx, em_c);
    /* final positions in the reference's list order: per read, sub-alignments in
       split order, each in reverse emission order */
    std::vector<int64_t> win_dst(nw);
    int64_t pos = 0;
    for (int64_t r = 0; r < b->n_reads; r++) {
        ReadOut &ro = res->reads[r];
        ro.first_pair = pos;
        ro.n_sub = b->reads[r].n_tasks;
        for (int64_t ti = b->reads[r].first_task; ti < b->reads[r].first_task + b->reads[r].n_tasks; ti++) {
            int64_t E = 0;
            for (int32_t w = 0; w < nwin[ti]; w++) E += wcount[task_first_win[ti] + w];
            int64_t s = 0;
            for (int32_t w = 0; w < nwin[ti]; w++) {
                win_dst[task_first_win[ti] + w] = pos + E - 1 - s;
                s += wcount[task_first_win[ti] + w];
            }
            pos += E;
            ro.cells += cells[ti];
            ro.n_windows += nwin[ti];
        }
        ro.n_pairs = pos - ro.first_pair;
    }
    const int64_t total_pairs = pos;
    if ((rc = upload(ctx, b->d_win_dst, win_dst))) return rc;
    SA_HIP(ctx, b->d_packed.ensure(sizeof(int32_t) * 3 * std::max<int64_t>(total_pairs, 1)));
    SA_HIP(ctx, hipEventRecord(ctx->ev[5], st));
    SA_HIP(ctx, sa::launch_collect(b->d_wins.as<WinDev>(), b->d_win_count.as<int32_t>(), b->d_win_dst.as<int64_t>(),
                                   b->d_out.as<int32_t>(), b->d_packed.as<int32_t>(), b->d_x1.as<int32_t>(),
                                   b->d_y1.as<int32_t>(), (int)nw, st));
    SA_HIP(ctx, hipEventRecord(ctx->ev[6], st));
    std::vector<double> wtot(nw);
    if ((rc = fetch_pairs(ctx, res, b->d_packed.p, total_pairs, st))) return rc;
    SA_HIP(ctx, hipMemcpyAsync(wtot.data(), b->d_win_total.p, sizeof(double) * nw, hipMemcpyDeviceToHost, st));
    SA_HIP(ctx, hipStreamSynchronize(st));
    finish_pairs(ctx, res, total_pairs);
    if (expect) {
        /* per read: sub-alignments in split order, each window's records in call order */
        std::vector<double> etr(10 * nw), es0(nt * nk), es1(nt * nk);
        std::vector<int8_t> eobs(nt * nk);
        std::vector<int32_t> ecnt(nw);
        std::vector<sa::ExRec> erec(std::max<int64_t>(rec_total, 1));
        SA_HIP(ctx, hipMemcpyAsync(etr.data(), b->d_ex_trans.p, sizeof(double) * 10 * nw, hipMemcpyDeviceToHost, st));
        SA_HIP(ctx, hipMemcpyAsync(es0.data(), b->d_ex_s0.p, sizeof(double) * nt * nk, hipMemcpyDeviceToHost, st));
        SA_HIP(ctx, hipMemcpyAsync(es1.data(), b->d_ex_s1.p, sizeof(double) * nt * nk, hipMemcpyDeviceToHost, st));
        SA_HIP(ctx, hipMemcpyAsync(eobs.data(), b->d_ex_obs.p, nt * nk, hipMemcpyDeviceToHost, st));
        SA_HIP(ctx, hipMemcpyAsync(ecnt.data(), b->d_ex_count.p, sizeof(int32_t) * nw, hipMemcpyDeviceToHost, st));
        if (rec_total)
            SA_HIP(ctx, hipMemcpyAsync(erec.data(), b->d_ex_rec.p, sizeof(sa::ExRec) * rec_total,
                                       hipMemcpyDeviceToHost, st));
        SA_HIP(ctx, hipStreamSynchronize(st));
        res->expect.assign(b->n_reads, ReadExpect());
        for (int64_t r = 0; r < b->n_reads; r++) {
            ReadExpect &re = res->expect[r];
            re.s0.assign(nk, 0.0);
            re.s1.assign(nk, 0.0);
            re.obs.assign(nk, 0);
            for (int64_t ti = b->reads[r].first_task; ti < b->reads[r].first_task + b->reads[r].n_tasks; ti++) {
                for (int64_t k = 0; k < nk; k++) {
                    re.s0[k] += es0[ti * nk + k];
                    re.s1[k] += es1[ti * nk + k];
                    re.obs[k] |= eobs[ti * nk + k];
                }
                for (int32_t w = 0; w < nwin[ti]; w++) {
                    const int64_t wi = task_first_win[ti] + w;
                    for (int k = 0; k < 9; k++) re.trans[k] += etr[10 * wi + k];
                    re.likelihood += etr[10 * wi + 9];
                    const int64_t n = std::min<int64_t>(ecnt[wi], wins[wi].rec_cap);
                    for (int64_t q = 0; q < n; q++) {
                        const sa::ExRec &x = erec[wins[wi].rec_off + q];
                        sa_match_record m;
                        m.x = x.x + b->thost[ti].x1;
                        m.y = x.y + b->thost[ti].y1;
                        m.p[0] = x.p[0];
                        m.p[1] = x.p[1];
                        m.p[2] = x.p[2];
                        re.recs.push_back(m);
                    }
                }
            }
        }
    }
    for (int64_t r = 0; r < b->n_reads; r++) {
        double ll = 0.0;
        for (int64_t ti = b->reads[r].first_task; ti < b->reads[r].first_task + b->reads[r].n_tasks; ti++)
            if (nwin[ti] > 0) ll += wtot[task_first_win[ti] + nwin[ti] - 1];
        res->reads[r].log_likelihood = ll;
    }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]); tm.prep_ms = ms;
    (void)hipEventElapsedTime(&ms, ctx->ev[5], ctx->ev[6]); tm.collect_ms = ms;
    (void)hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[6]); tm.total_ms = ms;
    tm.device_start_ms = sa_rt::timeline_ms(ctx->device, ctx->ev[0]);
    tm.device_end_ms = sa_rt::timeline_ms(ctx->device, ctx->ev[6]);
    tm.cells = total_cells;

    /* record the launch sequence of this (now planned) batch as a graph: later runs
       replay it without the host round trips between the steps */
    auto &gc = b->gc[sizeof(T) == 4 ? 1 : 0];
    gc.reset();
    /* only a batch that is run again is worth recording (sa_align runs once); an
       expectations run reads its accumulators back after each run, so it is not */
    if (b->tune.use_graph && !force_lds && !expect && redo.empty() && b->nruns[sizeof(T) == 4 ? 1 : 0] >= 1) {
        hipGraph_t g = nullptr;
        bool ok = hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed) == hipSuccess;
        if (ok) {
            const bool enq =
                sa::launch_zero(b->d_err.as<int32_t>(), 1, st) == hipSuccess &&
                hipEventRecord(ctx->ev[0], st) == hipSuccess && sa::launch_prep<T>(a, (int)nt, st) == hipSuccess &&
                hipEventRecord(ctx->ev[1], st) == hipSuccess && run_sweeps(true) == SA_OK &&
                hipEventRecord(ctx->ev[5], st) == hipSuccess &&
                sa::launch_collect(b->d_wins.as<WinDev>(), b->d_win_count.as<int32_t>(), b->d_win_dst.as<int64_t>(),
                                   b->d_out.as<int32_t>(), b->d_packed.as<int32_t>(), b->d_x1.as<int32_t>(),
                                   b->d_y1.as<int32_t>(), (int)nw, st) == hipSuccess &&
                hipEventRecord(ctx->ev[6], st) == hipSuccess;
            const hipError_t ec = hipStreamEndCapture(st, &g);
            ok = enq && ec == hipSuccess && g != nullptr &&
                 hipGraphInstantiate(&gc.exec, g, nullptr, nullptr, 0) == hipSuccess;
            if (g) (void)hipGraphDestroy(g);
        }
        if (ok) {
            gc.reads = res->reads;
            gc.last_win.assign(nt, -1);
            for (int64_t ti = 0; ti < nt; ti++)
                if (nwin[ti] > 0) gc.last_win[ti] = task_first_win[ti] + nwin[ti] - 1;
            gc.launches = res->launches;
            gc.nw = nw;
            gc.total_pairs = total_pairs;
            gc.cells = total_cells;
            gc.ek = ek;
            std::copy(em_c, em_c + 5, gc.em_c);
        } else {
            gc.reset();
            (void)hipGetLastError();
            ctx->err[0] = 0;
        }
    }
    return SA_OK;
}

/* A batch with degenerate references: the multi-path sweeps (sa_paths.hip), FP64,
 * every sub-alignment's forward sweep in one launch, then every window's backward
 * pass; pairs carry their path.  The path tables are built here from the reference
 * bytes: per column its 'X' sites and paths, per entry its substituted k-mer's id. */
int run_paths(sa_batch *b, sa_result *res) {
    sa_context *ctx = b->ctx;
    const sa_model *model = b->model;
    hipStream_t st = ctx->stream;
    const int64_t nt = (int64_t)b->tasks.size();
    sa_timing &tm = res->timing;
    std::memset(&tm, 0, sizeof(tm));
    res->reads.assign(b->n_reads, ReadOut());
    if (b->expect) return fail(ctx, SA_ERR_UNSUPPORTED, "expectations of degenerate references");
    if (nt == 0) {
        for (int64_t r = 0; r < b->n_reads; r++) res->reads[r].n_sub = 0;
        return SA_OK;
    }
    const int k = model->k, A = model->A;
    const int B = (int)b->options.size();
    SA_HIP(ctx, b->d_kid.ensure(sizeof(int32_t) * (b->total_pos + 64)));
    SA_HIP(ctx, b->d_evm.ensure(sizeof(double) * (b->total_ev + 64)));
    SA_HIP(ctx, b->d_evn.ensure(sizeof(double) * (b->total_ev + 64)));
    SA_HIP(ctx, b->d_evl.ensure(sizeof(double) * (b->total_ev + 64)));
    SA_HIP(ctx, b->d_evi.ensure(sizeof(double) * (b->total_ev + 64)));
    SA_HIP(ctx, hipMemsetAsync(b->d_evm.p, 0, sizeof(double) * (b->total_ev + 64), st));
    SA_HIP(ctx, hipMemsetAsync(b->d_evn.p, 0, sizeof(double) * (b->total_ev + 64), st));
    SA_HIP(ctx, hipMemsetAsync(b->d_evl.p, 0, sizeof(double) * (b->total_ev + 64), st));
    SA_HIP(ctx, hipMemsetAsync(b->d_evi.p, 0, sizeof(double) * (b->total_ev + 64), st));
    SA_HIP(ctx, b->d_xlo.ensure(sizeof(int32_t) * b->total_diag));
    SA_HIP(ctx, b->d_xhi.ensure(sizeof(int32_t) * b->total_diag));
    SA_HIP(ctx, b->d_doff.ensure(sizeof(int32_t) * b->total_diag));
    SA_HIP(ctx, b->d_foff.ensure(sizeof(double) * b->total_diag));
    SA_HIP(ctx, b->d_dflag.ensure(sizeof(int8_t) * b->total_diag + 4));
    SA_HIP(ctx, b->d_dcls.ensure(sizeof(int8_t) * b->total_diag + 4));
    SA_HIP(ctx, b->d_doffxy.ensure(sizeof(int32_t) * b->total_diag));
    SA_HIP(ctx, b->d_task_cells.ensure(sizeof(int64_t) * nt));
    SA_HIP(ctx, b->d_task_xycells.ensure(sizeof(int64_t) * nt));
    SA_HIP(ctx, b->d_task_maxw.ensure(sizeof(int32_t) * nt));
    SA_HIP(ctx, b->d_task_span.ensure(sizeof(int32_t) * nt));
    SA_HIP(ctx, b->d_task_nwin.ensure(sizeof(int32_t) * nt));
    SA_HIP(ctx, b->d_win_slots.ensure(sizeof(WinDev) * std::max<int64_t>(b->total_wslots, 1)));
    SA_HIP(ctx, b->d_err.ensure(sizeof(int32_t)));
    SA_HIP(ctx, hipMemsetAsync(b->d_err.p, 0, sizeof(int32_t), st));

    const DevTables<double> &tab = tables<double>(model);
    sa::Args<double> a;
    std::memset(&a, 0, sizeof(a));
    a.m = device_model<double>(model, tab);
    a.tasks = b->d_tasks.as<TaskDev>();
    a.task_order = b->d_prep_order.as<int32_t>(); /* (the sweeps' launches set their own) */
    a.ref = b->d_ref.as<char>();
    a.events = b->d_events.as<double>();
    a.event_stride = b->stride;
    a.anc_x = b->d_ancx.as<int32_t>();
    a.anc_y = b->d_ancy.as<int32_t>();
    a.lut = model->d_lut.as<int8_t>();
    a.kid = b->d_kid.as<int32_t>();
    a.evm = b->d_evm.as<double>(); a.evn = b->d_evn.as<double>(); a.evl = b->d_evl.as<double>();
    a.evi = b->d_evi.as<double>();
    a.xlo = b->d_xlo.as<int32_t>(); a.xhi = b->d_xhi.as<int32_t>(); a.doff = b->d_doff.as<int32_t>();
    a.foff = b->d_foff.as<double>();
    a.dflag = b->d_dflag.as<int8_t>();
    a.dcls = b->d_dcls.as<int8_t>();
    a.doffxy = b->d_doffxy.as<int32_t>();
    a.task_cells = b->d_task_cells.as<int64_t>();
    a.task_xycells = b->d_task_xycells.as<int64_t>();
    a.task_maxw = b->d_task_maxw.as<int32_t>();
    a.task_span = b->d_task_span.as<int32_t>();
    a.task_nwin = b->d_task_nwin.as<int32_t>();
    a.win_slots = b->d_win_slots.as<WinDev>();
    a.err = b->d_err.as<int32_t>();
    a.threshold = b->params.threshold;
    a.min_diags = b->params.min_diags_between_traceback;
    a.traceback = b->params.traceback_diagonals;
    a.expansion = b->params.diagonal_expansion;
    a.paths = 1;

    SA_HIP(ctx, hipEventRecord(ctx->ev[0], st));
    SA_HIP(ctx, sa::launch_prep<double>(a, (int)nt, st));
    SA_HIP(ctx, hipEventRecord(ctx->ev[1], st));
    std::vector<int64_t> cells(nt);
    std::vector<int32_t> nwin(nt), xlo(b->total_diag), xhi(b->total_diag);
    std::vector<WinDev> slots(b->total_wslots);
    int32_t err = 0;
    SA_HIP(ctx, hipMemcpyAsync(cells.data(), b->d_task_cells.p, sizeof(int64_t) * nt, hipMemcpyDeviceToHost, st));
    SA_HIP(ctx, hipMemcpyAsync(nwin.data(), b->d_task_nwin.p, sizeof(int32_t) * nt, hipMemcpyDeviceToHost, st));
    SA_HIP(ctx, hipMemcpyAsync(slots.data(), b->d_win_slots.p, sizeof(WinDev) * b->total_wslots,
                               hipMemcpyDeviceToHost, st));
    SA_HIP(ctx, hipMemcpyAsync(xlo.data(), b->d_xlo.p, sizeof(int32_t) * b->total_diag, hipMemcpyDeviceToHost, st));
    SA_HIP(ctx, hipMemcpyAsync(xhi.data(), b->d_xhi.p, sizeof(int32_t) * b->total_diag, hipMemcpyDeviceToHost, st));
    SA_HIP(ctx, hipMemcpyAsync(&err, b->d_err.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    SA_HIP(ctx, hipStreamSynchronize(st));
    if (err & sa::ERR_BAD_KMER) return fail(ctx, SA_ERR_BAD_KMER, "reference k-mer with a character outside the alphabet");
    if (err & sa::ERR_BAND) return fail(ctx, SA_ERR_ANCHORS, "invalid band (anchors outside the matrix or out of order)");
    if (err & sa::ERR_WINDOWS) return fail(ctx, SA_ERR_HIP, "window schedule overflow");

    /* path tables, tasks spread over worker threads, in two passes: the entries of
       each task (per column, one path per combination of its k-mer's 'X' options),
       then every table written in place: per column its first entry and site range,
       per entry its column, k-mer id and legal predecessors, per diagonal its band,
       entries and plane offset */
    auto ipow = [](int64_t base, int e) {
        int64_t r = 1;
        for (int i = 0; i < e; i++) r *= base;
        return r;
    };
    const int nthreads = (int)std::max<int64_t>(1, std::min<int64_t>(nt, 16));
    auto parallel = [&](auto &&fn) {
        std::atomic<int64_t> next{0};
        auto worker = [&]() {
            for (int64_t ti; (ti = next.fetch_add(1)) < nt;) fn(ti);
        };
        std::vector<std::thread> th;
        for (int i = 1; i < nthreads; i++) th.emplace_back(worker);
        worker();
        for (auto &x : th) x.join();
    };
    /* the 'X' sites of a task's bases and, per column x >= 1, the sites [s0, s1) of
       its k-mer ref[x-1 .. x-1+k) */
    auto sites_of = [&](int64_t ti, std::vector<int32_t> &site) {
        const TaskDev &t = b->tasks[ti];
        const char *ref = b->href.data() + t.ref_off;
        site.clear();
        for (int64_t i = 0; i < (int64_t)t.lX + k - 1; i++)
            if (ref[i] == 'X') site.push_back((int32_t)i);
    };
    std::vector<int64_t> n_ent(nt), planes_t(nt);
    std::vector<int32_t> maxent(nt), terr(nt);
    parallel([&](int64_t ti) {
        std::vector<int32_t> site;
        sites_of(ti, site);
        int64_t n = 1; /* column 0 */
        size_t s0 = 0, s1 = 0;
        for (int64_t x = 1; x <= b->tasks[ti].lX; x++) {
            while (s0 < site.size() && site[s0] < x - 1) s0++;
            while (s1 < site.size() && site[s1] < x - 1 + k) s1++;
            const int64_t np = ipow(B, (int)(s1 - s0));
            if (np > 729 || n + np > INT32_MAX / 2) {
                terr[ti] = 1;
                return;
            }
            n += np;
        }
        n_ent[ti] = n;
    });
    for (int64_t ti = 0; ti < nt; ti++)
        if (terr[ti] == 1) return fail(ctx, SA_ERR_UNSUPPORTED, "too many paths in a degenerate reference");
    std::vector<int64_t> col_off(nt), ent_off(nt), plane_off(nt);
    int64_t ncol = 0, nent = 0;
    for (int64_t ti = 0; ti < nt; ti++) {
        col_off[ti] = ncol;
        ent_off[ti] = nent;
        ncol += (int64_t)b->tasks[ti].lX + 2;
        nent += n_ent[ti];
    }
    std::vector<int32_t> pp(ncol), psa(ncol), psb(ncol), ecol(nent), ekid(nent);
    std::vector<int64_t> doffE(b->total_diag);
    SA_HIP(ctx, b->h_pentr.ensure(sizeof(sa::PathEnt) * (size_t)std::max<int64_t>(nent, 1)));
    SA_HIP(ctx, b->h_pdiag.ensure(sizeof(sa::PathDiag) * (size_t)std::max<int64_t>(b->total_diag, 1)));
    sa::PathEnt *pent = b->h_pentr.as<sa::PathEnt>();
    sa::PathDiag *pdiag = b->h_pdiag.as<sa::PathDiag>();
    parallel([&](int64_t ti) {
        const TaskDev &t = b->tasks[ti];
        const char *ref = b->href.data() + t.ref_off;
        const int64_t lX = t.lX;
        std::vector<int32_t> site;
        sites_of(ti, site);
        int32_t *tpp = pp.data() + col_off[ti], *tsa = psa.data() + col_off[ti], *tsb = psb.data() + col_off[ti];
        int32_t *tcol = ecol.data() + ent_off[ti], *tkid = ekid.data() + ent_off[ti];
        sa::PathEnt *tent = pent + ent_off[ti];
        int32_t P = 0;
        size_t s0 = 0, s1 = 0;
        for (int64_t x = 0; x <= lX; x++) {
            int64_t np = 1;
            if (x >= 1) {
                while (s0 < site.size() && site[s0] < x - 1) s0++;
                while (s1 < site.size() && site[s1] < x - 1 + k) s1++;
                np = ipow(B, (int)(s1 - s0));
            }
            tpp[x] = P;
            tsa[x] = x >= 1 ? (int32_t)s0 : 0;
            tsb[x] = x >= 1 ? (int32_t)s1 : 0;
            for (int64_t q = 0; q < np; q++) {
                sa::PathEnt e{(int32_t)x, (int32_t)q, 0, 1, 0, 0};
                int32_t id = 0;
                if (x >= 1) {
                    /* the path's k-mer: the first 'X' varies slowest */
                    char km[16];
                    std::memcpy(km, ref + x - 1, (size_t)k);
                    int64_t rem = q;
                    for (int j = k - 1; j >= 0; j--)
                        if (km[j] == 'X') {
                            km[j] = b->options[(size_t)(rem % B)];
                            rem /= B;
                        }
                    for (int j = 0; j < k; j++) {
                        const int c = model->lut[(uint8_t)km[j]];
                        if (c < 0) {
                            terr[ti] = 2;
                            return;
                        }
                        id = id * A + c;
                    }
                    /* legal predecessors in column x-1: the shared sites [sa0, sb1) are
                       q's leading digits; every option of the sites leaving, [sa1, sa0) */
                    if (x == 1) {
                        e.pn = 1;
                    } else {
                        const int sa1 = tsa[x - 1], sb1 = tsb[x - 1], sa0 = (int)s0, sb0 = (int)s1;
                        e.pfirst = tpp[x - 1] + (int32_t)(q / ipow(B, sb0 - sb1));
                        e.pstep = (int32_t)ipow(B, sb1 - sa0);
                        e.pn = (int32_t)ipow(B, sa0 - sa1);
                    }
                }
                e.kk = id;
                tcol[P + q] = (int32_t)x;
                tkid[P + q] = id;
                tent[P + q] = e;
            }
            P += (int32_t)np;
        }
        tpp[lX + 1] = P;
        tsa[lX + 1] = 0;
        tsb[lX + 1] = 0;
        const int64_t N = (int64_t)t.lX + t.lY;
        const int32_t *l = xlo.data() + t.diag_off, *h = xhi.data() + t.diag_off;
        int64_t run = 0;
        int32_t mx = 0;
        for (int64_t d = 0; d <= N; d++) {
            const int32_t e0 = tpp[l[d]], e1 = tpp[h[d] + 1];
            doffE[t.diag_off + d] = run;
            pdiag[t.diag_off + d] = sa::PathDiag{l[d], h[d], e0, e1, run}; /* + the task's plane offset below */
            run += e1 - e0;
            mx = std::max<int32_t>(mx, e1 - e0);
        }
        planes_t[ti] = run;
        maxent[ti] = mx;
    });
    for (int64_t ti = 0; ti < nt; ti++)
        if (terr[ti] == 2) return fail(ctx, SA_ERR_BAD_KMER, "degenerate option outside the alphabet");
    int64_t planes = 0, max_ent = 1;
    for (int64_t ti = 0; ti < nt; ti++) {
        plane_off[ti] = planes;
        planes += planes_t[ti];
        max_ent = std::max<int64_t>(max_ent, maxent[ti]);
    }
    parallel([&](int64_t ti) {
        const TaskDev &t = b->tasks[ti];
        for (int64_t d = 0; d <= (int64_t)t.lX + t.lY; d++) pdiag[t.diag_off + d].foff += plane_off[ti];
    });
    /* the ring index wraps with & (ring - 1): a diagonal with more (cell, path) entries
       than the ring holds would alias its own entries */
    if (max_ent > (1 << 24)) return fail(ctx, SA_ERR_UNSUPPORTED, "a diagonal holds more than 2^24 (cell, path) entries");
    const int ring = next_pow2((int)max_ent);
    /* windows and their pair slots: a first guess of a few pairs per posterior
       diagonal (a diagonal's match posteriors over cells and paths sum to at most 1);
       windows that overflow run again with their exact count */
    std::vector<WinDev> wins;
    std::vector<int32_t> task_first_win(nt);
    int64_t out_total = 0;
    for (int64_t ti = 0; ti < nt; ti++) {
        const TaskDev &t = b->tasks[ti];
        task_first_win[ti] = (int32_t)wins.size();
        for (int32_t w = 0; w < nwin[ti]; w++) {
            WinDev W = slots[t.win_off + w];
            W.task = (int32_t)ti;
            W.out_off = out_total;
            W.out_cap = 3 * (W.TF - W.T) + 64;
            out_total += W.out_cap;
            wins.push_back(W);
        }
    }
    const int64_t nw = (int64_t)wins.size();
    /* two groups: the tasks at least half as long as the longest (their forward sweeps
       are the critical path) on the context's stream, the rest on a second stream,
       whose backward windows then run while the long forward sweeps continue */
    int64_t n_max = 0;
    for (int64_t ti = 0; ti < nt; ti++) n_max = std::max<int64_t>(n_max, (int64_t)b->tasks[ti].lX + b->tasks[ti].lY);
    std::vector<int32_t> task_order, win_order;
    for (int g = 0; g < 2; g++)
        for (int64_t ti = 0; ti < nt; ti++) {
            const bool longt = 2 * ((int64_t)b->tasks[ti].lX + b->tasks[ti].lY) >= n_max;
            if (longt != (g == 0)) continue;
            task_order.push_back((int32_t)ti);
        }
    int64_t n_long = 0;
    for (int64_t ti = 0; ti < nt; ti++) n_long += 2 * ((int64_t)b->tasks[ti].lX + b->tasks[ti].lY) >= n_max;
    int64_t nw_long = 0;
    for (int64_t i = 0; i < nt; i++) {
        const int32_t ti = task_order[i];
        for (int32_t w = 0; w < nwin[ti]; w++) win_order.push_back(task_first_win[ti] + w);
        if (i + 1 == n_long) nw_long = (int64_t)win_order.size();
    }
    int rc;
    if ((rc = upload(ctx, b->d_pcol, col_off)) || (rc = upload(ctx, b->d_pent, ent_off)) ||
        (rc = upload(ctx, b->d_pplane, plane_off)) || (rc = upload(ctx, b->d_pp, pp)) ||
        (rc = upload(ctx, b->d_psa, psa)) || (rc = upload(ctx, b->d_psb, psb)) ||
        (rc = upload(ctx, b->d_pecol, ecol)) || (rc = upload(ctx, b->d_pekid, ekid)) ||
        (rc = upload(ctx, b->d_pdoff, doffE)) || (rc = upload_raw(ctx, b->d_pdiag, pdiag, b->total_diag)) ||
        (rc = upload_raw(ctx, b->d_pentr, pent, nent)) || (rc = upload(ctx, b->d_pmaxent, maxent)) ||
        (rc = upload(ctx, b->d_wins, wins)) ||
        (rc = upload(ctx, b->d_win_order, win_order)) || (rc = upload(ctx, b->d_task_order, task_order)))
        return rc;
    SA_HIP(ctx, b->d_fM.ensure(sizeof(double) * std::max<int64_t>(planes, 1)));
    SA_HIP(ctx, b->d_fX.ensure(sizeof(double) * std::max<int64_t>(planes, 1)));
    SA_HIP(ctx, b->d_fY.ensure(sizeof(double) * std::max<int64_t>(planes, 1)));
    SA_HIP(ctx, b->d_out.ensure(sizeof(int32_t) * 4 * std::max<int64_t>(out_total, 1)));
    SA_HIP(ctx, b->d_win_count.ensure(sizeof(int32_t) * std::max<int64_t>(nw, 1)));
    SA_HIP(ctx, b->d_win_total.ensure(sizeof(double) * std::max<int64_t>(nw, 1)));
    sa::PathTab P;
    P.diag = b->d_pdiag.as<sa::PathDiag>();
    P.ent = b->d_pentr.as<sa::PathEnt>();
    P.maxent = b->d_pmaxent.as<int32_t>();
    P.col_off = b->d_pcol.as<int64_t>();
    P.ent_off = b->d_pent.as<int64_t>();
    P.plane_off = b->d_pplane.as<int64_t>();
    P.pp = b->d_pp.as<int32_t>();
    P.sa = b->d_psa.as<int32_t>();
    P.sb = b->d_psb.as<int32_t>();
    P.ecol = b->d_pecol.as<int32_t>();
    P.ekid = b->d_pekid.as<int32_t>();
    P.doffE = b->d_pdoff.as<int64_t>();
    P.nopt = B;
    P.ring = ring;
    a.wins = b->d_wins.as<WinDev>();
    a.win_order = b->d_win_order.as<int32_t>();
    a.fM = b->d_fM.as<double>(); a.fX = b->d_fX.as<double>(); a.fY = b->d_fY.as<double>();
    a.out = b->d_out.as<int32_t>();
    a.win_count = b->d_win_count.as<int32_t>();
    a.win_total = b->d_win_total.as<double>();
    /* rings in LDS up to 64 KiB per workgroup, else per-workgroup slices of HBM */
    constexpr size_t kLdsMax = 65536;
    const size_t fb = sa::paths_forward_bytes(ring), bb = sa::paths_backward_bytes(ring);
    const size_t fs = (fb + 255) / 256 * 256, bs = (bb + 255) / 256 * 256;
    /* HBM rings: one slice per forward workgroup, then one per backward workgroup (the
       two streams run forward and backward sweeps at the same time) */
    const size_t f_region = fb > kLdsMax ? (size_t)nt * fs : 0, b_region = bb > kLdsMax ? (size_t)nw * bs : 0;
    if (f_region + b_region) SA_HIP(ctx, b->d_ring_g.ensure(f_region + b_region));
    unsigned char *ring_f = fb > kLdsMax ? b->d_ring_g.as<unsigned char>() : nullptr;
    unsigned char *ring_b = bb > kLdsMax ? b->d_ring_g.as<unsigned char>() + f_region : nullptr;
    const int32_t *d_to = b->d_task_order.as<int32_t>(), *d_wo = b->d_win_order.as<int32_t>();
    hipStream_t st2 = ctx->aux[0]; /* the context's stream itself in a one-stream context */
    SA_HIP(ctx, hipEventRecord(ctx->ev[2], st));
    if (st2 != st) SA_HIP(ctx, hipStreamWaitEvent(st2, ctx->ev[2], 0));
    auto sweeps = [&](hipStream_t sg, int64_t t0, int64_t t1, int64_t w0, int64_t w1, bool mark) -> int {
        if (t1 > t0) {
            sa::Args<double> af = a;
            af.task_order = d_to + t0;
            af.ring_g = ring_f ? ring_f + (size_t)t0 * fs : nullptr;
            af.ring_g_stride = (int64_t)fs;
            SA_HIP(ctx, sa::launch_paths_forward(af, P, (int)(t1 - t0), sg));
        }
        if (mark) SA_HIP(ctx, hipEventRecord(ctx->ev[3], sg));
        if (w1 > w0) {
            sa::Args<double> ab = a;
            ab.win_order = d_wo + w0;
            ab.ring_g = ring_b ? ring_b + (size_t)w0 * bs : nullptr;
            ab.ring_g_stride = (int64_t)bs;
            SA_HIP(ctx, sa::launch_paths_backward(ab, P, (int)(w1 - w0), sg));
        }
        return SA_OK;
    };
    rc = sweeps(st, 0, n_long, 0, nw_long, true);
    if (!rc) rc = sweeps(st2, n_long, nt, nw_long, nw, false);
    if (st2 != st) {
        /* join st2 on every exit: buffers are freed stream-ordered on st, and a free
           must not overtake the sweeps still queued on st2 */
        const hipError_t e1 = hipEventRecord(ctx->ev[9], st2);
        const hipError_t e2 = e1 == hipSuccess ? hipStreamWaitEvent(st, ctx->ev[9], 0) : e1;
        if (e2 != hipSuccess) {
            (void)hipStreamSynchronize(st2);
            if (!rc) rc = sa_rt::fail_hip(ctx, e2, "joining the second stream");
        }
    }
    if (rc) return rc;
    SA_HIP(ctx, hipEventRecord(ctx->ev[4], st));
    std::vector<int32_t> wcount(nw);
    std::vector<double> wtot(nw);
    std::vector<int32_t> outv(4 * std::max<int64_t>(out_total, 1));
    SA_HIP(ctx, hipMemcpyAsync(wcount.data(), b->d_win_count.p, sizeof(int32_t) * nw, hipMemcpyDeviceToHost, st));
    SA_HIP(ctx, hipMemcpyAsync(wtot.data(), b->d_win_total.p, sizeof(double) * nw, hipMemcpyDeviceToHost, st));
    if (out_total)
        SA_HIP(ctx, hipMemcpyAsync(outv.data(), b->d_out.p, sizeof(int32_t) * 4 * out_total, hipMemcpyDeviceToHost,
                                   st));
    SA_HIP(ctx, hipStreamSynchronize(st));
    /* overflowing windows: the backward sweep again for them alone (the forward
       planes are kept), into slots of their exact size after the first pass's */
    std::vector<int32_t> redo;
    for (int64_t w = 0; w < nw; w++)
        if (wcount[w] > wins[w].out_cap) redo.push_back((int32_t)w);
    if (!redo.empty()) {
        int64_t extra = 0;
        for (int32_t w : redo) {
            wins[w].out_off = extra;
            wins[w].out_cap = wcount[w];
            extra += wcount[w];
        }
        SA_HIP(ctx, b->d_packed.ensure(sizeof(int32_t) * 4 * extra));
        if ((rc = upload(ctx, b->d_wins, wins)) || (rc = upload(ctx, b->d_win_order, redo))) return rc;
        sa::Args<double> ab = a;
        ab.wins = b->d_wins.as<WinDev>();
        ab.win_order = b->d_win_order.as<int32_t>();
        ab.out = b->d_packed.as<int32_t>();
        ab.ring_g = ring_b;
        ab.ring_g_stride = (int64_t)bs;
        SA_HIP(ctx, sa::launch_paths_backward(ab, P, (int)redo.size(), st));
        /* the timed span ends after the second pass (the host round trip included) */
        SA_HIP(ctx, hipEventRecord(ctx->ev[4], st));
        std::vector<int32_t> extra_v(4 * extra);
        SA_HIP(ctx, hipMemcpyAsync(extra_v.data(), b->d_packed.p, sizeof(int32_t) * 4 * extra, hipMemcpyDeviceToHost,
                                   st));
        SA_HIP(ctx, hipStreamSynchronize(st));
        /* the redone windows' pairs go to the end of the host copy */
        for (int32_t w : redo) wins[w].out_off += out_total;
        outv.resize(4 * (out_total + extra));
        std::memcpy(outv.data() + 4 * out_total, extra_v.data(), sizeof(int32_t) * 4 * extra);
        tm.backward_launches++;
    }
    /* the reference's list order: per read, sub-alignments in split order, each
       sub-alignment's pairs in reverse emission order, shifted by (x1, y1) */
    for (int64_t r = 0; r < b->n_reads; r++) {
        ReadOut &ro = res->reads[r];
        ro.first_pair = (int64_t)res->pairs.size();
        ro.n_sub = b->reads[r].n_tasks;
        for (int64_t ti = b->reads[r].first_task; ti < b->reads[r].first_task + b->reads[r].n_tasks; ti++) {
            std::vector<int64_t> idx;
            for (int32_t w = 0; w < nwin[ti]; w++) {
                const WinDev &W = wins[task_first_win[ti] + w];
                const int64_t n = std::min<int64_t>(wcount[task_first_win[ti] + w], W.out_cap);
                for (int64_t q = 0; q < n; q++) idx.push_back(W.out_off + q);
            }
            for (auto it = idx.rbegin(); it != idx.rend(); ++it) {
                const int32_t *o = outv.data() + 4 * *it;
                sa_pair pr;
                pr.p = o[0];
                pr.x = o[1] + b->thost[ti].x1;
                pr.y = o[2] + b->thost[ti].y1;
                res->pairs.push_back(pr);
                res->paths.push_back(o[3]);
            }
            if (nwin[ti] > 0) ro.log_likelihood += wtot[task_first_win[ti] + nwin[ti] - 1];
            ro.cells += cells[ti];
            ro.n_windows += nwin[ti];
        }
        ro.n_pairs = (int64_t)res->pairs.size() - ro.first_pair;
    }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]); tm.prep_ms = ms;
    (void)hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]); tm.forward_ms = ms;
    (void)hipEventElapsedTime(&ms, ctx->ev[3], ctx->ev[4]); tm.backward_ms = ms;
    (void)hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[4]); tm.fb_ms = ms;
    (void)hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[4]); tm.total_ms = ms;
    tm.device_start_ms = sa_rt::timeline_ms(ctx->device, ctx->ev[0]);
    tm.device_end_ms = sa_rt::timeline_ms(ctx->device, ctx->ev[4]);
    int64_t tc = 0;
    for (int64_t c : cells) tc += c;
    tm.cells = tc;
    tm.forward_launches = (n_long > 0) + (nt > n_long);
    tm.backward_launches += (nw_long > 0) + (nw > nw_long);
    return SA_OK;
}

/* A run from the recorded graph: same launches, same buffers, same results. */
template <typename T>
int replay_batch(sa_batch *b, sa_result *res) {
    sa_context *ctx = b->ctx;
    hipStream_t st = ctx->stream;
    auto &gc = b->gc[sizeof(T) == 4 ? 1 : 0];
    sa_timing &tm = res->timing;
    std::memset(&tm, 0, sizeof(tm));
    SA_HIP(ctx, hipEventRecord(ctx->ev[12], st));
    SA_HIP(ctx, hipGraphLaunch(gc.exec, st));
    SA_HIP(ctx, hipEventRecord(ctx->ev[13], st));
    res->reads = gc.reads;
    std::vector<double> wtot(gc.nw);
    int32_t err = 0;
    if (int rc = fetch_pairs(ctx, res, b->d_packed.p, gc.total_pairs, st)) return rc;
    SA_HIP(ctx, hipMemcpyAsync(wtot.data(), b->d_win_total.p, sizeof(double) * gc.nw, hipMemcpyDeviceToHost, st));
    SA_HIP(ctx, hipMemcpyAsync(&err, b->d_err.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    SA_HIP(ctx, hipStreamSynchronize(st));
    finish_pairs(ctx, res, gc.total_pairs);
    if (err) return fail(ctx, SA_ERR_HIP, "replayed run raised device error flags " + std::to_string(err));
    for (int64_t r = 0; r < b->n_reads; r++) {
        double ll = 0.0;
        for (int64_t ti = b->reads[r].first_task; ti < b->reads[r].first_task + b->reads[r].n_tasks; ti++)
            if (gc.last_win[ti] >= 0) ll += wtot[gc.last_win[ti]];
        res->reads[r].log_likelihood = ll;
    }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ctx->ev[12], ctx->ev[13]); tm.total_ms = ms;
    tm.device_start_ms = sa_rt::timeline_ms(ctx->device, ctx->ev[12]);
    tm.device_end_ms = sa_rt::timeline_ms(ctx->device, ctx->ev[13]);
    /* per-step times from the events the graph recorded (left 0 if unavailable) */
    if (hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]) == hipSuccess) tm.prep_ms = ms;
    if (hipEventElapsedTime(&ms, ctx->ev[5], ctx->ev[6]) == hipSuccess) tm.collect_ms = ms;
    float fb = 0.f;
    const bool have_fb = hipEventElapsedTime(&fb, ctx->ev[2], ctx->ev[4]) == hipSuccess;
    if (have_fb) tm.fb_ms = fb;
    double fwd_end = 0.0, bwd_start = fb;
    res->launches = gc.launches;
    size_t nf = 0;
    for (const sa_launch &L : gc.launches) nf += L.direction == 0;
    size_t kf = 0, kb = 0;
    for (sa_launch &L : res->launches) {
        const size_t e = L.direction == 0 ? 2 * kf++ : 2 * (nf + kb++);
        float t0 = 0.f, t1 = 0.f;
        if (have_fb && hipEventElapsedTime(&t0, ctx->ev[2], ctx->gev[e]) == hipSuccess &&
            hipEventElapsedTime(&t1, ctx->ev[2], ctx->gev[e + 1]) == hipSuccess) {
            L.start_ms = t0;
            L.ms = t1 - t0;
            if (L.direction == 0) fwd_end = std::max(fwd_end, (double)t1);
            else bwd_start = std::min(bwd_start, (double)t0);
        } else {
            L.start_ms = 0.0;
            L.ms = 0.0;
        }
    }
    (void)hipGetLastError();
    tm.forward_ms = fwd_end;
    tm.backward_ms = have_fb ? fb - bwd_start : 0.0;
    tm.forward_launches = (int64_t)nf;
    tm.backward_launches = (int64_t)(gc.launches.size() - nf);
    tm.cells = gc.cells;
    tm.emission_path = gc.ek;
    if (gc.ek == 3) tm.emission_ms = plane_ms(ctx, gc.em_c);
    tm.replayed = 1;
    return SA_OK;
}

} // namespace

extern "C" {

int sa_batch_run(sa_batch *b, sa_precision precision, sa_result **out) {
    *out = nullptr;
    if (!b) return SA_ERR_INVALID_ARGUMENT;
    sa_context *ctx = b->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    sa_rt::enter(ctx);
    AllocStream as(ctx->stream);
    sa_result *res = new sa_result();
    const char *kern = getenv("SA_KERNELS");
    const bool lds = kern && std::strcmp(kern, "lds") == 0;
    int rc;
    if (b->paths) {
        /* degenerate references run in FP64 */
        rc = precision == SA_FP64 ? run_paths(b, res)
                                  : fail(ctx, SA_ERR_UNSUPPORTED, "degenerate references run in FP64");
        if (rc) {
            delete res;
            return rc;
        }
        b->nruns[0]++;
        *out = res;
        return SA_OK;
    }
    if (!lds && b->tune.use_graph && b->gc[precision == SA_FP32 ? 1 : 0].exec)
        rc = precision == SA_FP32 ? replay_batch<float>(b, res) : replay_batch<double>(b, res);
    else
        rc = precision == SA_FP32 ? run_batch<float>(b, res, lds) : run_batch<double>(b, res, lds);
    if (rc == kRetryLds) { /* a band escaped a register window: redo with the LDS sweeps */
        delete res;
        res = new sa_result();
        rc = precision == SA_FP32 ? run_batch<float>(b, res, true) : run_batch<double>(b, res, true);
    }
    if (rc) {
        delete res;
        return rc;
    }
    b->nruns[precision == SA_FP32 ? 1 : 0]++;
    *out = res;
    return SA_OK;
}

int sa_align_batch(sa_context *ctx, const sa_model *model, const sa_read *reads, int64_t n_reads,
                   const sa_params *params, sa_precision precision, sa_result **out) {
    *out = nullptr;
    if (!ctx) return SA_ERR_INVALID_ARGUMENT;
    /* one batch per context, refilled call after call: its device buffers only grow, so
       repeated calls do not allocate again (a large batch's planes are tens of GB, and
       allocating them from the stream-ordered pool can take seconds per call) */
    int rc;
    if (ctx->batch) {
        rc = sa_batch_refill(ctx->batch, model, reads, n_reads, params);
    } else {
        sa_batch *b = nullptr;
        rc = sa_batch_create(ctx, model, reads, n_reads, params, &b);
        if (rc == SA_OK) ctx->batch = b;
    }
    if (rc == SA_OK) rc = sa_batch_run(ctx->batch, precision, out);
    if (rc != SA_OK && ctx->batch) { /* start the next call from a fresh batch */
        sa_batch_destroy(ctx->batch);
        ctx->batch = nullptr;
    }
    return rc;
}

int sa_align(sa_context *ctx, const sa_model *model, const sa_read *read, const sa_params *params,
             sa_precision precision, sa_result **out) {
    return sa_align_batch(ctx, model, read, 1, params, precision, out);
}

int sa_expect_batch(sa_context *ctx, const sa_model *model, const sa_read *reads, int64_t n_reads,
                    const sa_params *params, double record_threshold, sa_expect_result **out) {
    if (!out) return SA_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (!(record_threshold >= 0.0) || n_reads < 0) return SA_ERR_INVALID_ARGUMENT;
    sa_batch *b = nullptr;
    int rc = sa_batch_create(ctx, model, reads, n_reads, params, &b);
    if (rc) return rc;
    b->expect = 1;
    b->ex_rec_threshold = record_threshold;
    sa_result *res = nullptr;
    rc = sa_batch_run(b, SA_FP64, &res);
    sa_batch_destroy(b);
    if (rc) return rc;
    sa_expect_result *er = new sa_expect_result();
    er->reads = std::move(res->expect);
    er->timing = res->timing;
    er->reads.resize(n_reads);
    for (ReadExpect &e : er->reads)
        if (e.s0.empty()) { /* a read with no cells */
            e.s0.assign(model->n_kmers, 0.0);
            e.s1.assign(model->n_kmers, 0.0);
            e.obs.assign(model->n_kmers, 0);
        }
    delete res;
    *out = er;
    return SA_OK;
}

int sa_expect(sa_context *ctx, const sa_model *model, const sa_read *read, const sa_params *params,
              double record_threshold, sa_expect_result **out) {
    return sa_expect_batch(ctx, model, read, 1, params, record_threshold, out);
}

int sa_expect_read(const sa_expect_result *res, int64_t i, sa_expectations *out) {
    if (!res || !out || i < 0 || i >= (int64_t)res->reads.size()) return SA_ERR_INVALID_ARGUMENT;
    const ReadExpect &e = res->reads[i];
    for (int k = 0; k < 9; k++) out->transitions[k] = e.trans[k];
    out->likelihood = e.likelihood;
    out->n_kmers = (int64_t)e.s0.size();
    out->match_posterior = e.s0.data();
    out->match_event_sum = e.s1.data();
    out->observed = e.obs.data();
    out->records = e.recs.data();
    out->n_records = (int64_t)e.recs.size();
    return SA_OK;
}

int sa_expect_get(const sa_expect_result *res, sa_expectations *out) { return sa_expect_read(res, 0, out); }

int sa_expect_timing(const sa_expect_result *res, sa_timing *out) {
    if (!res || !out) return SA_ERR_INVALID_ARGUMENT;
    *out = res->timing;
    return SA_OK;
}

void sa_expect_free(sa_expect_result *res) { delete res; }

int sa_band_geometry(const int64_t *anchors, int64_t n_anchors, int64_t lX, int64_t lY, int64_t expansion,
                     int32_t *xlo, int32_t *xhi) {
    if (lX < 0 || lY < 0 || expansion < 0 || expansion % 2 != 0 || n_anchors < 0) return SA_ERR_INVALID_ARGUMENT;
    std::vector<int32_t> ax(n_anchors), ay(n_anchors);
    for (int64_t i = 0; i < n_anchors; i++) {
        const int64_t x = anchors[2 * i] + 1, y = anchors[2 * i + 1] + 1;
        if (x <= 0 || y <= 0 || x > lX || y > lY || (i > 0 && (x <= ax[i - 1] || y <= ay[i - 1])))
            return SA_ERR_ANCHORS;
        ax[i] = (int32_t)x;
        ay[i] = (int32_t)y;
    }
    for (int64_t d = 0; d <= lX + lY; d++) {
        int64_t L, R;
        if (!sa::band_at(ax.data(), ay.data(), (int32_t)n_anchors, lX, lY, expansion, d, &L, &R)) return SA_ERR_ANCHORS;
        xlo[d] = (int32_t)sa::div2(d + L);
        xhi[d] = (int32_t)sa::div2(d + R);
    }
    return SA_OK;
}

int64_t sa_window_schedule(const int32_t *xlo, const int32_t *xhi, int64_t lX, int64_t lY, const sa_params *p,
                           int32_t *out, int64_t cap) {
    const int64_t N = lX + lY;
    std::vector<WinDev> w(sa::window_cap(N, p->min_diags_between_traceback, p->traceback_diagonals));
    auto width = [&](int64_t d) { return (int64_t)(xhi[d] - xlo[d] + 1); };
    int32_t n = sa::window_schedule(N, p->min_diags_between_traceback, p->traceback_diagonals, p->diagonal_expansion,
                                    width, w.data(), (int32_t)w.size(), 0);
    for (int32_t i = 0; i < n && i < cap; i++) {
        out[4 * i + 0] = w[i].T;
        out[4 * i + 1] = w[i].D;
        out[4 * i + 2] = w[i].TF;
        out[4 * i + 3] = w[i].last;
    }
    return n;
}

int64_t sa_split_points(const int64_t *anchors, int64_t n_anchors, int64_t lX, int64_t lY, const sa_params *p,
                        int32_t ragged_left, int32_t ragged_right, int64_t *out, int64_t cap) {
    std::vector<int64_t> sp;
    if (!split_points(anchors, n_anchors, lX, lY, p->split_matrix_bigger_than_this, ragged_left, ragged_right, sp))
        return SA_ERR_ANCHORS;
    const int64_t n = (int64_t)sp.size() / 4;
    for (int64_t i = 0; i < 4 * std::min(n, cap); i++) out[i] = sp[i];
    return n;
}

int64_t sa_result_n_reads(const sa_result *res) { return res ? (int64_t)res->reads.size() : 0; }

int sa_result_read(const sa_result *res, int64_t i, sa_read_result *out) {
    if (!res || i < 0 || i >= (int64_t)res->reads.size()) return SA_ERR_INVALID_ARGUMENT;
    const ReadOut &r = res->reads[i];
    out->pairs = r.n_pairs ? res->pair_data() + r.first_pair : nullptr;
    out->n_pairs = r.n_pairs;
    out->log_likelihood = r.log_likelihood;
    out->cells = r.cells;
    out->n_windows = r.n_windows;
    out->n_subalignments = r.n_sub;
    out->paths = res->paths.empty() || !r.n_pairs ? nullptr : res->paths.data() + r.first_pair;
    return SA_OK;
}

int64_t sa_result_launches(const sa_result *res, sa_launch *out, int64_t cap) {
    if (!res) return SA_ERR_INVALID_ARGUMENT;
    const int64_t n = (int64_t)res->launches.size();
    for (int64_t i = 0; i < n && i < cap; i++) out[i] = res->launches[i];
    return n;
}

int sa_result_timing(const sa_result *res, sa_timing *out) {
    if (!res) return SA_ERR_INVALID_ARGUMENT;
    *out = res->timing;
    return SA_OK;
}

void sa_result_free(sa_result *res) { delete res; }

} // extern "C"
