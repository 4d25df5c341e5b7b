/*
 * sa_noalloc.cpp -- sa_align_noalloc: the plain per-read call (sa_align) for callers
 * that cannot allocate memory (code that must not call hipMalloc, tight serving
 * loops, harnesses that forbid allocation).
 *
 * Every buffer the call needs beyond the caller's inputs and output -- the staged
 * inputs, the per-position, per-event and per-diagonal constants, the forward
 * planes, the traceback windows, the pair slots and the compacted pairs -- is
 * carved out of one file-scope device array, `scratch`, of kScratchBytes per device.
 * Its device address is resolved once per device (scratch_setup, at context
 * creation) with hipGetSymbolAddress; host code only ever hands that address to
 * kernels, copies and memsets.  The host side plans in fixed per-device arrays, so
 * the call allocates nothing on the device or on the host.
 *
 * The array is shared by every call on its device and keeps the last call's
 * contents: each call clears what must start at zero (the k-mer and event
 * constants with their padding, the error flag).  Calls on one device are
 * serialised by a per-device lock, since they share the array; they never run
 * concurrently on one GPU.
 *
 * The kernels are the library's own (sa_kernels.hip, sa_fb_reg.hip), launched on
 * the context's stream: band prep, the forward sweep of each sub-alignment (several
 * wavefronts per sweep, or the LDS sweep for bands no register window holds), the
 * backward sweep of each traceback window, and the pair compaction.  Results equal
 * sa_align's for the same read.
 */
#include "sa_internal.h"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>

using sa::TaskDev;
using sa::WinDev;
using sa_rt::fail;

namespace {

constexpr size_t kScratchBytes = (size_t)SA_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 16;
constexpr int kMaxTasks = 256;        /* split blocks of one read */
constexpr int kMaxAnchors = 1 << 16;  /* anchors of one read */
constexpr int kMaxWins = 8192;        /* traceback window slots of one read */

/* The scratch array, one copy per device (its host name is only a handle for
   hipGetSymbolAddress, never a device address). */
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

/* Host-side planning state of one device, in fixed storage. */
struct HostPlan {
    std::mutex mu; /* one call at a time per device: the calls share `scratch` */
    unsigned char *dev = nullptr; /* device address of `scratch` on this device */
    TaskDev tasks[kMaxTasks];
    int32_t x1[kMaxTasks], y1[kMaxTasks];
    int64_t cells[kMaxTasks], xycells[kMaxTasks];
    int32_t maxw[kMaxTasks], span[kMaxTasks], hspan[kMaxTasks], nwin[kMaxTasks], first_win[kMaxTasks];
    int32_t fgrp[kMaxTasks];
    int32_t ancx[kMaxAnchors], ancy[kMaxAnchors];
    WinDev slots[kMaxWins];
    WinDev wins[kMaxWins];
    int32_t wgrp[kMaxWins];
    int32_t task_order[kMaxTasks], win_order[kMaxWins];
    int32_t wcount[kMaxWins];
    int64_t win_dst[kMaxWins];
    double wtot[kMaxWins];
};
HostPlan g_plan[kMaxDevices];
std::mutex g_setup_mu;

/* Bump allocation inside the scratch array (256-byte aligned slices). */
struct Arena {
    unsigned char *base;
    size_t used = 0;
    bool over = false;
    explicit Arena(unsigned char *b) : base(b) {}
    template <typename U> U *take(int64_t n) {
        const size_t bytes = ((size_t)(n > 0 ? n : 1) * sizeof(U) + 255) & ~(size_t)255;
        if (bytes > kScratchBytes - used) { over = true; return nullptr; }
        U *p = reinterpret_cast<U *>(base + used);
        used += bytes;
        return p;
    }
};

int next_pow2(int v) {
    int r = 1;
    while (r < v) r <<= 1;
    return r;
}

constexpr int kCpt[] = {1, 2, 3, 4, 6, 8};
constexpr int kMW[4][2] = {{2, 1}, {3, 1}, {4, 1}, {4, 2}}; /* several-wavefront shapes (waves, cpt) */

int cpt_class(int sp) { /* index into kCpt of the narrowest one-wavefront window, or -1 */
    for (int k = 0; k < 6; k++)
        if (64 * kCpt[k] >= sp + kCpt[k] - 1) return k;
    return -1;
}
int mw_class(int sp) {
    for (int k = 0; k < 4; k++)
        if (64 * kMW[k][0] * kMW[k][1] >= sp + kMW[k][1] - 1) return k;
    return -1;
}

int capacity_error(sa_context *ctx, const char *what, int64_t need, int64_t cap) {
    std::snprintf(ctx->err, sizeof(ctx->err), "sa_align_noalloc: %s: %lld needed, at most %lld", what,
                  (long long)need, (long long)cap);
    return SA_ERR_CAPACITY;
}

template <typename T>
int align_noalloc(sa_context *ctx, HostPlan &H, const sa_model *model, const sa_read &rd, const sa_params &p,
                  sa_pair *pairs, int64_t pairs_cap, int64_t *n_pairs, double *log_likelihood) {
    hipStream_t st = ctx->stream;
    const int k = model->k;
    const int64_t lX = sa_rt::correct_length(rd.ref_length, k), lY = rd.n_events;
    const int64_t stride = rd.event_stride;

    /* split blocks and their anchors (matrix coordinates) -- as sa_batch_create */
    int nt = 0, n_split = 0;
    int64_t total_pos = 0, total_ev = 0, total_diag = 0, total_wslots = 0;
    int64_t na_all = 0, j = 0;
    bool anchors_ok = true, too_many = false;
    int64_t bounds[kMaxTasks][4];
    const bool split_ok = sa_rt::split_points_each(
        rd.anchors, rd.n_anchors, lX, lY, p.split_matrix_bigger_than_this, rd.ragged_left != 0, rd.ragged_right != 0,
        [&](int64_t a, int64_t b, int64_t c, int64_t d) {
            if (n_split >= kMaxTasks) { too_many = true; return; }
            bounds[n_split][0] = a; bounds[n_split][1] = b; bounds[n_split][2] = c; bounds[n_split][3] = d;
            n_split++;
        });
    if (!split_ok) return fail(ctx, SA_ERR_ANCHORS, "anchors out of order or range");
    if (too_many) return capacity_error(ctx, "split blocks", kMaxTasks + 1, kMaxTasks);
    for (int i = 0; i < n_split; i++) {
        const int64_t x1 = bounds[i][0], y1 = bounds[i][1], x2 = bounds[i][2], y2 = bounds[i][3];
        if (x2 - x1 > INT32_MAX / 4 || y2 - y1 > INT32_MAX / 4) return fail(ctx, SA_ERR_UNSUPPORTED, "sub-alignment too long");
        TaskDev t;
        std::memset(&t, 0, sizeof(t));
        t.lX = (int32_t)(x2 - x1);
        t.lY = (int32_t)(y2 - y1);
        t.ref_off = x1;
        t.ev_off = y1;
        t.pos_off = total_pos;
        t.evc_off = total_ev;
        t.diag_off = total_diag;
        t.anc_off = na_all;
        int32_t na = 0;
        int64_t px = -1, py = -1;
        while (j < rd.n_anchors) {
            const int64_t x = rd.anchors[2 * j], y = rd.anchors[2 * j + 1];
            if (x + y >= x2 + y2) break;
            if (!(x >= x1 && x < x2 && y >= y1 && y < y2) || x - x1 <= px || y - y1 <= py) { anchors_ok = false; break; }
            if (na_all >= kMaxAnchors) return capacity_error(ctx, "anchors", rd.n_anchors, kMaxAnchors);
            px = x - x1;
            py = y - y1;
            H.ancx[na_all] = (int32_t)(x - x1 + 1);
            H.ancy[na_all] = (int32_t)(y - y1 + 1);
            na_all++;
            na++;
            j++;
        }
        if (!anchors_ok) return fail(ctx, SA_ERR_ANCHORS, "anchors out of order");
        t.n_anc = na;
        t.ragged_left = rd.ragged_left || i > 0;
        t.ragged_right = rd.ragged_right || i < n_split - 1;
        t.read = 0;
        const int64_t N = (int64_t)t.lX + t.lY;
        t.win_off = (int32_t)total_wslots;
        t.win_cap = sa::window_cap(N, p.min_diags_between_traceback, p.traceback_diagonals);
        t.scale = rd.scale;
        t.shift = rd.shift;
        t.var = rd.var;
        total_pos += t.lX;
        total_ev += t.lY;
        total_diag += N + 1;
        if (N > 0) {
            total_wslots += t.win_cap;
            if (total_wslots > kMaxWins) return capacity_error(ctx, "traceback window slots", total_wslots, kMaxWins);
            H.tasks[nt] = t;
            H.x1[nt] = (int32_t)x1;
            H.y1[nt] = (int32_t)y1;
            nt++;
        }
    }
    if (j != rd.n_anchors) return fail(ctx, SA_ERR_ANCHORS, "anchors not consumed");
    *n_pairs = 0;
    *log_likelihood = 0.0;
    if (nt == 0) return SA_OK;

    /* scratch, first part: inputs, constants, band, prep outputs */
    Arena ar(H.dev);
    char *d_ref = ar.take<char>(rd.ref_length);
    double *d_events = ar.take<double>(lY * stride);
    int32_t *d_ancx = ar.take<int32_t>(na_all), *d_ancy = ar.take<int32_t>(na_all);
    TaskDev *d_tasks = ar.take<TaskDev>(nt);
    int32_t *d_x1 = ar.take<int32_t>(nt), *d_y1 = ar.take<int32_t>(nt);
    int32_t *d_kid = ar.take<int32_t>(total_pos + 64);
    T *d_evm = ar.take<T>(total_ev + 64), *d_evn = ar.take<T>(total_ev + 64);
    T *d_evl = ar.take<T>(total_ev + 64), *d_evi = ar.take<T>(total_ev + 64);
    int32_t *d_xlo = ar.take<int32_t>(total_diag), *d_xhi = ar.take<int32_t>(total_diag);
    int32_t *d_doff = ar.take<int32_t>(total_diag), *d_doffxy = ar.take<int32_t>(total_diag);
    double *d_foff = ar.take<double>(total_diag);
    int8_t *d_dflag = ar.take<int8_t>(total_diag + 4), *d_dcls = ar.take<int8_t>(total_diag + 4);
    int64_t *d_cells = ar.take<int64_t>(nt), *d_xycells = ar.take<int64_t>(nt);
    int32_t *d_maxw = ar.take<int32_t>(nt), *d_span = ar.take<int32_t>(nt), *d_nwin = ar.take<int32_t>(nt);
    int32_t *d_hspan = ar.take<int32_t>(nt);
    WinDev *d_slots = ar.take<WinDev>(total_wslots);
    int32_t *d_err = ar.take<int32_t>(1);
    if (ar.over) return capacity_error(ctx, "scratch bytes for the inputs and band", (int64_t)kScratchBytes + 1,
                                       (int64_t)kScratchBytes);

    if (rd.ref_length > 0) SA_HIP(ctx, hipMemcpyAsync(d_ref, rd.ref, (size_t)rd.ref_length, hipMemcpyHostToDevice, st));
    if (lY > 0)
        SA_HIP(ctx, hipMemcpyAsync(d_events, rd.events, sizeof(double) * (size_t)(lY * stride), hipMemcpyHostToDevice, st));
    if (na_all > 0) {
        SA_HIP(ctx, hipMemcpyAsync(d_ancx, H.ancx, sizeof(int32_t) * na_all, hipMemcpyHostToDevice, st));
        SA_HIP(ctx, hipMemcpyAsync(d_ancy, H.ancy, sizeof(int32_t) * na_all, hipMemcpyHostToDevice, st));
    }
    SA_HIP(ctx, hipMemcpyAsync(d_x1, H.x1, sizeof(int32_t) * nt, hipMemcpyHostToDevice, st));
    SA_HIP(ctx, hipMemcpyAsync(d_y1, H.y1, sizeof(int32_t) * nt, hipMemcpyHostToDevice, st));
    SA_HIP(ctx, hipMemsetAsync(d_kid, 0, sizeof(int32_t) * (total_pos + 64), st));
    SA_HIP(ctx, hipMemsetAsync(d_evm, 0, sizeof(T) * (total_ev + 64), st));
    SA_HIP(ctx, hipMemsetAsync(d_evn, 0, sizeof(T) * (total_ev + 64), st));
    SA_HIP(ctx, hipMemsetAsync(d_evl, 0, sizeof(T) * (total_ev + 64), st));
    SA_HIP(ctx, hipMemsetAsync(d_evi, 0, sizeof(T) * (total_ev + 64), st));
    SA_HIP(ctx, hipMemsetAsync(d_err, 0, sizeof(int32_t), st));

    sa::Args<T> a;
    std::memset(&a, 0, sizeof(a));
    a.m = sa_rt::device_model<T>(model, sa_rt::tables<T>(model));
    a.tasks = d_tasks;
    a.ref = d_ref;
    a.events = d_events;
    a.event_stride = stride;
    a.anc_x = d_ancx;
    a.anc_y = d_ancy;
    a.lut = model->d_lut.as<int8_t>();
    a.kid = d_kid;
    a.evm = d_evm; a.evn = d_evn; a.evl = d_evl; a.evi = d_evi;
    a.xlo = d_xlo; a.xhi = d_xhi; a.doff = d_doff;
    a.foff = d_foff;
    a.dflag = d_dflag;
    a.dcls = d_dcls;
    a.doffxy = d_doffxy;
    a.task_cells = d_cells;
    a.task_xycells = d_xycells;
    a.task_maxw = d_maxw;
    a.task_span = d_span;
    a.task_hspan = d_hspan;
    a.task_nwin = d_nwin;
    a.win_slots = d_slots;
    a.err = d_err;
    a.threshold = p.threshold;
    a.min_diags = p.min_diags_between_traceback;
    a.traceback = p.traceback_diagonals;
    a.expansion = p.diagonal_expansion;

    SA_HIP(ctx, hipMemcpyAsync(d_tasks, H.tasks, sizeof(TaskDev) * nt, hipMemcpyHostToDevice, st));
    SA_HIP(ctx, sa::launch_prep<T>(a, nt, st));
    int32_t err = 0;
    SA_HIP(ctx, hipMemcpyAsync(H.cells, d_cells, sizeof(int64_t) * nt, hipMemcpyDeviceToHost, st));
    SA_HIP(ctx, hipMemcpyAsync(H.xycells, d_xycells, sizeof(int64_t) * nt, hipMemcpyDeviceToHost, st));
    SA_HIP(ctx, hipMemcpyAsync(H.maxw, d_maxw, sizeof(int32_t) * nt, hipMemcpyDeviceToHost, st));
    SA_HIP(ctx, hipMemcpyAsync(H.span, d_span, sizeof(int32_t) * nt, hipMemcpyDeviceToHost, st));
    SA_HIP(ctx, hipMemcpyAsync(H.hspan, d_hspan, sizeof(int32_t) * nt, hipMemcpyDeviceToHost, st));
    SA_HIP(ctx, hipMemcpyAsync(H.nwin, d_nwin, sizeof(int32_t) * nt, hipMemcpyDeviceToHost, st));
    SA_HIP(ctx, hipMemcpyAsync(H.slots, d_slots, sizeof(WinDev) * total_wslots, hipMemcpyDeviceToHost, st));
    SA_HIP(ctx, hipMemcpyAsync(&err, d_err, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    SA_HIP(ctx, hipStreamSynchronize(st));
    if (err & sa::ERR_BAD_KMER) return fail(ctx, SA_ERR_BAD_KMER, "reference k-mer with a character outside the alphabet");
    if (err & sa::ERR_BAND) return fail(ctx, SA_ERR_ANCHORS, "invalid band (anchors outside the matrix or out of order)");
    if (err & sa::ERR_WINDOWS) return fail(ctx, SA_ERR_HIP, "window schedule overflow");

    /* plan: planes, windows and their pair slots (as sa_batch_run) */
    int64_t total_cells = 0, total_xy = 0, out_total = 0;
    int nw = 0;
    for (int i = 0; i < nt; i++) {
        H.tasks[i].cell_off = total_cells;
        H.tasks[i].xy_off = total_xy;
        total_cells += H.cells[i];
        total_xy += H.xycells[i];
        H.first_win[i] = nw;
        for (int32_t w = 0; w < H.nwin[i]; w++) {
            WinDev W = H.slots[H.tasks[i].win_off + w];
            W.task = i;
            W.out_off = out_total;
            W.out_cap = 2 * (W.TF - W.T) + 64;
            W.rec_off = 0;
            W.rec_cap = 0;
            out_total += W.out_cap;
            H.wins[nw++] = W;
        }
    }
    const size_t mark = ar.used;

    /* attempt 0: register sweeps where a register window holds the band; a band that
       escapes its window raises ERR_WINDOWS and the call runs again on the LDS sweeps */
    for (int attempt = 0; attempt < 2; attempt++) {
        const bool force_lds = attempt == 1;
        ar.used = mark;
        ar.over = false;
        int lds_maxw = 1;
        int n_lds_f = 0, n_lds_b = 0;
        for (int i = 0; i < nt; i++) {
            const int mw = force_lds ? -1 : mw_class(H.span[i]);
            H.fgrp[i] = mw >= 0 ? 1 + mw : 0; /* 0: LDS sweep */
            /* 5 + k: the halo sweep with k + 2 wavefronts, where its window holds the band */
            const char *halo_env = getenv("SA_HALO"); /* "0": no halo sweeps, as in sa_align */
            const bool use_halo = !(halo_env && std::strcmp(halo_env, "0") == 0);
            for (int hw = 2; use_halo && !force_lds && hw <= 4; hw++)
                if (sa::halo_window(hw) >= H.hspan[i]) { H.fgrp[i] = 5 + hw - 2; break; }
            if (H.fgrp[i] == 0) { n_lds_f++; lds_maxw = std::max(lds_maxw, H.maxw[i]); }
        }
        for (int i = 0; i < nt; i++)
            for (int32_t w = 0; w < H.nwin[i]; w++) {
                const int wi = H.first_win[i] + w;
                const int c = force_lds ? -1 : cpt_class(H.wins[wi].span);
                H.wgrp[wi] = c >= 0 ? 1 + c : 0;
                if (H.wgrp[wi] == 0) { n_lds_b++; lds_maxw = std::max(lds_maxw, H.maxw[i]); }
            }
        /* orders grouped by sweep shape (stable within a group) */
        int no = 0;
        for (int g = 0; g <= 7; g++)
            for (int i = 0; i < nt; i++)
                if (H.fgrp[i] == g) H.task_order[no++] = i;
        int nwo = 0;
        for (int g = 0; g <= 6; g++)
            for (int wi = 0; wi < nw; wi++)
                if (H.wgrp[wi] == g) H.win_order[nwo++] = wi;
        const int ring = next_pow2(lds_maxw);
        if ((n_lds_f || n_lds_b) && ring > (1 << 20)) return fail(ctx, SA_ERR_UNSUPPORTED, "band diagonal wider than 2^20");
        constexpr size_t kLdsMax = 65536;
        const size_t f_bytes = sa::forward_block_bytes<T>(ring), b_bytes = sa::backward_block_bytes<T>(ring);
        const bool f_hbm = n_lds_f && f_bytes > kLdsMax, b_hbm = n_lds_b && b_bytes > kLdsMax;
        const size_t f_stride = (f_bytes + 255) / 256 * 256, b_stride = (b_bytes + 255) / 256 * 256;
        const size_t ring_bytes = std::max(f_hbm ? (size_t)n_lds_f * f_stride : 0, b_hbm ? (size_t)n_lds_b * b_stride : 0);

        T *d_fM = ar.take<T>(total_cells);
        T *d_fX = ar.take<T>(total_xy), *d_fY = ar.take<T>(total_xy);
        WinDev *d_wins = ar.take<WinDev>(nw);
        int32_t *d_task_order = ar.take<int32_t>(nt), *d_win_order = ar.take<int32_t>(nw);
        int32_t *d_out = ar.take<int32_t>(3 * out_total);
        int32_t *d_wcount = ar.take<int32_t>(nw);
        double *d_wtotal = ar.take<double>(nw);
        int64_t *d_win_dst = ar.take<int64_t>(nw);
        unsigned char *d_ring = ring_bytes ? ar.take<unsigned char>((int64_t)ring_bytes) : nullptr;
        if (ar.over) {
            const int64_t need = (int64_t)(mark + sizeof(T) * (total_cells + 2 * total_xy) + 12 * out_total + ring_bytes);
            return capacity_error(ctx, "scratch bytes for the forward planes and pair slots", need,
                                  (int64_t)kScratchBytes);
        }
        SA_HIP(ctx, hipMemcpyAsync(d_tasks, H.tasks, sizeof(TaskDev) * nt, hipMemcpyHostToDevice, st));
        SA_HIP(ctx, hipMemcpyAsync(d_wins, H.wins, sizeof(WinDev) * nw, hipMemcpyHostToDevice, st));
        SA_HIP(ctx, hipMemcpyAsync(d_task_order, H.task_order, sizeof(int32_t) * nt, hipMemcpyHostToDevice, st));
        SA_HIP(ctx, hipMemcpyAsync(d_win_order, H.win_order, sizeof(int32_t) * nw, hipMemcpyHostToDevice, st));
        SA_HIP(ctx, hipMemsetAsync(d_err, 0, sizeof(int32_t), st));
        a.wins = d_wins;
        a.fM = d_fM; a.fX = d_fX; a.fY = d_fY;
        a.out = d_out;
        a.win_count = d_wcount;
        a.win_total = d_wtotal;
        a.ring = ring;
        a.prio = 3;
        const int ek = model->ek;
        const int lds_fnt = (lds_maxw <= 64 || n_lds_f >= 4 * std::max(ctx->n_cu, 1)) && lds_maxw <= 256 ? 64 : 256;
        /* forward sweeps, one launch per shape group */
        for (int s = 0; s < nt;) {
            const int g = H.fgrp[H.task_order[s]];
            int e = s;
            while (e < nt && H.fgrp[H.task_order[e]] == g) e++;
            if (g == 0) {
                sa::Args<T> al = a;
                al.task_order = d_task_order + s;
                al.ring_g = f_hbm ? d_ring : nullptr;
                al.ring_g_stride = (int64_t)f_stride;
                SA_HIP(ctx, sa::launch_forward<T>(al, e - s, lds_fnt, st));
            } else if (g >= 5) {
                SA_HIP(ctx, sa::launch_forward_halo<T>(a, d_task_order + s, e - s, g - 5 + 2, ek, st));
            } else {
                SA_HIP(ctx, sa::launch_forward_reg<T>(a, d_task_order + s, e - s, kMW[g - 1][1], kMW[g - 1][0], ek, 0, st));
            }
            s = e;
        }
        /* backward sweeps, one launch per window class; windows whose pairs overflow
           their slots run again with their exact counts */
        auto backward = [&]() -> int {
            for (int s = 0; s < nw;) {
                const int g = H.wgrp[H.win_order[s]];
                int e = s;
                while (e < nw && H.wgrp[H.win_order[e]] == g) e++;
                if (g == 0) {
                    sa::Args<T> al = a;
                    al.win_order = d_win_order + s;
                    al.ring_g = b_hbm ? d_ring : nullptr;
                    al.ring_g_stride = (int64_t)b_stride;
                    SA_HIP(ctx, sa::launch_backward<T>(al, e - s, st));
                } else {
                    SA_HIP(ctx, sa::launch_backward_reg<T>(a, d_win_order + s, e - s, kCpt[g - 1], ek, 0, st));
                }
                s = e;
            }
            SA_HIP(ctx, hipMemcpyAsync(H.wcount, d_wcount, sizeof(int32_t) * nw, hipMemcpyDeviceToHost, st));
            SA_HIP(ctx, hipMemcpyAsync(&err, d_err, sizeof(int32_t), hipMemcpyDeviceToHost, st));
            SA_HIP(ctx, hipStreamSynchronize(st));
            return SA_OK;
        };
        int rc = backward();
        if (rc) return rc;
        if (err & sa::ERR_WINDOWS) {
            if (force_lds) return fail(ctx, SA_ERR_HIP, "the LDS sweeps raised a window error");
            continue;
        }
        int64_t extra = 0;
        for (int wi = 0; wi < nw; wi++)
            if (H.wcount[wi] > H.wins[wi].out_cap) extra += H.wcount[wi];
        if (extra) {
            int32_t *d_more = ar.take<int32_t>(3 * (extra + 1));
            if (ar.over) return capacity_error(ctx, "scratch bytes for the pair slots", (int64_t)(ar.used + 12 * extra),
                                               (int64_t)kScratchBytes);
            /* the new slots follow the first pass's, counted in pairs from d_out */
            int64_t off = ((d_more - d_out) + 2) / 3;
            for (int wi = 0; wi < nw; wi++)
                if (H.wcount[wi] > H.wins[wi].out_cap) {
                    H.wins[wi].out_off = off;
                    H.wins[wi].out_cap = H.wcount[wi];
                    off += H.wcount[wi];
                }
            SA_HIP(ctx, hipMemcpyAsync(d_wins, H.wins, sizeof(WinDev) * nw, hipMemcpyHostToDevice, st));
            if ((rc = backward())) return rc;
        }
        /* final positions in the reference's list order: sub-alignments in split
           order, each in reverse emission order */
        int64_t pos = 0;
        for (int i = 0; i < nt; i++) {
            int64_t E = 0, s = 0;
            for (int32_t w = 0; w < H.nwin[i]; w++) E += H.wcount[H.first_win[i] + w];
            for (int32_t w = 0; w < H.nwin[i]; w++) {
                H.win_dst[H.first_win[i] + w] = pos + E - 1 - s;
                s += H.wcount[H.first_win[i] + w];
            }
            pos += E;
        }
        *n_pairs = pos;
        if (pos > pairs_cap) return capacity_error(ctx, "pairs (the caller's buffer)", pos, pairs_cap);
        int32_t *d_packed = ar.take<int32_t>(3 * pos);
        if (ar.over) return capacity_error(ctx, "scratch bytes for the pairs", (int64_t)(ar.used + 12 * pos),
                                           (int64_t)kScratchBytes);
        SA_HIP(ctx, hipMemcpyAsync(d_win_dst, H.win_dst, sizeof(int64_t) * nw, hipMemcpyHostToDevice, st));
        SA_HIP(ctx, sa::launch_collect(d_wins, d_wcount, d_win_dst, d_out, d_packed, d_x1, d_y1, nw, st));
        if (pos > 0)
            SA_HIP(ctx, hipMemcpyAsync(pairs, d_packed, sizeof(int32_t) * 3 * pos, hipMemcpyDeviceToHost, st));
        SA_HIP(ctx, hipMemcpyAsync(H.wtot, d_wtotal, sizeof(double) * nw, hipMemcpyDeviceToHost, st));
        SA_HIP(ctx, hipStreamSynchronize(st));
        double ll = 0.0;
        for (int i = 0; i < nt; i++)
            if (H.nwin[i] > 0) ll += H.wtot[H.first_win[i] + H.nwin[i] - 1];
        *log_likelihood = ll;
        return SA_OK;
    }
    return fail(ctx, SA_ERR_HIP, "sa_align_noalloc: no sweep held the band");
}

} // namespace

namespace sa_rt {

unsigned char *scratch_setup(int device, hipError_t *err) {
    *err = hipSuccess;
    if (device < 0 || device >= kMaxDevices) return nullptr; /* the entry point then reports it */
    std::lock_guard<std::mutex> lk(g_setup_mu);
    HostPlan &H = g_plan[device];
    if (!H.dev) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)); /* the current device's copy */
        if (e != hipSuccess) { *err = e; return nullptr; }
        H.dev = static_cast<unsigned char *>(p);
    }
    return H.dev;
}

} // namespace sa_rt

extern "C" int sa_align_noalloc(sa_context *ctx, const sa_model *model, const sa_read *read, const sa_params *params,
                                sa_precision precision, sa_pair *pairs, int64_t pairs_cap, int64_t *n_pairs,
                                double *log_likelihood) {
    if (!ctx || !model || !read || !n_pairs || !log_likelihood || (pairs_cap > 0 && !pairs) || pairs_cap < 0)
        return SA_ERR_INVALID_ARGUMENT;
    if (precision != SA_FP32 && precision != SA_FP64) return SA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (ctx->device < 0 || ctx->device >= kMaxDevices || !ctx->scratch)
        return fail(ctx, SA_ERR_UNSUPPORTED, "sa_align_noalloc: no scratch array on this device");
    HostPlan &H = g_plan[ctx->device];
    std::lock_guard<std::mutex> dk(H.mu);
    SA_HIP(ctx, hipSetDevice(ctx->device));
    (void)hipGetLastError(); /* no earlier call's error reaches the launch checks */
    sa_params p;
    if (params) p = *params; else sa_params_default(&p);
    /* getPosteriorProbsWithBanding prerequisites (impl/pairwiseAligner.c:1291-1295) */
    if (!(p.traceback_diagonals >= 1) || !(p.diagonal_expansion >= 0) || p.diagonal_expansion % 2 != 0 ||
        !(p.min_diags_between_traceback >= 2) || !(p.traceback_diagonals + 1 < p.min_diags_between_traceback) ||
        !(p.threshold >= 0.0 && p.threshold <= 1.0))
        return fail(ctx, SA_ERR_INVALID_ARGUMENT, "invalid banding parameters");
    const sa_read &rd = *read;
    if (rd.event_stride < 2 || rd.n_events < 0 || rd.ref_length < 0 || rd.n_anchors < 0 || !(rd.var != 0.0) ||
        (rd.ref_length > 0 && !rd.ref) || (rd.n_events > 0 && !rd.events) || (rd.n_anchors > 0 && !rd.anchors))
        return fail(ctx, SA_ERR_INVALID_ARGUMENT, "bad read sizes");
    if (rd.ref_length > 0 && std::memchr(rd.ref, 'X', (size_t)rd.ref_length))
        return fail(ctx, SA_ERR_UNSUPPORTED, "sa_align_noalloc: degenerate references run through sa_align");
    return precision == SA_FP32 ? align_noalloc<float>(ctx, H, model, rd, p, pairs, pairs_cap, n_pairs, log_likelihood)
                                : align_noalloc<double>(ctx, H, model, rd, p, pairs, pairs_cap, n_pairs, log_likelihood);
}
