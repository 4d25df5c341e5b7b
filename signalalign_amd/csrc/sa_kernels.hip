/*
 * sa_kernels.hip -- gfx950 kernels of the banded signal-to-reference pair-HMM
 * forward-backward (signalAlign's getAlignedPairsUsingAnchors path).
 *
 * Kernels, in launch order:
 *   prep_kernel      one workgroup per sub-alignment: k-mer ids, per-event
 *                    constants, band geometry (band_construct,
 *                    impl/pairwiseAligner.c:131-182), cell counts and the traceback
 *                    window schedule (getPosteriorProbsWithBanding, :1319-1410).
 *   forward_kernel   one workgroup per sub-alignment sweeps its band anti-diagonal
 *                    by anti-diagonal; the three states of the last two diagonals
 *                    live in an LDS ring indexed by x; every diagonal's cells are
 *                    streamed to HBM for the posterior pass.
 *   backward_kernel  one wavefront per traceback window: the windowed backward
 *                    pass (re-initialised with the end-state values at the window
 *                    top), the every-10th-diagonal total refresh with the
 *                    match-flux term (:1182-1200, :1370-1380), and the thresholded
 *                    posterior pairs (:1202-1253) compacted with wave ballots.
 *   collect_kernel   moves each window's pairs to their place in the reference's
 *                    list order (alignedPairCoordinateCorrectionFn, :1862-1869).
 *
 * Cell recurrence: stateMachine3_cellCalculate (impl/stateMachine.c:1194-1257) in
 * pull form; each state accumulates its terms in the order the reference's
 * doTransitionForward/Backward calls reach it, with the lookup log-add of
 * impl/pairwiseAligner.c:237-254.  No MFMA: the recurrence is not a contraction.
 */
#include "sa_device.h"

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace sa {

/* ------------------------------------------------------------------------ */
/* prep: k-mer ids, event constants, band, cell counts, windows             */
/* ------------------------------------------------------------------------ */

template <typename T>
__global__ void __launch_bounds__(256) prep_kernel(Args<T> a) {
    const int ti = a.task_order ? a.task_order[blockIdx.x] : (int)blockIdx.x;
    const TaskDev t = a.tasks[ti];
    const int tid = threadIdx.x;
    const int nt = blockDim.x;
    const int k = a.m.kmer_length;
    const int A = a.m.alphabet_size;
    __shared__ long long s_cells, s_xy;
    __shared__ int s_maxw;
    __shared__ int s_err;
    __shared__ int s_nwin;
    __shared__ int s_span;
    /* the character table and (up to kPrepAnc of) the sub-alignment's anchors in LDS:
       every k-mer reads the table k times and every diagonal binary-searches the
       anchors, each a chain of dependent loads */
    constexpr int kPrepAnc = 2048;
    __shared__ int8_t s_lut[256];
    __shared__ int32_t s_ax[kPrepAnc], s_ay[kPrepAnc];
    const bool anc_lds = t.n_anc <= kPrepAnc;
    for (int i = tid; i < 256; i += nt) s_lut[i] = a.lut[i];
    if (anc_lds)
        for (int i = tid; i < t.n_anc; i += nt) {
            s_ax[i] = a.anc_x[t.anc_off + i];
            s_ay[i] = a.anc_y[t.anc_off + i];
        }
    if (tid == 0) { s_cells = 0; s_xy = 0; s_maxw = 0; s_err = 0; s_nwin = 0; s_span = 1; }
    __syncthreads();
    int err = 0;
    for (int p = tid; p < t.lX; p += nt) {
        int32_t id = 0;
        bool bad = false;
        for (int j = 0; j < k; j++) {
            const char ch = a.ref[t.ref_off + p + j];
            int c = s_lut[(uint8_t)ch];
            bad |= c < 0 && !(a.paths && ch == 'X'); /* an 'X' k-mer's ids come per path */
            id = id * A + (c < 0 ? 0 : c);
        }
        a.kid[t.pos_off + p] = bad ? 0 : id;
        if (bad) err |= ERR_BAD_KMER;
    }
    const double inv_var = 1.0 / t.var;
    for (int j = tid; j < t.lY; j += nt) {
        const double *ev = a.events + (t.ev_off + j) * a.event_stride;
        double mean = ev[0], noise = ev[1];
        double evm = a.m.kind == 1 ? mean : (mean - t.shift) * inv_var;
        a.evm[t.evc_off + j] = (T)evm;
        a.evn[t.evc_off + j] = (T)noise;
        a.evl[t.evc_off + j] = (T)(-1.5 * log(noise));
        a.evi[t.evc_off + j] = (T)(1.0 / noise);
    }
    const int64_t N = (int64_t)t.lX + t.lY;
    long long cells = 0;
    int maxw = 0;
    for (int64_t d = tid; d <= N; d += nt) {
        int64_t L = 0, R = 0;
        const bool ok = anc_lds ? band_at(s_ax, s_ay, t.n_anc, t.lX, t.lY, a.expansion, d, &L, &R)
                                : band_at(a.anc_x + t.anc_off, a.anc_y + t.anc_off, t.n_anc, t.lX, t.lY, a.expansion,
                                          d, &L, &R);
        if (!ok) { err |= ERR_BAND; L = -d; R = -d; }
        int32_t lo = (int32_t)div2(d + L), hi = (int32_t)div2(d + R);
        a.xlo[t.diag_off + d] = lo;
        a.xhi[t.diag_off + d] = hi;
        int w = hi - lo + 1;
        cells += w;
        maxw = w > maxw ? w : maxw;
    }
    atomicAdd((unsigned long long *)&s_cells, (unsigned long long)cells);
    atomicMax(&s_maxw, maxw);
    if (err) atomicOr(&s_err, err);
    __syncthreads();
    if (a.em_plane) {
        /* exclusive prefix sums over i = 0..n-1 of w(i), stored by put(i, sum) */
        __shared__ int s_wsum[4];
        __shared__ int s_carry;
        auto block_scan = [&](int64_t n, auto w_of, auto put) {
            if (tid == 0) s_carry = 0;
            __syncthreads();
            for (int64_t base = 0; base < n; base += nt) {
                const int64_t i = base + tid;
                const int w = i < n ? w_of(i) : 0;
                int v = w; /* inclusive scan over the wavefront */
                for (int o = 1; o < 64; o <<= 1) {
                    const int u = __shfl_up(v, o, 64);
                    if ((tid & 63) >= o) v += u;
                }
                if ((tid & 63) == 63) s_wsum[tid >> 6] = v;
                __syncthreads();
                int pre = s_carry;
                for (int q = 0; q < (tid >> 6); q++) pre += s_wsum[q];
                if (i < n) put(i, pre + v - w);
                __syncthreads();
                if (tid == nt - 1) s_carry = pre + v;
                __syncthreads();
            }
        };
        const int32_t *bl = a.xlo + t.diag_off, *bh = a.xhi + t.diag_off;
        /* the forward match plane's per-diagonal offsets (prefix sums of the band
           widths), which the forward sweeps also write as they go: the HDP emission
           plane shares the layout and is filled before they run */
        block_scan(N + 1, [&](int64_t d) { return bh[d] - bl[d] + 1; },
                   [&](int64_t d, int v) { a.doff[t.diag_off + d] = v; });
        /* per column x = 0..lX the diagonals holding it, [first d with xhi(d) >= x, last
           d with xlo(d) <= x] (bounds are non-decreasing), and its offset in the column
           plane the emission plane is built through */
        int32_t *ca = a.col_a + t.pos_off + ti, *cb = a.col_b + t.pos_off + ti, *co = a.col_off + t.pos_off + ti;
        for (int64_t d = tid; d <= N; d += nt) {
            const int h = bh[d], hp = d > 0 ? bh[d - 1] : -1;
            for (int x = max(hp + 1, 0); x <= min(h, t.lX); x++) ca[x] = (int)d;
            const int l = bl[d], ln = d < N ? bl[d + 1] : t.lX + 1;
            for (int x = max(l, 0); x < min(ln, t.lX + 1); x++) cb[x] = (int)d;
        }
        __syncthreads();
        block_scan((int64_t)t.lX + 1, [&](int64_t x) { return max(0, cb[x] - ca[x] + 1); },
                   [&](int64_t x, int v) { co[x] = v; });
    }
    if (tid == 0) {
        const int32_t *xlo = a.xlo + t.diag_off;
        const int32_t *xhi = a.xhi + t.diag_off;
        auto width = [&](int64_t d) { return (int64_t)(xhi[d] - xlo[d] + 1); };
        int32_t nw = window_schedule(N, a.min_diags, a.traceback, a.expansion, width, a.win_slots + t.win_off,
                                     t.win_cap, ti);
        if (nw < 0) { s_err |= ERR_WINDOWS; nw = 0; }
        a.task_nwin[ti] = nw;
        a.task_cells[ti] = s_cells;
        a.task_maxw[ti] = s_maxw;
        if (s_err) atomicOr(a.err, s_err);
        s_nwin = nw;
    }
    /* the refresh at diagonal e (every 10th posterior diagonal of a window, counted
       down from tracedBackFrom) reads the forward gapX/gapY states of e and e-1 */
    /* (an expectations run keeps every diagonal's gapX/gapY states) */
    for (int64_t d = tid; d <= N; d += nt) a.dflag[t.diag_off + d] = a.expect ? 1 : 0;
    __syncthreads();
    if (tid == 0) {
        for (int w = 0; w < s_nwin; w++) {
            const WinDev W = a.win_slots[t.win_off + w];
            for (int64_t e = W.TF; e > W.T; e -= 10) {
                a.dflag[t.diag_off + e] = 1;
                a.dflag[t.diag_off + e - 1] = 1;
            }
        }
    }
    __syncthreads();
    long long xy = 0;
    for (int64_t d = tid; d <= N; d += nt)
        if (a.dflag[t.diag_off + d]) xy += a.xhi[t.diag_off + d] - a.xlo[t.diag_off + d] + 1;
    atomicAdd((unsigned long long *)&s_xy, (unsigned long long)xy);
    /* window span for the register sweeps: forward at d needs
       [min xlo(d-2..d), xhi(d)], backward at e needs [xlo(e), max xhi(e..e+2)].
       Per aligned block of 64 diagonals (one wavefront's diagonals of an iteration)
       the register-window class: the narrowest lane window of the register sweeps
       (1, 2 or 3 positions per lane; 4: wider) holding the forward span (bits 0-2)
       and the backward span (bits 3-5) of every diagonal in the block.  The sweeps
       change window width at block boundaries. */
    int span = 1, hspan = 1;
    const int32_t *XL = a.xlo + t.diag_off, *XH = a.xhi + t.diag_off;
    auto span_class = [](int sp) { return sp <= 64 ? 1 : sp <= 127 ? 2 : sp <= 190 ? 3 : 4; };
    auto wave_class = [](int c) {
        return __ballot(c >= 4) ? 4 : __ballot(c >= 3) ? 3 : __ballot(c >= 2) ? 2 : 1;
    };
    for (int64_t base = 0; base <= N; base += nt) {
        const int64_t d = base + tid;
        int cf = 1, cb = 1;
        if (d <= N) {
            int lmin = XL[d];
            if (d >= 1) lmin = min(lmin, XL[d - 1]);
            if (d >= 2) lmin = min(lmin, XL[d - 2]);
            int hmax = XH[d];
            if (d + 1 <= N) hmax = max(hmax, XH[d + 1]);
            if (d + 2 <= N) hmax = max(hmax, XH[d + 2]);
            const int fsp = XH[d] - lmin + 1, bsp = hmax - XL[d] + 1;
            span = max(span, max(fsp, bsp));
            if (d % kHaloK == 0) {
                /* the halo sweep's window at a synchronisation: d and d-1 kept, the
                   next kHaloK diagonals computed */
                const int lk = d >= 1 ? min(XL[d], XL[d - 1]) : XL[d];
                hspan = max(hspan, XH[d + kHaloK < N ? d + kHaloK : N] - lk + 1);
            }
            cf = span_class(fsp);
            cb = span_class(bsp);
        }
        cf = wave_class(cf);
        cb = wave_class(cb);
        if ((tid & 63) == 0 && d <= N) a.dcls[t.diag_off + d] = (int8_t)(cf | cb << 3);
    }
    atomicMax(&s_span, span);
    __shared__ int s_hspan;
    if (tid == 0) s_hspan = 1;
    __syncthreads();
    atomicMax(&s_hspan, hspan);
    __syncthreads();
    if (tid == 0) {
        a.task_xycells[ti] = s_xy;
        a.task_span[ti] = s_span;
        if (a.task_hspan) a.task_hspan[ti] = s_hspan;
    }
    /* per traceback window: the backward span over its diagonals (T, D] (each
       window's sweep picks the narrowest register window that holds it) and the
       cells of its posterior diagonals (T, TF] */
    for (int w = 0; w < s_nwin; w++) {
        WinDev *W = a.win_slots + t.win_off + w;
        const int64_t wT = W->T, wD = W->D, wTF = W->TF;
        __syncthreads();
        if (tid == 0) { s_span = 1; s_cells = 0; }
        __syncthreads();
        int wspan = 1;
        long long wcells = 0;
        for (int64_t e = wT + 1 + tid; e <= wD; e += nt) {
            int hmax = XH[e];
            if (e + 1 <= N) hmax = max(hmax, XH[e + 1]);
            if (e + 2 <= N) hmax = max(hmax, XH[e + 2]);
            wspan = max(wspan, hmax - XL[e] + 1);
            if (e <= wTF) wcells += XH[e] - XL[e] + 1;
        }
        atomicMax(&s_span, wspan);
        atomicAdd((unsigned long long *)&s_cells, (unsigned long long)wcells);
        __syncthreads();
        if (tid == 0) { W->span = s_span; W->cells = s_cells; }
    }
}

/* ------------------------------------------------------------------------ */
/* forward                                                                   */
/* ------------------------------------------------------------------------ */

template <typename T, int NT, bool LDSR>
__global__ void __launch_bounds__(NT) forward_kernel(Args<T> a) {
    constexpr bool NORM = sizeof(T) == 4;
    extern __shared__ __align__(16) unsigned char smem[];
    const int R = a.ring, RM = R - 1;
    /* the ring in LDS (LDSR: addressed as LDS, which a pointer that may be either is
       not), or (bands too wide for LDS) in this workgroup's HBM slice */
    unsigned char *mem = LDSR ? smem : a.ring_g + (size_t)blockIdx.x * a.ring_g_stride;
    T *ring = reinterpret_cast<T *>(mem);                /* [3 diag][3 state][R] */
    T *red = ring + 9 * R;                              /* NT/64 wave maxima */
    const int ti = a.task_order[blockIdx.x];
    const TaskDev t = a.tasks[ti];
    const Model<T> &m = a.m;
    const int tid = threadIdx.x;
    const int64_t N = (int64_t)t.lX + t.lY;
    const int32_t *xlo = a.xlo + t.diag_off;
    const int32_t *xhi = a.xhi + t.diag_off;
    int32_t *doff = a.doff + t.diag_off;
    int32_t *doffxy = a.doffxy + t.diag_off;
    const int8_t *dflag = a.dflag + t.diag_off;
    double *foff = a.foff + t.diag_off;
    const int64_t base = t.cell_off;
    const int64_t xybase = t.xy_off;
    const int32_t *kid = a.kid + t.pos_off;
    const int64_t evc = t.evc_off;
    const T s = (T)(m.kind == 1 ? 1.0 : t.scale / t.var);
    const T inv_var = (T)(1.0 / t.var);
    const T tMC = m.trans[0], tMFX = m.trans[1], tMFY = m.trans[2], tGOX = m.trans[3], tGOY = m.trans[4];
    const T tGEX = m.trans[5], tGEY = m.trans[6], tSWX = m.trans[7];
    const T NI = ninf<T>();

    /* diagonal 0: start or ragged-start state values (impl/stateMachine.c:1022-1031) */
    {
        const int lo = xlo[0], w = xhi[0] - lo + 1;
        for (int i = tid; i < w; i += NT) {
            const int x = lo + i;
            T v0 = t.ragged_left ? NI : (T)0, v12 = t.ragged_left ? (T)0 : NI;
            ring[0 * R + (x & RM)] = v0;
            ring[1 * R + (x & RM)] = v12;
            ring[2 * R + (x & RM)] = v12;
            a.fM[base + i] = v0;
            if (dflag[0]) {
                a.fX[xybase + i] = v12;
                a.fY[xybase + i] = v12;
            }
        }
        if (tid == 0) { doff[0] = 0; doffxy[0] = 0; foff[0] = 0.0; }
    }
    int64_t run = xhi[0] - xlo[0] + 1;
    int64_t runxy = dflag[0] ? run : 0;
    double off1 = 0.0, off2 = 0.0;
    __syncthreads();

    for (int64_t d = 1; d <= N; d++) {
        const int lo = xlo[d], hi = xhi[d], w = hi - lo + 1;
        const int lo1 = xlo[d - 1], hi1 = xhi[d - 1];
        const int lo2 = d >= 2 ? xlo[d - 2] : 1, hi2 = d >= 2 ? xhi[d - 2] : 0;
        const int c = (int)(d % 3), c1 = (int)((d + 2) % 3), c2 = (int)((d + 1) % 3);
        T *rc = ring + c * 3 * R, *r1 = ring + c1 * 3 * R, *r2 = ring + c2 * 3 * R;
        const T dl2 = NORM ? (T)(off2 - off1) : (T)0;
        const bool keep_xy = dflag[d] != 0;
        T mx = NI;
        /* cells of d: FP64 values go straight to the ring and the planes; FP32 values
           go to the ring unnormalised and are shifted by the diagonal's maximum in a
           second pass over the same cells (each thread re-reads its own) */
        for (int i = tid; i < w; i += NT) {
            const int x = lo + i;
            const int64_t y = d - x;
            const bool hasL = x - 1 >= lo1 && x - 1 <= hi1;
            const bool hasU = x >= lo1 && x <= hi1;
            const bool hasM = x - 1 >= lo2 && x - 1 <= hi2;
            const T LM = hasL ? r1[(x - 1) & RM] : NI, LX = hasL ? r1[R + ((x - 1) & RM)] : NI,
                    LY = hasL ? r1[2 * R + ((x - 1) & RM)] : NI;
            const T UM = hasU ? r1[x & RM] : NI, UY = hasU ? r1[2 * R + (x & RM)] : NI;
            T MM = hasM ? r2[(x - 1) & RM] : NI, MX = hasM ? r2[R + ((x - 1) & RM)] : NI,
              MY = hasM ? r2[2 * R + ((x - 1) & RM)] : NI;
            if (NORM) { MM += dl2; MX += dl2; MY += dl2; }
            const bool kx = x >= 1, ky = y >= 1;
            const int32_t kk = kid[kx ? x - 1 : 0];
            const int64_t jj = evc + (ky ? y - 1 : 0);
            T eX = kx ? Emit<T>::gapx(m, kk) : NI;
            T eM = NI, eY = NI;
            if (kx && ky && (hasM || hasU)) Emit<T>::eval(m, a, kk, jj, s, inv_var, eM, eY);
            /* lower -> gapX, middle -> match, upper -> gapY (impl/stateMachine.c:1209-1256) */
            T X = LM + (eX + tGOX);
            X = logadd(X, LX + (eX + tGEX));
            X = logadd(X, LY + (eX + tSWX));
            T M = MM + (eM + tMC);
            M = logadd(M, MX + (eM + tMFX));
            M = logadd(M, MY + (eM + tMFY));
            T Y = UM + (eY + tGOY);
            Y = logadd(Y, UY + (eY + tGEY));
            rc[x & RM] = M;
            rc[R + (x & RM)] = X;
            rc[2 * R + (x & RM)] = Y;
            if (NORM) {
                T m3 = M > X ? M : X;
                m3 = m3 > Y ? m3 : Y;
                mx = m3 > mx ? m3 : mx;
            } else {
                a.fM[base + run + i] = M;
                if (keep_xy) {
                    a.fX[xybase + runxy + i] = X;
                    a.fY[xybase + runxy + i] = Y;
                }
            }
        }
        double offd = 0.0;
        if (NORM) {
            mx = wave_max(mx);
            if (NT > 64) {
                if ((tid & 63) == 0) red[tid >> 6] = mx;
                __syncthreads();
                mx = red[0];
#pragma unroll
                for (int q = 1; q < NT / 64; q++) mx = red[q] > mx ? red[q] : mx;
            }
            const T sh = mx > NI ? mx : (T)0;
            offd = off1 + (double)sh;
            for (int i = tid; i < w; i += NT) {
                const int x = lo + i;
                const T vM = rc[x & RM] - sh, vX = rc[R + (x & RM)] - sh, vY = rc[2 * R + (x & RM)] - sh;
                rc[x & RM] = vM;
                rc[R + (x & RM)] = vX;
                rc[2 * R + (x & RM)] = vY;
                a.fM[base + run + i] = vM;
                if (keep_xy) {
                    a.fX[xybase + runxy + i] = vX;
                    a.fY[xybase + runxy + i] = vY;
                }
            }
        }
        if (tid == 0) {
            doff[d] = (int32_t)run;
            doffxy[d] = keep_xy ? (int32_t)runxy : -1;
            foff[d] = offd;
        }
        run += w;
        if (keep_xy) runxy += w;
        off2 = off1;
        off1 = offd;
        __syncthreads();
    }
}

/* ------------------------------------------------------------------------ */
/* backward + total refresh + posterior pairs, one wavefront per window     */
/* ------------------------------------------------------------------------ */

/* Sequential log-add chain of the reference's dpDiagonal_dotProduct
 * (impl/pairwiseAligner.c:1014-1027).  Terms at least 7.5 below the running
 * prefix maximum are exactly no-ops (logAdd returns its accumulator) and are
 * skipped without changing the result. */
__device__ __forceinline__ double chain(const double *term, int n) {
    double acc = -__builtin_huge_val();
    double pm = -__builtin_huge_val();
    for (int i = 0; i < n; i++) {
        const double v = term[i];
        if (!(v > pm - 7.5)) continue;
        acc = logadd(acc, v);
        pm = v > pm ? v : pm;
    }
    return acc;
}

/* diagonalCalculation_Expectations (impl/pairwiseAligner.c:1255-1275) for posterior
 * diagonal e of window W, FP64: each cell's transitions from the forward cells of
 * e-1 (lower, upper) and e-2 (middle) into its backward states (rc, the ring),
 * p = exp(F_from + B_to + (eP + tP) - total), in stateMachine3_cellCalculate's
 * order (cell_signal_updateExpectations, :763-790).  The reference has deleted
 * forward diagonal T-1 when window (T, D] runs, so diagonal T+1 sees no middle
 * cells.  Match transitions also add to the k-mer's posterior and descaled-event
 * sums (continuousPairHmm_addToEmissionExpectation, impl/continuousHmm.c:160-169)
 * and, above the record threshold, to the window's records in call order. */
__device__ __forceinline__ void expect_diagonal(const Args<double> &a, const TaskDev &t, const WinDev &W, int wi,
                                             int e, int lo, int w, const double *rc, int R, double total,
                                             double *ex, double &ex_like, int &ex_cursor) {
    const Model<double> &m = a.m;
    const int lane = threadIdx.x;
    const int RM = R - 1;
    const int32_t *xlo = a.xlo + t.diag_off;
    const int32_t *xhi = a.xhi + t.diag_off;
    const int32_t *doff = a.doff + t.diag_off;
    const int32_t *doffxy = a.doffxy + t.diag_off;
    const int32_t *kid = a.kid + t.pos_off;
    const int64_t evc = t.evc_off;
    const double s = m.kind == 1 ? 1.0 : t.scale / t.var;
    const double inv_var = 1.0 / t.var;
    const double tMC = m.trans[0], tMFX = m.trans[1], tMFY = m.trans[2], tGOX = m.trans[3], tGOY = m.trans[4];
    const double tGEX = m.trans[5], tGEY = m.trans[6], tSWX = m.trans[7];
    ex_like += total;
    const int lo1 = xlo[e - 1], hi1 = xhi[e - 1];
    const bool mid = e - 2 >= W.T && e >= 2;
    const int lo2 = mid ? xlo[e - 2] : 1, hi2 = mid ? xhi[e - 2] : 0;
    const int64_t f1 = t.cell_off + doff[e - 1], fxy1 = t.xy_off + doffxy[e - 1];
    const int64_t f2 = mid ? t.cell_off + doff[e - 2] : 0, fxy2 = mid ? t.xy_off + doffxy[e - 2] : 0;
    double *s0 = a.ex_s0 + (int64_t)W.task * m.n_kmers;
    double *s1 = a.ex_s1 + (int64_t)W.task * m.n_kmers;
    int8_t *obs = a.ex_obs + (int64_t)W.task * m.n_kmers;
    const double *mu = m.kind == 2 ? m.hmu : m.mmu;
    for (int i0 = 0; i0 < w; i0 += 64) {
        const int i = i0 + lane;
        bool keep = false;
        double q0 = 0.0, q1 = 0.0, q2 = 0.0;
        int x = 0, y = 0;
        if (i < w) {
            x = lo + i;
            y = e - x;
            const double BM = rc[x & RM], BX = rc[R + (x & RM)], BY = rc[2 * R + (x & RM)];
            if (x - 1 >= lo1 && x - 1 <= hi1) { /* lower (x-1, y) */
                const int64_t j = x - 1 - lo1;
                const double eX = Emit<double>::gapx(m, kid[x - 1]);
                ex[0 * 3 + 1] += exp(a.fM[f1 + j] + BX + (eX + tGOX) - total);
                ex[1 * 3 + 1] += exp(a.fX[fxy1 + j] + BX + (eX + tGEX) - total);
                ex[2 * 3 + 1] += exp(a.fY[fxy1 + j] + BX + (eX + tSWX) - total);
            }
            if (x - 1 >= lo2 && x - 1 <= hi2) { /* middle (x-1, y-1): x >= 1, y >= 1 */
                const int64_t j = x - 1 - lo2;
                const int32_t kk = kid[x - 1];
                const double eM = Emit<double>::match(m, a, kk, evc + (y - 1), s, inv_var);
                q0 = exp(a.fM[f2 + j] + BM + (eM + tMC) - total);
                q1 = exp(a.fX[fxy2 + j] + BM + (eM + tMFX) - total);
                q2 = exp(a.fY[fxy2 + j] + BM + (eM + tMFY) - total);
                ex[0] += q0;
                ex[1 * 3 + 0] += q1;
                ex[2 * 3 + 0] += q2;
                /* emissions_signal_descaleEventMean_JordanStyle (impl/stateMachine.c:285-289) */
                const double xm = a.events[(t.ev_off + y - 1) * a.event_stride];
                const double u = mu[kk];
                const double xd = (xm + t.var * u - t.scale * u - t.shift) / t.var;
                const double qs = q0 + q1 + q2;
                atomicAdd(&s0[kk], qs);
                atomicAdd(&s1[kk], qs * xd);
                obs[kk] = 1;
                const double qm = fmax(q0, fmax(q1, q2));
                keep = qm >= a.ex_rec_threshold && qm > 0.0; /* a zero posterior adds nothing */
            }
            if (x >= lo1 && x <= hi1 && x >= 1) { /* upper (x, y-1); k-mer x-1 (none at x = 0) */
                const int64_t j = x - lo1;
                const double eY = Emit<double>::gapy(m, a, kid[x - 1], evc + (y - 1), s, inv_var);
                ex[0 * 3 + 2] += exp(a.fM[f1 + j] + BY + (eY + tGOY) - total);
                ex[2 * 3 + 2] += exp(a.fY[fxy1 + j] + BY + (eY + tGEY) - total);
            }
        }
        const unsigned long long bal = __ballot(keep);
        if (keep) {
            const int idx = ex_cursor + __popcll(bal & ((1ull << lane) - 1ull));
            if (idx < W.rec_cap) {
                ExRec r;
                r.x = x - 1;
                r.y = y - 1;
                r.p[0] = q0;
                r.p[1] = q1;
                r.p[2] = q2;
                a.ex_rec[W.rec_off + idx] = r;
            }
        }
        ex_cursor += __popcll(bal);
    }
}

template <typename T, bool EX, bool LDSR>
__global__ void __launch_bounds__(64) backward_kernel(Args<T> a) {
    constexpr bool NORM = sizeof(T) == 4;
    extern __shared__ __align__(16) unsigned char smem[];
    const int R = a.ring, RM = R - 1;
    /* the ring and the refresh terms in LDS (LDSR: addressed as LDS), or (bands too
       wide for LDS) in this workgroup's HBM slice */
    unsigned char *mem = LDSR ? smem : a.ring_g + (size_t)blockIdx.x * a.ring_g_stride;
    double *termA = reinterpret_cast<double *>(mem);    /* R */
    double *termB = termA + R;                          /* R */
    double *bcast = termB + R;                          /* 2 */
    T *ring = reinterpret_cast<T *>(bcast + 2);         /* [3 diag][3 state][R] */
    const int wi = a.win_order[blockIdx.x];
    const WinDev W = a.wins[wi];
    const TaskDev t = a.tasks[W.task];
    const Model<T> &m = a.m;
    const int lane = threadIdx.x;
    const int32_t *xlo = a.xlo + t.diag_off;
    const int32_t *xhi = a.xhi + t.diag_off;
    const int32_t *doff = a.doff + t.diag_off;
    const int32_t *doffxy = a.doffxy + t.diag_off;
    const double *foff = a.foff + t.diag_off;
    const int64_t base = t.cell_off;
    const int64_t xybase = t.xy_off;
    const int32_t *kid = a.kid + t.pos_off;
    const int64_t evc = t.evc_off;
    const T s = (T)(m.kind == 1 ? 1.0 : t.scale / t.var);
    const T inv_var = (T)(1.0 / t.var);
    const T tMC = m.trans[0], tMFX = m.trans[1], tMFY = m.trans[2], tGOX = m.trans[3], tGOY = m.trans[4];
    const T tGEX = m.trans[5], tGEY = m.trans[6], tSWX = m.trans[7];
    const T NI = ninf<T>();
    const T thr = (T)a.threshold;
    const int D = W.D, TF = W.TF, TB = W.T;
    const bool ragged_end = W.last && t.ragged_right;
    const T e0 = ragged_end ? m.rendv[0] : m.endv[0];
    const T e1 = ragged_end ? m.rendv[1] : m.endv[1];
    const T e2 = ragged_end ? m.rendv[2] : m.endv[2];
    int32_t *out = a.out + 3 * W.out_off;
    int cursor = 0;
    int posterior_count = 0;
    double ex[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}; /* this lane's transition posterior sums */
    double ex_like = 0.0;
    int ex_cursor = 0;
    double total = -__builtin_huge_val();
    double ob1 = 0.0, ob2 = 0.0; /* offsets of B[e+1], B[e+2] */

    for (int e = D; e > TB; e--) {
        const int lo = xlo[e], hi = xhi[e], w = hi - lo + 1;
        const bool h1 = e + 1 <= D, h2 = e + 2 <= D;
        const int lo1 = h1 ? xlo[e + 1] : 1, hi1 = h1 ? xhi[e + 1] : 0;
        const int lo2 = h2 ? xlo[e + 2] : 1, hi2 = h2 ? xhi[e + 2] : 0;
        const int c = e % 3, c1 = (e + 1) % 3, c2 = (e + 2) % 3;
        T *rc = ring + c * 3 * R, *r1 = ring + c1 * 3 * R, *r2 = ring + c2 * 3 * R;
        T mx = NI;
        if (e == D) {
            for (int i = lane; i < w; i += 64) {
                const int x = lo + i;
                rc[x & RM] = e0;
                rc[R + (x & RM)] = e1;
                rc[2 * R + (x & RM)] = e2;
            }
        } else {
            const T dl2 = NORM ? (T)(ob2 - ob1) : (T)0;
            for (int i = lane; i < w; i += 64) {
                const int x = lo + i;
                const int y = e - x;
                const bool hasLow = x + 1 >= lo1 && x + 1 <= hi1; /* C = (x+1, y) on e+1 */
                const bool hasUp = x >= lo1 && x <= hi1;          /* C = (x, y+1) on e+1 */
                const bool hasMid = x + 1 >= lo2 && x + 1 <= hi2; /* C = (x+1, y+1) on e+2 */
                const T ClowX = hasLow ? r1[R + ((x + 1) & RM)] : NI;
                const T CupY = hasUp ? r1[2 * R + (x & RM)] : NI;
                T CmidM = hasMid ? r2[(x + 1) & RM] : NI;
                if (NORM) CmidM += dl2;
                const int32_t kx = kid[x < t.lX ? x : 0];
                const int32_t kxm1 = kid[x >= 1 ? x - 1 : 0];
                const int64_t jj = evc + (y < t.lY ? y : 0);
                const T eX = hasLow ? Emit<T>::gapx(m, kx) : NI;
                const T eM = hasMid ? Emit<T>::match(m, a, kx, jj, s, inv_var) : NI;
                const T eY = (hasUp && x >= 1) ? Emit<T>::gapy(m, a, kxm1, jj, s, inv_var) : NI;
                /* contributions arrive middle (from e+2), then upper, then lower */
                T M = CmidM + (eM + tMC);
                M = logadd(M, CupY + (eY + tGOY));
                M = logadd(M, ClowX + (eX + tGOX));
                T X = CmidM + (eM + tMFX);
                X = logadd(X, ClowX + (eX + tGEX));
                T Y = CmidM + (eM + tMFY);
                Y = logadd(Y, CupY + (eY + tGEY));
                Y = logadd(Y, ClowX + (eX + tSWX));
                /* the ring slot of e holds e+3, which no cell of e reads */
                rc[x & RM] = M;
                rc[R + (x & RM)] = X;
                rc[2 * R + (x & RM)] = Y;
                if (NORM) {
                    T m3 = M > X ? M : X;
                    m3 = m3 > Y ? m3 : Y;
                    mx = m3 > mx ? m3 : mx;
                }
            }
        }
        double obe = 0.0;
        if (NORM && e != D) {
            mx = wave_max(mx);
            const T sh = mx > NI ? mx : (T)0;
            for (int i = lane; i < w; i += 64) { /* each lane shifts the cells it wrote */
                const int x = lo + i;
                rc[x & RM] -= sh;
                rc[R + (x & RM)] -= sh;
                rc[2 * R + (x & RM)] -= sh;
            }
            obe = ob1 + (double)sh;
        }
        __syncthreads();

        if (e <= TF) {
            const int64_t fb = base + doff[e];
            const double offe = NORM ? foff[e] + obe : 0.0;
            if (posterior_count % 10 == 0) {
                /* diagonalCalculationTotalProbability (impl/pairwiseAligner.c:1182-1200) */
                const int64_t fxy = xybase + doffxy[e];
                for (int i = lane; i < w; i += 64) {
                    const int x = lo + i;
                    T cell = a.fM[fb + i] + rc[x & RM];
                    cell = logadd(cell, a.fX[fxy + i] + rc[R + (x & RM)]);
                    cell = logadd(cell, a.fY[fxy + i] + rc[2 * R + (x & RM)]);
                    termA[i] = (double)cell + offe;
                }
                int w1 = 0;
                if (h1) {
                    const int lo1e = xlo[e + 1];
                    w1 = xhi[e + 1] - lo1e + 1;
                    const int lom = xlo[e - 1], him = xhi[e - 1];
                    const int64_t fbm = base + doff[e - 1];
                    const int64_t fxym = xybase + doffxy[e - 1];
                    const double offm = NORM ? foff[e - 1] + ob1 : 0.0;
                    for (int i = lane; i < w1; i += 64) {
                        const int x = lo1e + i;        /* cell (x, y) on e+1, middle (x-1, y-1) on e-1 */
                        const int y = e + 1 - x;
                        const bool hm = x - 1 >= lom && x - 1 <= him && x >= 1 && y >= 1;
                        T Mv = NI;
                        if (hm) {
                            const int64_t fi = x - 1 - lom;
                            const T eM = Emit<T>::match(m, a, kid[x - 1], evc + (y - 1), s, inv_var);
                            Mv = a.fM[fbm + fi] + (eM + tMC);
                            Mv = logadd(Mv, a.fX[fxym + fi] + (eM + tMFX));
                            Mv = logadd(Mv, a.fY[fxym + fi] + (eM + tMFY));
                        }
                        termB[i] = (double)(Mv + r1[x & RM]) + offm;
                    }
                }
                __syncthreads();
                if (lane == 0) {
                    double tot = chain(termA, w);
                    if (h1) tot = logadd(tot, chain(termB, w1));
                    bcast[0] = tot;
                }
                __syncthreads();
                total = bcast[0];
                if (e == TF && lane == 0) a.win_total[wi] = total;
            }
            posterior_count++;
            /* diagonalCalculationPosteriorMatchProbs (impl/pairwiseAligner.c:1202-1253) */
            const T cst = NORM ? (T)(offe - total) : (T)0;
            for (int i0 = 0; i0 < w; i0 += 64) {
                const int i = i0 + lane;
                bool keep = false;
                double p = 0.0;
                int x = 0, y = 0;
                if (i < w) {
                    x = lo + i;
                    y = e - x;
                    if (x > 0 && y > 0) {
                        T pv;
                        if (NORM) pv = fexp((a.fM[fb + i] + rc[x & RM]) + cst);
                        else pv = fexp((a.fM[fb + i] + rc[x & RM]) - (T)total);
                        keep = pv >= thr;
                        p = (double)pv;
                    }
                }
                const unsigned long long bal = __ballot(keep);
                if (keep) {
                    const int idx = cursor + __popcll(bal & ((1ull << lane) - 1ull));
                    if (idx < W.out_cap) {
                        if (p > 1.0) p = 1.0;
                        out[3 * idx + 0] = (int32_t)floor(p * (double)kPairProb1);
                        out[3 * idx + 1] = x - 1;
                        out[3 * idx + 2] = y - 1;
                    }
                }
                cursor += __popcll(bal);
            }
            if constexpr (EX) expect_diagonal(a, t, W, wi, e, lo, w, rc, R, total, ex, ex_like, ex_cursor);
        }
        ob2 = ob1;
        ob1 = obe;
    }
    if (lane == 0) a.win_count[wi] = cursor;
    if constexpr (EX) {
        for (int k = 0; k < 9; k++) {
            double v = ex[k];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
            if (lane == 0) a.ex_trans[10 * (int64_t)wi + k] = v;
        }
        if (lane == 0) {
            a.ex_trans[10 * (int64_t)wi + 9] = ex_like;
            a.ex_count[wi] = ex_cursor;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* collect: reference list order                                             */
/* ------------------------------------------------------------------------ */

__global__ void __launch_bounds__(256) collect_kernel(const WinDev *wins, const int32_t *win_count,
                                                      const int64_t *win_dst, const int32_t *src, int32_t *dst,
                                                      const int32_t *task_x1, const int32_t *task_y1, int n_wins) {
    const int wi = blockIdx.x;
    if (wi >= n_wins) return;
    const WinDev W = wins[wi];
    const int n = win_count[wi];
    const int64_t last = win_dst[wi];
    const int32_t x1 = task_x1[W.task], y1 = task_y1[W.task];
    const int32_t *s = src + 3 * W.out_off;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const int64_t o = last - j;
        dst[3 * o + 0] = s[3 * j + 0];
        dst[3 * o + 1] = s[3 * j + 1] + x1;
        dst[3 * o + 2] = s[3 * j + 2] + y1;
    }
}

/* ------------------------------------------------------------------------ */
/* launchers                                                                 */
/* ------------------------------------------------------------------------ */

template <typename T>
hipError_t launch_prep(const Args<T> &a, int n_tasks, hipStream_t st) {
    hipLaunchKernelGGL(prep_kernel<T>, dim3(n_tasks), dim3(256), 0, st, a);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_forward(const Args<T> &a, int n_tasks, int nt, hipStream_t st) {
    const size_t lds = a.ring_g ? 0 : forward_block_bytes<T>(a.ring);
    const bool g = a.ring_g != nullptr;
    if (nt == 64 && !g) hipLaunchKernelGGL((forward_kernel<T, 64, true>), dim3(n_tasks), dim3(64), lds, st, a);
    else if (nt == 64) hipLaunchKernelGGL((forward_kernel<T, 64, false>), dim3(n_tasks), dim3(64), 0, st, a);
    else if (nt == 256 && !g) hipLaunchKernelGGL((forward_kernel<T, 256, true>), dim3(n_tasks), dim3(256), lds, st, a);
    else if (nt == 256) hipLaunchKernelGGL((forward_kernel<T, 256, false>), dim3(n_tasks), dim3(256), 0, st, a);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

template <typename T>
hipError_t launch_backward(const Args<T> &a, int n_wins, hipStream_t st) {
    const size_t lds = a.ring_g ? 0 : backward_block_bytes<T>(a.ring);
    /* EM expectations: FP64 only (the runtime refuses FP32 expectations runs) */
    const bool g = a.ring_g != nullptr;
    if constexpr (sizeof(T) == 8) {
        if (a.expect) {
            if (g) hipLaunchKernelGGL((backward_kernel<T, true, false>), dim3(n_wins), dim3(64), 0, st, a);
            else hipLaunchKernelGGL((backward_kernel<T, true, true>), dim3(n_wins), dim3(64), lds, st, a);
            return hipGetLastError();
        }
    }
    if (g) hipLaunchKernelGGL((backward_kernel<T, false, false>), dim3(n_wins), dim3(64), 0, st, a);
    else hipLaunchKernelGGL((backward_kernel<T, false, true>), dim3(n_wins), dim3(64), lds, st, a);
    return hipGetLastError();
}

__global__ void zero_kernel(int32_t *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

hipError_t launch_zero(int32_t *p, int n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(zero_kernel, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

hipError_t launch_collect(const WinDev *wins, const int32_t *win_count, const int64_t *win_dst, const int32_t *src,
                          int32_t *dst, const int32_t *task_x1, const int32_t *task_y1, int n_wins, hipStream_t st) {
    if (n_wins == 0) return hipSuccess;
    hipLaunchKernelGGL(collect_kernel, dim3(n_wins), dim3(256), 0, st, wins, win_count, win_dst, src, dst, task_x1,
                       task_y1, n_wins);
    return hipGetLastError();
}

template hipError_t launch_prep<float>(const Args<float> &, int, hipStream_t);
template hipError_t launch_prep<double>(const Args<double> &, int, hipStream_t);
template hipError_t launch_forward<float>(const Args<float> &, int, int, hipStream_t);
template hipError_t launch_forward<double>(const Args<double> &, int, int, hipStream_t);
template hipError_t launch_backward<float>(const Args<float> &, int, hipStream_t);
template hipError_t launch_backward<double>(const Args<double> &, int, hipStream_t);

} // namespace sa
