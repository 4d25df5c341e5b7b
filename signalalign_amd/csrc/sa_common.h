/*
 * sa_common.h -- data layout shared by the host runtime and the HIP kernels,
 * plus the integer band/window geometry (host and device).
 */
#ifndef SA_COMMON_H
#define SA_COMMON_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#define SA_HD __host__ __device__ __forceinline__

namespace sa {

constexpr int kMaxKmer = 12;
constexpr int kPairProb1 = 10000000;      /* PAIR_ALIGNMENT_PROB_1 */
constexpr int kMaxRing = 1024;           /* widest band diagonal the kernels accept */

/* Error bits raised by device prep. */
enum : int32_t { ERR_BAD_KMER = 1, ERR_BAND = 2, ERR_WINDOWS = 4, ERR_QUEUE = 8 };

/* One sub-alignment (a split block of one read). */
struct TaskDev {
    int64_t ref_off;    /* first nucleotide of this block in the packed reference bytes */
    int64_t ev_off;     /* first event record (in records) in the packed event array */
    int64_t pos_off;    /* first k-mer slot in the per-position arrays */
    int64_t evc_off;    /* first slot in the per-event constant arrays */
    int64_t diag_off;   /* first slot in the per-diagonal arrays (N+1 entries) */
    int64_t anc_off;    /* first anchor (matrix coordinates) of this block */
    int64_t cell_off;   /* first cell of the forward match plane (set after prep) */
    int64_t xy_off;     /* first cell of the forward gapX/gapY planes (set after prep) */
    int32_t lX, lY;
    int32_t n_anc;
    int32_t ragged_left, ragged_right;
    int32_t read;
    int32_t win_off;    /* first window slot (capacity win_cap) */
    int32_t win_cap;
    double scale, shift, var;
};

/* One traceback window of one task (getPosteriorProbsWithBanding's outer
 * traceback, impl/pairwiseAligner.c:1319-1410). */
struct WinDev {
    int32_t task;
    int32_t T;     /* tracedBackTo */
    int32_t D;     /* diagonal where the backward pass starts */
    int32_t TF;    /* tracedBackFrom: posteriors for T < d <= TF */
    int32_t last;  /* atEnd */
    int32_t out_cap;
    int32_t span;  /* positions a register window must cover over (T, D] (set by prep) */
    int32_t rec_cap; /* expectation records this window's slot holds */
    int64_t out_off;
    int64_t cells; /* cells of the diagonals (T, TF] (set by prep) */
    int64_t rec_off; /* first expectation record of this window (EM expectations runs) */
};

/* One cell's match-transition posteriors of an EM expectations run: k-mer position
 * x, event y (0-based, sub-alignment coordinates on the device) and the posteriors
 * of M->M, X->M, Y->M into the cell (cell_signal_updateExpectations,
 * impl/pairwiseAligner.c:763-790). */
struct ExRec {
    int32_t x, y;
    double p[3];
};

SA_HD int64_t div2(int64_t v) { return v / 2; } /* C truncation, as the reference */

SA_HD int64_t bound_coord(int64_t z, int64_t lz) { return z < 0 ? 0 : (z > lz ? lz : z); }

/* band_setCurrentDiagonal (impl/pairwiseAligner.c:108-126).  Returns false where
 * the reference would throw or assert. */
SA_HD bool band_diag(int64_t xay, int64_t xL, int64_t yL, int64_t xU, int64_t yU, int64_t *outL, int64_t *outR) {
    int64_t xmyL = xL - yL;
    int64_t xmyR = xU - yU;
    if (!(xay >= xL + yU) || !(xay <= xU + yL)) return false;
    if ((xay + xmyL) % 2 != 0) xmyL += 1;
    if ((xay + xmyR) % 2 != 0) xmyR += 1;
    int64_t x = div2(xay + xmyL);
    if (x < xL) xmyL += 2 * (xL - x);
    int64_t y = div2(xay - xmyL);
    if (yL < y) xmyL += 2 * (y - yL);
    x = div2(xay + xmyR);
    if (xU < x) xmyR -= 2 * (x - xU);
    y = div2(xay - xmyR);
    if (y < yU) xmyR -= 2 * (yU - y);
    if ((xay + xmyL) % 2 != 0 || (xay + xmyR) % 2 != 0 || xmyL > xmyR) return false;
    *outL = xmyL;
    *outR = xmyR;
    return true;
}

/* Diagonal d of band_construct (impl/pairwiseAligner.c:131-182) for anchors given
 * in matrix coordinates (x+1, y+1), strictly increasing.  Diagonal d >= 1 is bounded
 * by the previous constraint point P[j-1] and the next P[j], where P[0] = (0,0),
 * P[1..n] = anchors and P[n+1] = (lX, lY), and xay(P[j-1]) < d <= xay(P[j]). */
SA_HD bool band_at(const int32_t *ax, const int32_t *ay, int32_t n, int64_t lX, int64_t lY, int64_t e, int64_t d,
                   int64_t *xmyL, int64_t *xmyR) {
    if (d == 0) return band_diag(0, 0, 0, 0, 0, xmyL, xmyR);
    /* smallest j in [1, n+1] with xay(P[j]) >= d */
    int32_t lo = 1, hi = n + 1;
    while (lo < hi) {
        int32_t mid = (lo + hi) >> 1;
        int64_t mx = mid <= n ? (int64_t)ax[mid - 1] + ay[mid - 1] : lX + lY;
        if (mx >= d) hi = mid; else lo = mid + 1;
    }
    int32_t j = lo;
    int64_t px = j - 1 >= 1 ? ax[j - 2] : 0, py = j - 1 >= 1 ? ay[j - 2] : 0;
    int64_t nx = j <= n ? ax[j - 1] : lX, ny = j <= n ? ay[j - 1] : lY;
    int64_t pxay = px + py, pxmy = px - py, nxay = nx + ny, nxmy = nx - ny;
    int64_t xL = bound_coord(div2(pxay + pxmy - e), lX);
    int64_t yL = bound_coord(div2(nxay - (nxmy - e)), lY);
    int64_t xU = bound_coord(div2(nxay + nxmy + e), lX);
    int64_t yU = bound_coord(div2(pxay - (pxmy + e)), lY);
    return band_diag(d, xL, yL, xU, yU, xmyL, xmyR);
}

/* The traceback schedule of getPosteriorProbsWithBanding: a window ends at the
 * first diagonal d >= tracedBackTo + minDiags whose width is <= 2e+1, or at the last
 * diagonal; its posteriors cover (tracedBackTo, d - (traceBack+1)] (all of it at the
 * end).  `width(d)` is supplied by the caller.  Returns the number of windows, or -1
 * if `cap` is too small. */
template <typename WidthFn>
SA_HD int32_t window_schedule(int64_t N, int64_t min_diags, int64_t traceback, int64_t e, WidthFn width,
                              WinDev *out, int32_t cap, int32_t task) {
    int32_t n = 0;
    int64_t T = 0;
    if (N == 0) return 0;
    int64_t d = T + min_diags;
    for (;;) {
        int64_t D = -1;
        for (; d <= N; d++) {
            if (d == N || width(d) <= 2 * e + 1) { D = d; break; }
        }
        if (D < 0) D = N;
        int last = D == N;
        int64_t TF = D - (last ? 0 : traceback + 1);
        if (n >= cap) return -1;
        WinDev w;
        w.task = task;
        w.T = (int32_t)T;
        w.D = (int32_t)D;
        w.TF = (int32_t)TF;
        w.last = last;
        w.out_cap = 0;
        w.span = 1;
        w.rec_cap = 0;
        w.rec_off = 0;
        w.out_off = 0;
        w.cells = 0;
        out[n++] = w;
        if (last) break;
        T = TF;
        d = T + min_diags;
        if (d <= D) d = D + 1; /* the forward sweep only moves forward */
    }
    return n;
}

/* Upper bound on the windows of an N-diagonal block. */
SA_HD int32_t window_cap(int64_t N, int64_t min_diags, int64_t traceback) {
    int64_t step = min_diags - traceback - 1;
    return (int32_t)(N / (step > 0 ? step : 1) + 2);
}

} // namespace sa

#endif
