/*
 * sa_kernels.h -- kernel argument blocks and launchers (internal to the library).
 */
#ifndef SA_KERNELS_H
#define SA_KERNELS_H

#include "sa_common.h"

namespace sa {

template <typename T>
struct Model {
    int32_t kind;          /* sa_emission_kind */
    int32_t kmer_length;
    int32_t alphabet_size;
    int32_t n_kmers;
    /* Gaussian: per-k-mer constants of the match (m*) and extra-event (y*) tables:
     * mu = level_mean, isd = 1/level_sd (0 when sd == 0), ca = -log(sqrt(2 pi)) - log(sd)
     * + (log(lambda) - log(2 pi))/2 (-inf when sd == 0), nmu = noise_mean,
     * nb = lambda / (2 noise_mean^2); gx = log skip emission. */
    const T *mmu, *misd, *mca, *mnmu, *mnb;
    const T *ymu, *yisd, *yca, *ynmu, *ynb;
    const T *gx;
    /* HDP: per-k-mer level mean and density row; grid knots and interleaved
     * (density, slope) pairs per row. */
    const T *hmu;
    const int32_t *hrow;
    const T *gridx;
    const T *ys;
    int32_t grid_length;
    int32_t use_swx;        /* GAP_SWITCH_TO_X above -inf: the gapY -> gapX term is summed */
    int32_t std_y;          /* Gaussian tables as a model file gives them: the extra-event
                               emission comes from the match table's constants */
    T inv_dx;
    T grid_x0, grid_xn, grid_dx; /* first and last knot, knot spacing (knot 1 - knot 0) */
    T gx_const;             /* the skip emission when every k-mer has the same one */
    T trans[9];
    T endv[3];
    T rendv[3];
};

template <typename T>
struct Args {
    Model<T> m;
    const TaskDev *tasks;
    const WinDev *wins;
    const int32_t *task_order;
    const int32_t *win_order;
    /* raw inputs, packed for the batch */
    const char *ref;
    const double *events;
    int64_t event_stride;
    const int32_t *anc_x;   /* matrix coordinates (x+1) */
    const int32_t *anc_y;
    const int8_t *lut;      /* character -> sorted-alphabet index, -1 outside */
    /* derived per position / event / diagonal */
    int32_t *kid;
    T *evm, *evn, *evl, *evi;
    int32_t *xlo, *xhi, *doff;
    int8_t *dflag;          /* 1 where the forward gapX/gapY states are kept */
    int8_t *dcls;           /* at d = 64b: register-window classes of diagonals 64b..64b+63
                               (forward bits 0-2, backward bits 3-5; see prep_kernel) */
    int32_t *doffxy;        /* offset of a kept diagonal in the gapX/gapY planes */
    double *foff;
    int64_t *task_cells;
    int64_t *task_xycells;
    int32_t *task_maxw;
    int32_t *task_span;     /* positions a register window must cover (see prep_kernel) */
    int32_t *task_hspan;    /* positions the halo sweep's window must cover (nullable) */
    int32_t *task_nwin;
    WinDev *win_slots;
    int32_t *err;
    /* forward planes: match state of every cell; gapX/gapY states of the
       diagonals the total-probability refresh reads (d and d-1 of each refresh) */
    T *fM, *fX, *fY;
    /* HDP emission plane (register sweeps with EK 3): the density emission of each
       cell, laid out like fM; em_plane asks prep_kernel for the per-diagonal plane
       offsets (doff) the plane kernel needs before any forward sweep runs */
    T *em;
    int32_t em_plane;
    /* the plane's column pass: per position slot (pos_off + task index + x, x = 0..lX)
       the first and last diagonal holding column x and the column's offset in the
       column plane (prep); the column plane (cell_off-based, column after column) */
    int32_t *col_a, *col_b, *col_off;
    T *emc;
    /* window outputs: 3 int32 (p, x, y) per pair */
    int32_t *out;
    int32_t *win_count;
    double *win_total;
    double threshold;
    int64_t min_diags, traceback, expansion;
    int32_t ring;
    int32_t prio;           /* forward sweeps: wave priority on shared SIMDs (0-3) */
    /* LDS sweeps whose ring does not fit in LDS: per-workgroup slices of HBM */
    unsigned char *ring_g;
    int64_t ring_g_stride;
    /* EM expectations (FP64 LDS sweeps only; diagonalCalculation_Expectations,
       impl/pairwiseAligner.c:1255-1275): every diagonal keeps its gapX/gapY states;
       per window 9 transition posterior sums (from * 3 + to) and the likelihood
       (sum of the window total over its posterior diagonals); per task and k-mer the
       sum of match-transition posteriors and of posterior x descaled event mean, and
       the k-mers a match transition was evaluated for; per window the cells whose
       largest match-transition posterior reaches ex_rec_threshold, in call order */
    int32_t expect;
    double ex_rec_threshold;
    double *ex_trans;       /* 10 per window */
    double *ex_s0, *ex_s1;  /* n_kmers per task */
    int8_t *ex_obs;         /* n_kmers per task */
    ExRec *ex_rec;          /* W.rec_off.. (capacity W.rec_cap) */
    int32_t *ex_count;      /* records per window */
    /* degenerate references: 'X' is a valid reference character (prep) */
    int32_t paths;
};

/* Degenerate references ('X' bases, multi-path cells): per sub-alignment, column x
 * (0..lX; column x >= 1 is k-mer x-1) holds path entries pp[x] .. pp[x+1]-1 (numbered
 * per sub-alignment from 0), its k-mer covers the 'X' sites sa[x] .. sb[x]-1 (site
 * index order = position order); per entry its column and the k-mer id of its
 * substituted k-mer; per diagonal the first cell of its entries in the forward
 * planes (all three states of every diagonal are kept). */
/* Per diagonal of a degenerate task: its band [lo, hi], its entries [e0, e1)
 * (task-relative), and the absolute offset of its entries in the forward planes. */
struct PathDiag {
    int32_t lo, hi, e0, e1;
    int64_t foff;
};
/* Per entry (column x, path p) of a degenerate task: the path's k-mer id and its
 * legal predecessors in column x-1, entries pfirst + i * pstep for i < pn
 * (task-relative; pn = 0 in column 0). */
struct PathEnt {
    int32_t x, p, pfirst, pstep, pn, kk;
};
struct PathTab {
    const PathDiag *diag;     /* diag_off + d */
    const PathEnt *ent;       /* ent_off[task] + entry */
    const int32_t *maxent;    /* per task: the most entries on one diagonal */
    const int64_t *col_off;   /* per task: first slot of its column arrays (lX + 2) */
    const int64_t *ent_off;   /* per task: first slot of its entry arrays */
    const int64_t *plane_off; /* per task: first cell of its forward planes */
    const int32_t *pp, *sa, *sb;
    const int32_t *ecol, *ekid;
    const int64_t *doffE;     /* diag_off + d */
    int32_t nopt;             /* options per 'X' site */
    int32_t ring;             /* entries per ring row (power of two, >= the entries of any diagonal) */
};
inline size_t paths_forward_bytes(int R) { return (size_t)9 * R * sizeof(double); }
inline size_t paths_backward_bytes(int R) { return (size_t)11 * R * sizeof(double) + 2 * sizeof(double); }
hipError_t launch_paths_forward(const Args<double> &a, const PathTab &P, int n_tasks, hipStream_t st);
hipError_t launch_paths_backward(const Args<double> &a, const PathTab &P, int n_wins, hipStream_t st);

template <typename T> hipError_t launch_prep(const Args<T> &a, int n_tasks, hipStream_t st);
/* bytes of LDS (or HBM slice) one workgroup of the LDS sweeps needs for a ring of R
   positions: three diagonals of three states (+ wave maxima); the backward sweep also
   keeps the two refresh term rows and a broadcast slot in FP64 */
template <typename T> constexpr size_t forward_block_bytes(int R) { return (size_t)9 * R * sizeof(T) + 16 * sizeof(T); }
template <typename T> constexpr size_t backward_block_bytes(int R) {
    return (size_t)2 * R * sizeof(double) + 2 * sizeof(double) + (size_t)9 * R * sizeof(T);
}
template <typename T> hipError_t launch_forward(const Args<T> &a, int n_tasks, int nt, hipStream_t st);
template <typename T> hipError_t launch_backward(const Args<T> &a, int n_wins, hipStream_t st);
template <typename T>
hipError_t launch_forward_reg(const Args<T> &a, const int32_t *order, int n, int cpt, int nw, int ek, size_t shmem,
                              hipStream_t st);
/* several wavefronts (2, 3 or 4) per sub-alignment, one position per lane, synchronised
   every kHaloK diagonals through halo lanes (forward_halo_kernel) */
constexpr int kHaloH = 8, kHaloK = 8, kHaloS = 64 - kHaloH;
inline int halo_window(int nw) { return 64 + kHaloS * (nw - 1); }
template <typename T>
hipError_t launch_forward_halo(const Args<T> &a, const int32_t *order, int n, int nw, int ek, hipStream_t st);
/* Work queue of one persistent launch (fb_persistent_kernel): ctr[0] next forward item,
   ctr[1] publication slots taken, ctr[2] the publication slot whose windows are being
   claimed; done[k] = the k-th sub-alignment to finish its forward sweep + 1 (0 until
   published); wclaim per sub-alignment: windows claimed.  ctr, done and wclaim start
   at zero (3 + n_fwd + n_tasks ints, contiguous). */
struct PersistQ {
    int32_t *ctr;
    int32_t *done;
    int32_t *wclaim;
    const int32_t *first_win; /* per sub-alignment: its first window in Args::wins */
    const int32_t *order;     /* the n_fwd forward items, longest first */
    int32_t n_fwd;
    int32_t zero;             /* 0 (read-modify-write reads OR it in) */
    int64_t wait_ticks;       /* 100 MHz real-time ticks a worker waits without progress before it gives up */
};
template <typename T>
hipError_t launch_fb_persistent(const Args<T> &a, const PersistQ &q, int workers, int ek, hipStream_t st);
template <typename T>
hipError_t launch_backward_reg(const Args<T> &a, const int32_t *order, int n, int cpt, int ek, size_t shmem,
                               hipStream_t st);
/* The HDP emission plane of a set of sub-alignments, in two passes: the column pass
   over column groups (3 ints each: density row or -1 for columns with no k-mer, first
   entry, entries; entries are 2 ints: task, column x, grouped by row), then the
   transpose into the plane's layout over chunks c0..c1-1 of kPlaneDiags diagonals
   (chunk c belongs to task ctask[c], whose first chunk is chunk0[task]). */
constexpr int kPlaneDiags = 32;
constexpr int kColumnGroup = 64;
template <typename T>
hipError_t launch_hdp_plane(const Args<T> &a, const int32_t *entries, const int32_t *groups, int n_groups,
                            const int32_t *ctask, const int32_t *chunk0, int c0, int c1, int n_cu, hipStream_t st);
/* The column entries of the plane's sub-alignments, grouped on the device: one
   PlaneTask per sub-alignment (in stream-lane order, lane 0-3), whose columns 0..lX
   take the flat indices col0..col0+lX.  Writes the entries (task, x int32 pairs)
   stably sorted by key lane * (nrow + 1) + density row + 1, and the runs of equal
   keys (key, count; their number in *n_runs) in ascending key order. */
constexpr int kPlaneLanes = 4;
struct PlaneTask {
    int64_t pos_off, col0;
    int32_t task, lane, lX, pad;
};
size_t plane_sort_temp_bytes(int64_t n_cols, int32_t nrow); /* 0: the sort cannot be sized */
hipError_t launch_plane_sort(const PlaneTask *pt, int n_tasks, int64_t n_cols, const int32_t *kid,
                             const int32_t *hrow, int32_t nrow, uint32_t *keys, uint32_t *keys_sorted,
                             uint64_t *vals, uint64_t *entries, uint32_t *run_keys, int32_t *run_counts,
                             int32_t *n_runs, void *temp, size_t temp_bytes, hipStream_t st);
/* zeroes n ints (a kernel, so that it can sit in a captured graph) */
hipError_t launch_zero(int32_t *p, int n, hipStream_t st);
hipError_t launch_collect(const WinDev *wins, const int32_t *win_count, const int64_t *win_dst, const int32_t *src,
                          int32_t *dst, const int32_t *task_x1, const int32_t *task_y1, int n_wins, hipStream_t st);

} // namespace sa

#endif
