/*
 * signalalign_amd.h -- C-ABI of the MI355X-native banded signal-to-reference
 * pair-HMM forward-backward (libsignalalign_amd.so).
 *
 * The library replaces signalAlign's getAlignedPairsUsingAnchors path
 * (impl/pairwiseAligner.c:1871-1899 and everything below it down to
 * stateMachine3_cellCalculate, impl/stateMachine.c:1194-1325) with HIP kernels
 * for gfx950.  Plain pointers and sizes only; all functions return SA_OK (0) or
 * a negative status and never exit the process.
 *
 * Threading: a context serialises its own calls with an internal mutex; use one
 * context per host thread for concurrency.
 */
#ifndef SIGNALALIGN_AMD_H
#define SIGNALALIGN_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SA_OK 0
#define SA_ERR_INVALID_ARGUMENT (-1) /* bad sizes, odd expansion, bad parameters */
#define SA_ERR_HIP (-2)              /* a HIP runtime call failed */
#define SA_ERR_OUT_OF_MEMORY (-3)
#define SA_ERR_BAD_KMER (-4)         /* a reference k-mer has a character outside the alphabet
                                        (the reference exits, impl/nanopore_hdp.c:376-381) */
#define SA_ERR_ANCHORS (-5)          /* anchors not strictly increasing / outside the matrix
                                        (asserts at impl/pairwiseAligner.c:161-167) */
#define SA_ERR_UNSUPPORTED (-6)      /* band wider than the kernels support, unknown emission; a degenerate
                                        reference in FP32 or in an expectations run */
#define SA_ERR_NO_DEVICE (-7)
#define SA_ERR_CAPACITY (-8)         /* sa_align_noalloc: the read needs more than its fixed scratch or
                                        the caller's pair buffer holds (the message names which) */
#define SA_PENDING 1 /* sa_server_wait: the request has not finished */

/* Arithmetic of the recurrence.  SA_FP64 follows the reference's doubles;
 * SA_FP32 runs the cell recurrence in float log space relative to a per-diagonal
 * FP64 offset, with FP64 totals and posterior normalisation. */
typedef enum { SA_FP64 = 0, SA_FP32 = 1 } sa_precision;

/* Emission functions of the reference's three-state machines:
 *   SA_EMIT_GAUSS_DESCALED  emissions_signal_strawManGetKmerEventMatchProbWithDescaling
 *                           (impl/stateMachine.c:498-545), signalMachine's choice
 *                           (getStateMachine3_descaled, :1627-1643)
 *   SA_EMIT_GAUSS           emissions_signal_strawManGetKmerEventMatchProb (:547-589)
 *                           (getStateMachine3, :1645-1653)
 *   SA_EMIT_HDP             emissions_signal_getHdpKmerDensity (:468-495) with
 *                           stateMachine3HDP_cellCalculate (:1259-1325) */
typedef enum { SA_EMIT_GAUSS_DESCALED = 0, SA_EMIT_GAUSS = 1, SA_EMIT_HDP = 2 } sa_emission_kind;

/* Transition slots (natural-log probabilities) in the field order of struct
 * _StateMachine3 (inc/stateMachine.h:141-160). */
enum {
    SA_T_MATCH_CONTINUE = 0, SA_T_MATCH_FROM_GAP_X, SA_T_MATCH_FROM_GAP_Y,
    SA_T_GAP_OPEN_X, SA_T_GAP_OPEN_Y, SA_T_GAP_EXTEND_X, SA_T_GAP_EXTEND_Y,
    SA_T_GAP_SWITCH_TO_X, SA_T_GAP_SWITCH_TO_Y
};

/* Mirrors PairwiseAlignmentParameters (inc/pairwiseAligner.h:123-134): the
 * fields the banded forward-backward reads. */
typedef struct {
    double threshold;                      /* minimum posterior match probability kept */
    int64_t min_diags_between_traceback;   /* 1000 */
    int64_t traceback_diagonals;           /* 40 */
    int64_t diagonal_expansion;            /* even */
    int64_t split_matrix_bigger_than_this; /* 3000*3000 */
} sa_params;

/* Defaults of pairwiseAlignmentBandingParameters_construct (impl/pairwiseAligner.c:1843-1856). */
void sa_params_default(sa_params *p);

/* Model tables.  Gaussian kinds: match and gapy hold 5 doubles per k-mer
 * (level_mean, level_sd, noise_mean, noise_sd, noise_lambda) indexed by kmer_id over
 * the sorted alphabet; gapx holds the log skip emission per k-mer (NULL = log 0.1).
 * HDP kind: a uniform density grid (grid_length knots from grid[0]), per-row density
 * and spline slopes, the row of each k-mer's first observed ancestor DP (kmer_row),
 * and the per-k-mer level mean used for descaling (level_mean). */
typedef struct {
    int32_t kind;           /* sa_emission_kind */
    int32_t kmer_length;
    const char *alphabet;   /* NUL-terminated, any order, distinct characters */
    double transitions[9];
    const double *match;
    const double *gapy;
    const double *gapx;
    int64_t grid_length;
    const double *grid;
    int64_t n_rows;
    const double *density;  /* n_rows * grid_length */
    const double *slopes;   /* n_rows * grid_length */
    const int64_t *kmer_row;
    const double *level_mean;
} sa_model_desc;

/* One read (one strand) to align, the arguments of getAlignedPairsUsingAnchors
 * in flat form.  Anchors are 0-based (k-mer index, event index) pairs, strictly
 * increasing in both coordinates. */
typedef struct {
    const char *ref;          /* nucleotides; k-mer i is ref[i .. i+k) */
    int64_t ref_length;       /* nucleotides: lX = ref_length - (k - 1) k-mers */
    const double *events;     /* n_events records of event_stride doubles: mean at [0], noise at [1] */
    int64_t n_events;
    int64_t event_stride;     /* 4 = NB_EVENT_PARAMS (inc/nanopore.h:4) */
    const int64_t *anchors;   /* 2 * n_anchors */
    int64_t n_anchors;
    double scale, shift, var; /* per-read descaling (StateMachine scale/shift/var) */
    int32_t ragged_left, ragged_right;
    /* The bases a degenerate reference base 'X' stands for (the Sequence's
     * degenerateBases, e.g. "CEO"), or NULL for a reference without 'X'.  A k-mer
     * with 'X' is a multi-path cell: one path per combination of these options
     * (hdCell_construct, impl/pairwiseAligner.c:599-656).  Batches holding such
     * reads run in FP64; every read with 'X' in a batch takes the same options. */
    const char *degenerate_options;
} sa_read;

/* Aligned pair: the reference tuple (floor(p*1e7), x, y, pathKmer) with the
 * k-mer pointer replaced by x (the k-mer is ref[x .. x+k)). */
typedef struct {
    int32_t p;
    int32_t x;
    int32_t y;
} sa_pair;

typedef struct {
    const sa_pair *pairs;   /* in the reference's list order: sub-alignments in split order, each
                               sub-alignment's pairs in reverse emission order */
    int64_t n_pairs;
    double log_likelihood;  /* sum over sub-alignments of the total at the final diagonal */
    int64_t cells;          /* band cells = work units */
    int64_t n_windows;
    int64_t n_subalignments;
    /* per pair, the index of its path in the cell (the first 'X' varying slowest);
     * NULL unless the batch held degenerate references */
    const int32_t *paths;
} sa_read_result;

typedef struct {
    double prep_ms;        /* band, schedule, k-mer ids, event constants (GPU) */
    /* The forward and backward launches of different sub-alignment groups overlap on
     * the streams, so these two are spans on the device timeline, not sums:
     *   forward_ms   start of the sweeps to the end of the last forward launch (in a
     *                batch with degenerate references: to the end of the long group's
     *                forward launch, the shorter group's sweeps running beside it);
     *   backward_ms  start of the first backward launch to the end of the sweeps (on a
     *                one-stream context this span includes forward launches queued after
     *                that first backward launch; with degenerate references: from the end
     *                of the long group's forward launch to the end of the sweeps). */
    double forward_ms;
    double backward_ms;
    double collect_ms;     /* pair compaction */
    double total_ms;       /* whole call on the device, inputs resident to outputs resident */
    int64_t cells;
    int64_t forward_launches;
    int64_t backward_launches;
    double fb_ms;            /* forward + backward sweeps on the device timeline (they overlap) */
    int32_t emission_path;   /* 0: standard Gaussian tables, 1: general Gaussian, 2: HDP grid (gathered by
                                the sweeps), 3: HDP grid through the emission plane */
    int32_t replayed;        /* 1: the run replayed the batch's recorded graph */
    /* the run's first and last device event on the device's timeline, in ms after an
       event the library records once per device (at its first context): runs of
       different contexts on one device are placed on one timeline (their overlap,
       the span of several concurrent runs) */
    double device_start_ms;
    double device_end_ms;
    double emission_ms;      /* HDP emission plane, filled before the sweeps (part of total_ms) */
} sa_timing;

/* One kernel launch of a run: forward sweeps of a group of sub-alignments, or
 * backward sweeps of a group of traceback windows. */
typedef struct {
    int32_t direction;  /* 0 forward, 1 backward */
    int32_t cpt;        /* register window of 64*cpt positions per wavefront; 0 = LDS sweep */
    int32_t waves;      /* wavefronts per sub-alignment (forward) or window (backward) */
    int32_t stream;     /* 0..3 */
    int64_t items;      /* sub-alignments (forward) or windows (backward) */
    int64_t cells;      /* band cells those items cover */
    double start_ms;    /* from the start of the sweeps (HIP events on its stream) */
    double ms;          /* duration */
    int32_t sync_every; /* several-wavefront forward sweeps: diagonals between two
                           synchronisations of the wavefronts (1, or 8 with halo lanes);
                           0 otherwise */
    int32_t persistent; /* 1: the persistent launch (direction 0), which also runs the
                           traceback windows of its items (cells = forward cells) */
} sa_launch;

typedef struct sa_context sa_context;
typedef struct sa_model sa_model;
typedef struct sa_result sa_result;

int sa_context_create(int device, sa_context **out);
/* n_streams 4 (as sa_context_create): sub-alignment groups run on four streams;
 * 1: every launch on one stream (for callers that run several contexts side by side). */
int sa_context_create_ex(int device, int n_streams, sa_context **out);
void sa_context_destroy(sa_context *ctx);
int sa_device_count(int *count);

/* Uploads a model's tables to HBM (resident until sa_model_destroy). */
int sa_model_create(sa_context *ctx, const sa_model_desc *desc, sa_model **out);
void sa_model_destroy(sa_model *model);
int sa_model_kmer_length(const sa_model *model);

/* Plain per-read call: the drop-in for one getAlignedPairsUsingAnchors call. */
int sa_align(sa_context *ctx, const sa_model *model, const sa_read *read, const sa_params *params,
             sa_precision precision, sa_result **out);

/* Plain per-read call that allocates nothing.  No counterpart in the reference: it is
 * sa_align for callers that cannot allocate memory (code that must not call hipMalloc,
 * tight serving loops, harnesses that forbid allocation).  The caller passes the read
 * and a host buffer for the pairs; every intermediate buffer comes from one fixed
 * device array of SA_NOALLOC_SCRATCH_BYTES per device, held until the process ends, and
 * the host side plans in fixed storage.  Pairs, their order and the log-likelihood are
 * sa_align's.  Limits: the read's staged inputs, band, forward planes (4 B per band cell
 * in FP32, 8 B in FP64, plus the refresh diagonals' gap states) and pair slots must fit
 * the scratch -- a 50 kb read at band expansion 100 fits, FP32 or FP64 -- with at most 256
 * split blocks, 65536 anchors and 8192 traceback windows; beyond them, or when more
 * pairs than pairs_cap result, it returns SA_ERR_CAPACITY (*n_pairs then holds the
 * read's pair count when that is what fell short).  No degenerate 'X' references.
 * Calls on one device share the scratch and are serialised by a per-device lock. */
#define SA_NOALLOC_SCRATCH_BYTES (256LL << 20)
int sa_align_noalloc(sa_context *ctx, const sa_model *model, const sa_read *read, const sa_params *params,
                     sa_precision precision, sa_pair *pairs, int64_t pairs_cap, int64_t *n_pairs,
                     double *log_likelihood);

/* Batched call: all reads in one length-sorted launch sequence.  The context keeps the
   call's batch and refills it on its next sa_align_batch or sa_align call, so its device
   buffers (sized for the largest call so far) are allocated once and freed with the
   context. */
int sa_align_batch(sa_context *ctx, const sa_model *model, const sa_read *reads, int64_t n_reads,
                   const sa_params *params, sa_precision precision, sa_result **out);

/* Split form of sa_align_batch for callers that keep inputs resident: create
 * splits the reads into sub-alignments and uploads the raw inputs to HBM; run does
 * all device work (band, schedule, forward, backward, pairs) and copies the pairs
 * back; a batch can be run any number of times. */
typedef struct sa_batch sa_batch;
int sa_batch_create(sa_context *ctx, const sa_model *model, const sa_read *reads, int64_t n_reads,
                    const sa_params *params, sa_batch **out);
int sa_batch_run(sa_batch *batch, sa_precision precision, sa_result **out);
/* Replaces the batch's reads (and model/params), reusing its device buffers. */
int sa_batch_refill(sa_batch *batch, const sa_model *model, const sa_read *reads, int64_t n_reads,
                    const sa_params *params);
/* (a batch is destroyed before the context it was created on) */
void sa_batch_destroy(sa_batch *batch);

/* Launch tuning of a batch.  use_graph: the first run records the launch sequence
 * (band prep, every forward and backward sweep on its stream, pair compaction) as a
 * HIP graph and later runs of the same batch replay it.  waves_per_simd: occupancy
 * cap of the one-wavefront sweeps (0: none).  long_factor: sub-alignments with
 * lX + lY >= max(20000, long_factor x median) run their forward sweep on two to four
 * wavefronts, the fewest whose window holds the band (0: 2.5; negative: never).  stream_split: fraction of the one-wavefront forward
 * cells (longest sub-alignments first) on stream 0, the rest on stream 3 (0: 0.5).
 * persist_waves: 0 or -1, a forward launch and then backward launches per stream
 * (stream_split applies).  1-8 would ask for a persistent launch (the one-wavefront
 * sub-alignments' forward sweeps longest first, then the traceback windows of those
 * done, by worker wavefronts, this many per SIMD); it hung on the device with its
 * cause not found, so sa_batch_set_tuning refuses it with SA_ERR_UNSUPPORTED. */
typedef struct {
    int32_t use_graph;
    int32_t waves_per_simd;
    double long_factor;
    double stream_split;
    int32_t persist_waves;
    int32_t reserved;
} sa_tuning;
void sa_tuning_default(sa_tuning *t);
int sa_batch_set_tuning(sa_batch *batch, const sa_tuning *t);
int sa_batch_get_tuning(const sa_batch *batch, sa_tuning *t);
/* Times runs of the batch under candidate tunings (one knob at a time: long_factor,
 * stream_split, waves_per_simd; each the best of two runs of the forward-backward
 * sweeps, a candidate kept only when at least 1% faster), keeps the fastest on the
 * batch and returns it with its sweep time.  use_graph is left as set. */
int sa_batch_autotune(sa_batch *batch, sa_precision precision, sa_tuning *best, double *best_ms);

int64_t sa_result_n_reads(const sa_result *res);
int sa_result_read(const sa_result *res, int64_t i, sa_read_result *out);
int sa_result_timing(const sa_result *res, sa_timing *out);
/* Launches of the run (up to cap); returns how many there were. */
int64_t sa_result_launches(const sa_result *res, sa_launch *out, int64_t cap);
void sa_result_free(sa_result *res);

/* ---- EM expectations: getExpectationsUsingAnchors -----------------------------
 * Replaces getExpectationsUsingAnchors (impl/pairwiseAligner.c:1983-2003) driven by
 * diagonalCalculation_Expectations (:1255-1275) and cell_signal_updateExpectations /
 * cell_signal_updateExpectationsAndAssignments (:763-811): the same banded,
 * windowed forward-backward (FP64), and at every posterior diagonal the posterior
 * p = exp(F_from + B_to + eP + tP - total) of each transition into each cell.
 *   transitions[from * 3 + to]  summed p (match 0, gapX 1, gapY 2), no pseudocount
 *   likelihood                  the window total summed over the posterior diagonals
 *   match_posterior[k]          summed p of the transitions into the match state of
 *                               cells of k-mer k
 *   match_event_sum[k]          summed p x descaled event mean, (e + var mu_k -
 *                               scale mu_k - shift) / var with the read's parameters
 *                               and mu_k the model's level mean
 *   observed[k]                 1 where a match transition into k-mer k was evaluated
 *   records                     the cells whose largest match-transition posterior is
 *                               >= record_threshold and above zero (threshold 0: every
 *                               cell that adds to the sums), in the reference's call order
 *                               (sub-alignments in order, traceback windows in
 *                               order, each from its top posterior diagonal down,
 *                               cells by ascending x): k-mer x, event y, p of M->M,
 *                               X->M, Y->M.  They carry what the order-dependent
 *                               accumulators (running variance, HDP assignments) need.
 * The result stays valid until sa_expect_free. */
typedef struct {
    int32_t x, y;
    double p[3];
} sa_match_record;
typedef struct sa_expect_result sa_expect_result;
typedef struct {
    double transitions[9];
    double likelihood;
    int64_t n_kmers;
    const double *match_posterior;
    const double *match_event_sum;
    const int8_t *observed;
    const sa_match_record *records;
    int64_t n_records;
} sa_expectations;
int sa_expect(sa_context *ctx, const sa_model *model, const sa_read *read, const sa_params *params,
              double record_threshold, sa_expect_result **out);
/* Many reads in one launch sequence (the EM training pass over a read set); each
 * read's expectations separately.  Record storage is one record per posterior cell
 * at most: large batches pass record_threshold 1 (no records) and use the sums. */
int sa_expect_batch(sa_context *ctx, const sa_model *model, const sa_read *reads, int64_t n_reads,
                    const sa_params *params, double record_threshold, sa_expect_result **out);
int sa_expect_read(const sa_expect_result *res, int64_t i, sa_expectations *out);
int sa_expect_get(const sa_expect_result *res, sa_expectations *out); /* read 0 */
int sa_expect_timing(const sa_expect_result *res, sa_timing *out);
void sa_expect_free(sa_expect_result *res);

/* Host-side views of the integer geometry the kernels compute on the device
 * (same code, sa_common.h), for validation and for callers that plan work:
 *   sa_band_geometry: band_construct (impl/pairwiseAligner.c:131-182) as per-diagonal
 *     x ranges xlo[d]..xhi[d], d = 0..lX+lY; anchors 0-based, strictly increasing.
 *   sa_window_schedule: the traceback windows of getPosteriorProbsWithBanding
 *     (impl/pairwiseAligner.c:1319-1410), 4 values per window (tracedBackTo, the
 *     diagonal the backward pass starts on, tracedBackFrom, last); returns the count.
 *   sa_split_points: getSplitPoints (impl/pairwiseAligner.c:1706-1757), 4 values per
 *     block (x1, y1, x2, y2); returns the count. */
int sa_band_geometry(const int64_t *anchors, int64_t n_anchors, int64_t lX, int64_t lY, int64_t expansion,
                     int32_t *xlo, int32_t *xhi);
int64_t sa_window_schedule(const int32_t *xlo, const int32_t *xhi, int64_t lX, int64_t lY, const sa_params *params,
                           int32_t *out, int64_t cap);
int64_t sa_split_points(const int64_t *anchors, int64_t n_anchors, int64_t lX, int64_t lY, const sa_params *params,
                        int32_t ragged_left, int32_t ragged_right, int64_t *out, int64_t cap);

/* Text of the last error on this context (valid until the next call). */
const char *sa_last_error(const sa_context *ctx);

/* ---- serving: request loop with continuous batching --------------------------
 * Reads are submitted one at a time (inputs copied) and grouped by a scheduler
 * thread into waves of the same model, parameters and precision, admitted while
 * the band-matrix bytes of the waves in flight fit the HBM budget; up to `lanes`
 * waves run at once on the device.  Models may come from any context on the same
 * device and must outlive the requests that use them.  A result does not depend
 * on the wave it ran in. */
typedef struct sa_server sa_server;
typedef struct {
    int64_t hbm_budget_bytes; /* band-matrix bytes of the waves in flight (0: 32 GiB) */
    int32_t lanes;            /* waves in flight, 1-4 (0: 4) */
    int32_t max_wave_reads;   /* reads per wave at most (0: 2048) */
    double length_ratio;      /* a wave takes reads whose lX + lY is within this factor of its
                                 first request's (a wave lasts as long as its longest read);
                                 0: 2; negative: any length */
    double batch_delay_ms;    /* while other waves keep the device busy, a free lane waits up to
                                 this long after the oldest pending request's arrival for
                                 max_wave_reads requests to gather; 0: 30; negative: no wait.
                                 A wave's first request is the longest of those that arrived
                                 within this delay of the oldest pending one */
} sa_server_config;
typedef struct {
    double submit_s, admit_s, done_s; /* seconds on the server's clock */
    int64_t wave;                     /* wave the request ran in */
    int32_t wave_reads;               /* reads in that wave */
    int32_t pad;
} sa_request_timing;
typedef struct {
    int64_t waves, reads, cells, max_wave_reads;
    double device_ms;        /* device time summed over waves (they overlap) */
    double now_s;
    int64_t pending;         /* submitted, not yet admitted */
    int64_t inflight_bytes;  /* band-matrix bytes admitted and not finished */
    double fill_ms;          /* lane host time splitting and uploading waves, summed */
    double run_ms;           /* lane wall time running waves (device work and host planning), summed */
} sa_server_stats;

int sa_server_create(int device, const sa_server_config *config, sa_server **out);
/* Copies the read and queues it.  A degenerate reference (degenerate_options set and
 * an 'X' in ref) runs in FP64 (SA_ERR_UNSUPPORTED for SA_FP32), in waves of reads with
 * the same options. */
int sa_server_submit(sa_server *server, const sa_model *model, const sa_read *read, const sa_params *params,
                     sa_precision precision, int64_t *ticket);
/* Queues n reads at once (copied on several threads), tickets[i] for reads[i]; the
 * scheduler sees them together, as one burst.  No counterpart in the reference: for
 * callers holding many reads.  A bad read fails the call before any is queued. */
int sa_server_submit_batch(sa_server *server, const sa_model *model, const sa_read *reads, int64_t n,
                           const sa_params *params, sa_precision precision, int64_t *tickets);
/* Waits up to timeout_s (0: poll) for a request: SA_OK with its result (valid until
 * sa_server_release), SA_PENDING, or the request's error status. */
int sa_server_wait(sa_server *server, int64_t ticket, double timeout_s, sa_read_result *out,
                   sa_request_timing *timing);
int sa_server_release(sa_server *server, int64_t ticket);
int sa_server_get_stats(sa_server *server, sa_server_stats *out);
/* Finishes the waves in flight, drops pending requests, frees everything. */
void sa_server_destroy(sa_server *server);

#ifdef __cplusplus
}
#endif

#endif
