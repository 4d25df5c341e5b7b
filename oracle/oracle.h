/*
 * oracle.h -- FP64 CPU restatement of signalAlign's banded signal-to-reference
 * pair-HMM forward-backward (the getAlignedPairsUsingAnchors path).
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the correctness checker for the HIP
 * implementation in signalalign_amd/.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  The product never links it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the signalAlign repository root).  Parity status of each function is listed
 * in the header comment of oracle.c.
 */
#ifndef SA_ORACLE_H
#define SA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Emission kinds.
 *   OR_GAUSS_DESCALED: emissions_signal_strawManGetKmerEventMatchProbWithDescaling
 *                      (impl/stateMachine.c:498-545), the signalMachine choice.
 *   OR_GAUSS:          emissions_signal_strawManGetKmerEventMatchProb
 *                      (impl/stateMachine.c:547-589), used by the reference tests.
 *   OR_HDP:            emissions_signal_getHdpKmerDensity (impl/stateMachine.c:468-495)
 *                      with stateMachine3HDP_cellCalculate (:1259-1325). */
enum { OR_GAUSS_DESCALED = 0, OR_GAUSS = 1, OR_HDP = 2 };

/* Transition slots, in the field order of struct _StateMachine3
 * (inc/stateMachine.h:141-160).  Values are natural-log probabilities. */
enum {
    OR_T_MATCH_CONTINUE = 0, OR_T_MATCH_FROM_GAP_X, OR_T_MATCH_FROM_GAP_Y,
    OR_T_GAP_OPEN_X, OR_T_GAP_OPEN_Y, OR_T_GAP_EXTEND_X, OR_T_GAP_EXTEND_Y,
    OR_T_GAP_SWITCH_TO_X, OR_T_GAP_SWITCH_TO_Y
};

typedef struct {
    int32_t kind;
    int32_t kmer_length;
    int32_t alphabet_size;
    char alphabet[16];        /* sorted, NUL terminated */
    double trans[9];
    const double *match;      /* 5 per k-mer: level_mean, level_sd, noise_mean, noise_sd, noise_lambda */
    const double *gapy;       /* 5 per k-mer (level_sd x 1.75 in the .model loader) */
    const double *gapx;       /* 1 per k-mer, log-space (log 0.1 from the constructors) */
    double scale, shift, var; /* per-read descaling parameters */
    /* HDP density grid (OR_HDP only) */
    int64_t grid_length;
    const double *grid;       /* grid_length knots, uniform spacing */
    const double *density;    /* n_rows x grid_length */
    const double *slopes;     /* n_rows x grid_length */
    const int64_t *kmer_row;  /* k-mer id -> row of the first observed ancestor DP */
} or_model;

typedef struct {
    double threshold;
    int64_t min_diags_between_traceback;
    int64_t traceback_diagonals;
    int64_t diagonal_expansion;
    int64_t split_matrix_bigger_than_this;
} or_params;

/* Pair record: the reference tuple (floor(p*1e7), x, y, kmer*) with the k-mer
 * pointer replaced by its x index (the k-mer string is ref[x .. x+k)). */
typedef struct {
    int64_t p;
    int64_t x;
    int64_t y;
    int64_t path;   /* path of the cell (0 unless the k-mer holds degenerate 'X' bases) */
} or_pair;

/* Per-window diagnostics of one aligned sub-problem. */
typedef struct {
    int64_t n_windows;
    int64_t cells;            /* sum of diagonal widths = work units of this call */
    double log_likelihood;    /* sum over sub-alignments of the total at their last diagonal */
} or_stats;

/* EM expectations of one read: getExpectationsUsingAnchors (impl/pairwiseAligner.c:1983-2003)
 * with diagonalCalculation_Expectations (:1255-1275) and cell_signal_updateExpectations
 * (:763-790), filling the accumulators of a ContinuousPairHmm (impl/continuousHmm.c:84-170).
 * The caller sizes the per-k-mer arrays (alphabet_size^kmer_length) and sets the
 * reference's starting values: transitions to the transition pseudocount,
 * posteriors to the emission pseudocount, the rest to zero. */
typedef struct {
    double *transitions;       /* 9: from * 3 + to (match 0, gapX 1, gapY 2) */
    double likelihood;         /* += the window total at every posterior diagonal */
    double *event_exp;         /* 2 per k-mer: sum p x~, running sum p (x~ - u)^2 */
    double *posteriors;        /* per k-mer: pseudocount + sum p */
    unsigned char *observed;   /* per k-mer */
    /* match-transition posteriors in call order: 5 doubles per cell (k-mer position
       x, event y, p of M->M, X->M, Y->M), cells whose largest p >= rec_threshold
       and > 0 */
    double *rec;
    int64_t rec_cap, n_rec;
    double rec_threshold;
    int64_t x1, y1;            /* offset of the current sub-alignment (set by the driver) */
} or_expect;

int or_get_expectations(const or_model *m, const char *ref, int64_t lX, const double *events, int64_t lY,
                        const int64_t *anchors, int64_t n_anchors, const or_params *p, int ragged_left,
                        int ragged_right, or_expect *ex);

void or_set_exact_logadd(int on);
void or_record_windows(int64_t *buf, int64_t cap);
int64_t or_recorded_windows(void);
double or_lookup(double x);
double or_logadd(double x, double y);

int64_t or_kmer_id(const char *kmer, const char *alphabet, int64_t alphabet_size, int64_t kmer_length);

double or_log_gauss_pdf(double x, double mu, double sigma);
double or_log_inv_gauss_pdf(double x, double mu, double lambda);
double or_emission(const or_model *m, const char *kmer, const double *event, int match);

double or_grid_spline_interp(double q, const double *x, const double *y, const double *slope, int64_t length);
void or_spline_knot_slopes(const double *x, const double *y, int64_t length, double *k_out);
void or_linspace(double start, double stop, int64_t length, double *out);
double or_hdp_expected_value(const double *grid, const double *distr, int64_t length);

int or_band(const int64_t *anchors, int64_t n_anchors, int64_t lX, int64_t lY, int64_t expansion,
            int64_t *xmyL, int64_t *xmyR);

int64_t or_split_points(const int64_t *anchors, int64_t n_anchors, int64_t lX, int64_t lY,
                        int64_t split_matrix, int ragged_left, int ragged_right,
                        int64_t *out, int64_t cap);

int or_get_aligned_pairs(const or_model *m, const char *ref, int64_t lX,
                         const double *events, int64_t lY,
                         const int64_t *anchors, int64_t n_anchors,
                         const or_params *p, int ragged_left, int ragged_right,
                         or_pair *out, int64_t cap, int64_t *n_out, or_stats *stats);

/* The same with degenerate 'X' bases in ref: each X position takes every base of
 * `options` (the Sequence's degenerateBases, e.g. "CEO"), a cell one path per
 * combination (hdCell_construct, impl/pairwiseAligner.c:599-656); each pair
 * carries the index of its path in that order (the first X varying slowest). */
int or_get_aligned_pairs_degenerate(const or_model *m, const char *ref, int64_t lX,
                                    const double *events, int64_t lY,
                                    const int64_t *anchors, int64_t n_anchors,
                                    const or_params *p, int ragged_left, int ragged_right, const char *options,
                                    or_pair *out, int64_t cap, int64_t *n_out, or_stats *stats);

/* The path k-mers of one k-mer (its 'X' positions substituted by every combination
 * of options, the first X varying slowest): returns how many; writes k characters
 * per path to out when cap allows.  -1 for an 'X' without options. */
int64_t or_path_kmers(const char *kmer, int k, const char *options, char *out, int64_t cap);
/* path_checkLegal (impl/pairwiseAligner.c:540-552) on two path k-mers (NULL: column 0). */
int or_path_legal(const char *from, const char *to, int k);

int or_full_forward_backward(const or_model *m, const char *ref, int64_t lX,
                             const double *events, int64_t lY,
                             const int64_t *anchors, int64_t n_anchors, int64_t expansion,
                             double threshold,
                             or_pair *out, int64_t cap, int64_t *n_out,
                             double *total_forward, double *total_backward,
                             double *diagonal_totals);

#ifdef __cplusplus
}
#endif

#endif
