/*
 * oracle.c -- FP64 CPU restatement of signalAlign's banded signal-to-reference
 * pair-HMM forward-backward, written from the reference's behaviour.
 *
 * TEST INFRASTRUCTURE ONLY: the correctness checker for the HIP path.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Reference files followed (paths relative to the signalAlign repository):
 *   impl/pairwiseAligner.c  band, logAdd, diagonal DP, windowed traceback,
 *                           total probability, posterior pairs, split points
 *   impl/stateMachine.c     3-state recurrence, start/end priors, emissions
 *   impl/nanopore_hdp.c     k-mer ids, HDP density query
 *   impl/hdp.c              density clamp, expected value
 *   impl/hdp_math_utils.c   cubic Hermite spline slopes and interpolation
 *
 * Parity pins (tests/test_oracle_*.py): band KAT (tests/signalPairwiseAlignerTest.c:447-510),
 * split-point KAT (:376-445), logAdd tolerance test (:121-131) and closed form, emission
 * pdfs against scipy closed forms, k-mer id KATs, the 5-mer tiny forward/backward KAT
 * (tests/stateMachineTests.c:565-692: 7 pairs, forward==backward total, per-diagonal
 * totals), brute-force path enumeration with exact log-sum-exp, spline invariants, and
 * window-schedule invariants (every diagonal gets exactly one posterior); the real-read
 * counts of tests/stateMachineTests.c:823-851 and :896-947 (1076 pairs; 7349 with every
 * C a degenerate 'X'); for degenerate cells the path KATs
 * (tests/signalPairwiseAlignerTest.c:556-581, tests/variableOrderPairwiseAlignerTests.c:107-123)
 * and the mixture identity over site options (tests/test_degenerate.py); for the
 * expectations, explicit enumeration of the 3-state paths (tests/test_expectations.py).
 * The windowed traceback (window schedule, the end-prior seed of each window's top
 * diagonal, the ragged-end prior only at the end of the matrix, the posterior range
 * of each window) is pinned by explicit enumeration of the band paths that end on each
 * window's top diagonal (tests/test_oracle_kats.py, test_windowed_traceback_brute_force).
 * Parity unpinned: the every-10th-diagonal total refresh (impl/pairwiseAligner.c:1370):
 * with exact arithmetic every posterior diagonal of a window has the same total, so
 * only a run of the reference could fix which diagonal's lookup-logAdd total is used.
 *
 * The restatement keeps the reference's operation order: the per-cell transition
 * order of stateMachine3_cellCalculate, push-form backward accumulation, xmy-ascending
 * diagonal walks, the lookup logAdd with float-literal coefficients, the every-10th
 * diagonal total refresh and the floor(p*1e7) output.  It replaces the per-cell heap
 * objects (HDCell/Path) by flat per-diagonal arrays of path entries: column x holds
 * one entry per path of its k-mer (the 'X' positions substituted by the degenerate
 * base options, hdCell_construct, impl/pairwiseAligner.c:599-656), one entry when the
 * k-mer has no 'X'.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OR_LOG_ZERO (-INFINITY)
#define OR_PAIR_PROB_1 10000000
#define OR_NUM_OF_KMERS 46656
#define OR_LOG_UNDERFLOW 7.5

enum { M_ST = 0, X_ST = 1, Y_ST = 2 };

static int g_exact_logadd = 0;
static int64_t *g_win_buf = NULL;
static int64_t g_win_cap = 0, g_win_n = 0;

void or_set_exact_logadd(int on) { g_exact_logadd = on; }

/* Test hook: record (tracedBackTo, traceback diagonal, tracedBackFrom, atEnd) of
 * every traceback window run by the next alignments; NULL stops recording. */
void or_record_windows(int64_t *buf, int64_t cap) {
    g_win_buf = buf;
    g_win_cap = cap;
    g_win_n = 0;
}
int64_t or_recorded_windows(void) { return g_win_n; }

/* impl/pairwiseAligner.c:237-248 -- piecewise cubic approximation of log(exp(x)+1)
 * on [0, 7.5].  The coefficients are float literals promoted to double. */
double or_lookup(double x) {
    if (x <= 1.00f)
        return ((-0.009350833524763f * x + 0.130659527668286f) * x + 0.498799810682272f) * x + 0.693203116424741f;
    if (x <= 2.50f)
        return ((-0.014532321752540f * x + 0.139942324101744f) * x + 0.495635523139337f) * x + 0.692140569840976f;
    if (x <= 4.50f)
        return ((-0.004605031767994f * x + 0.063427417320019f) * x + 0.695956496475118f) * x + 0.514272634594009f;
    return ((-0.000458661602210f * x + 0.009695946122598f) * x + 0.930734667215156f) * x + 0.168037164329057f;
}

/* impl/pairwiseAligner.c:250-254.  The exact variant (used only to pin the oracle
 * against brute-force enumeration) is log(exp(x)+exp(y)). */
double or_logadd(double x, double y) {
    if (g_exact_logadd) {
        if (x == OR_LOG_ZERO) return y;
        if (y == OR_LOG_ZERO) return x;
        return x > y ? x + log1p(exp(y - x)) : y + log1p(exp(x - y));
    }
    if (x < y)
        return (x == OR_LOG_ZERO || y - x >= OR_LOG_UNDERFLOW) ? y : or_lookup(y - x) + x;
    return (y == OR_LOG_ZERO || x - y >= OR_LOG_UNDERFLOW) ? x : or_lookup(x - y) + y;
}

/* impl/nanopore_hdp.c:360-393 (kmer_to_word + word_id): big-endian base-|alphabet|
 * number of the k-mer over the sorted alphabet.  Returns -1 where the reference
 * exits on a character outside the alphabet. */
int64_t or_kmer_id(const char *kmer, const char *alphabet, int64_t alphabet_size, int64_t kmer_length) {
    int64_t id = 0;
    for (int64_t i = 0; i < kmer_length; i++) {
        int64_t j = 0;
        while (j < alphabet_size && alphabet[j] != kmer[i]) j++;
        if (j == alphabet_size) return -1;
        id = id * alphabet_size + j;
    }
    return id;
}

/* impl/stateMachine.c:237-247 */
double or_log_gauss_pdf(double x, double mu, double sigma) {
    if (sigma == 0.0) return OR_LOG_ZERO;
    double log_inv_sqrt_2pi = -0.91893853320467267;
    double l_sigma = log(sigma);
    double a = (x - mu) / sigma;
    return log_inv_sqrt_2pi - l_sigma + (-0.5 * a * a);
}

/* impl/stateMachine.c:226-235 */
double or_log_inv_gauss_pdf(double x, double mu, double lambda) {
    double l_twoPi = 1.8378770664093453;
    double l_x = log(x);
    double a = (x - mu) / mu;
    double l_lambda = log(lambda);
    return (l_lambda - l_twoPi - 3 * l_x - lambda * a * a / x) / 2;
}

/* impl/hdp_math_utils.c:471-496, with the left index clamped to length-2 (the
 * reference reads past the grid when rounding puts a query just below the last
 * knot into the last interval; the clamp changes nothing elsewhere). */
double or_grid_spline_interp(double q, const double *x, const double *y, const double *slope, int64_t length) {
    if (q <= x[0]) return y[0] - slope[0] * (x[0] - q);
    if (q >= x[length - 1]) {
        int64_t n = length - 1;
        return y[n] + slope[n] * (q - x[n]);
    }
    double dx = x[1] - x[0];
    int64_t il = (int64_t)((q - x[0]) / dx);
    if (il > length - 2) il = length - 2;
    int64_t ir = il + 1;
    double dy = y[ir] - y[il];
    double a = slope[il] * dx - dy;
    double b = dy - slope[ir] * dx;
    double tl = (q - x[il]) / dx;
    double tr = 1.0 - tl;
    return tr * y[il] + tl * y[ir] + tl * tr * (a * tr + b * tl);
}

/* impl/hdp_math_utils.c:402-442: tridiagonal solve for the knot slopes of a C2
 * cubic Hermite spline with natural end conditions.  The reference recurses
 * forward (elimination) and assigns on the way back (substitution); this loop
 * performs the same arithmetic in the same order. */
void or_spline_knot_slopes(const double *x, const double *y, int64_t length, double *k) {
    int64_t f = length - 1;
    double *cc = (double *)malloc(sizeof(double) * (size_t)length);
    double *rc = (double *)malloc(sizeof(double) * (size_t)length);
    double *rh = (double *)malloc(sizeof(double) * (size_t)length);
    double right0 = 1.0 / (x[1] - x[0]);
    double center0 = 2.0 * right0;
    double rhs0 = 3.0 * (y[1] - y[0]) * right0 * right0;
    double c_prev = center0, r_prev = right0, rhs_prev = rhs0;
    for (int64_t idx = 1; idx < f; idx++) {
        double left = 1.0 / (x[idx] - x[idx - 1]);
        double right = 1.0 / (x[idx + 1] - x[idx]);
        double center = 2.0 * (left + right);
        double rhs = 3.0 * ((y[idx] - y[idx - 1]) * left * left + (y[idx + 1] - y[idx]) * right * right);
        center -= left * r_prev / c_prev;
        rhs -= left * rhs_prev / c_prev;
        cc[idx] = center; rc[idx] = right; rh[idx] = rhs;
        c_prev = center; r_prev = right; rhs_prev = rhs;
    }
    {
        double left = 1.0 / (x[f] - x[f - 1]);
        double center = 2.0 * left;
        double rhs = 3.0 * (y[f] - y[f - 1]) * left * left;
        k[f] = (rhs * c_prev - rhs_prev * left) / (center * c_prev - r_prev * left);
    }
    for (int64_t idx = f - 1; idx >= 1; idx--) k[idx] = (rh[idx] - rc[idx] * k[idx + 1]) / cc[idx];
    k[0] = (rhs0 - right0 * k[1]) / center0;
    free(cc); free(rc); free(rh);
}

/* impl/hdp_math_utils.c:498-510 */
void or_linspace(double start, double stop, int64_t length, double *out) {
    int64_t n = length - 1;
    double dx = (stop - start) / ((double)n);
    for (int64_t i = 0; i < n; i++) out[i] = start + i * dx;
    out[n] = stop;
}

/* impl/hdp.c:2772-2784 */
double or_hdp_expected_value(const double *grid, const double *distr, int64_t length) {
    double e = 0.0;
    for (int64_t i = 1; i < length; i++) e += grid[i] * distr[i] * (grid[i] - grid[i - 1]);
    return e;
}

static int64_t model_kmer_id(const or_model *m, const char *kmer) {
    return or_kmer_id(kmer, m->alphabet, m->alphabet_size, m->kmer_length);
}

/* Match / extra-event emission for the k-mer starting at `kmer` (NULL for the
 * x = 0 column) and event record `event` (mean, noise, ...).
 *   OR_GAUSS_DESCALED  impl/stateMachine.c:498-545
 *   OR_GAUSS           impl/stateMachine.c:547-589
 *   OR_HDP             impl/stateMachine.c:468-495 -> impl/nanopore_hdp.c:403-412
 *                      -> impl/hdp.c:2583-2607 (walk to observed ancestor, clamp >= 0) */
double or_emission(const or_model *m, const char *kmer, const double *event, int match) {
    if (kmer == NULL) return OR_LOG_ZERO;
    int64_t kid = model_kmer_id(m, kmer);
    if (kid < 0) return NAN;
    double event_mean = event[0];
    if (m->kind == OR_HDP) {
        double level_mean = kid > OR_NUM_OF_KMERS ? 0.0 : m->match[kid * 5];
        double u = (event_mean + m->var * level_mean - m->scale * level_mean - m->shift) / m->var;
        int64_t row = m->kmer_row[kid];
        const double *y = m->density + row * m->grid_length;
        const double *s = m->slopes + row * m->grid_length;
        double interp = or_grid_spline_interp(u, m->grid, y, s, m->grid_length);
        double dens = interp > 0.0 ? interp : 0.0;
        double density = (1 / m->var) * dens;
        return log(density);
    }
    double event_noise = event[1];
    const double *tab = match ? m->match : m->gapy;
    double level_mean = kid > OR_NUM_OF_KMERS ? 0.0 : tab[kid * 5 + 0];
    double level_sd = kid > OR_NUM_OF_KMERS ? 0.0 : tab[kid * 5 + 1];
    double noise_mean = kid > OR_NUM_OF_KMERS ? 0.0 : tab[kid * 5 + 2];
    double noise_lambda = kid > OR_NUM_OF_KMERS ? 0.0 : tab[kid * 5 + 4];
    if (m->kind == OR_GAUSS_DESCALED)
        event_mean = (event_mean + m->var * level_mean - m->scale * level_mean - m->shift) / m->var;
    double l_mean = or_log_gauss_pdf(event_mean, level_mean, level_sd);
    double l_noise = or_log_inv_gauss_pdf(event_noise, noise_mean, noise_lambda);
    return l_mean + l_noise;
}

/* gapX (k-mer skip) emission.  Gaussian machine: emissions_kmer_getGapProb
 * (impl/stateMachine.c:155-172) over EMISSION_GAP_X_PROBS; HDP machine: the
 * constant log(0.1) (impl/stateMachine.c:1282). */
static double gapx_emission(const or_model *m, const char *kmer) {
    if (m->kind == OR_HDP) return -2.3025850929940455;
    if (kmer == NULL) return OR_LOG_ZERO;
    int64_t kid = model_kmer_id(m, kmer);
    return kid > OR_NUM_OF_KMERS ? OR_LOG_ZERO : m->gapx[kid];
}

/* ------------------------------------------------------------------------- */
/* Band geometry: impl/pairwiseAligner.c:92-182                               */
/* ------------------------------------------------------------------------- */

static int64_t band_avoid_off_by_one(int64_t xay, int64_t xmy) {
    return (xay + xmy) % 2 == 0 ? xmy : xmy + 1;
}

static void band_set_p(int64_t *xmy, int64_t i, int64_t j, int64_t k) {
    if (i < j) *xmy += (int64_t)(2 * (j - i) * k);
}

static int64_t x_coord(int64_t xay, int64_t xmy) { return (xay + xmy) / 2; }
static int64_t y_coord(int64_t xay, int64_t xmy) { return (xay - xmy) / 2; }

static int band_set_current(int64_t xay, int64_t xL, int64_t yL, int64_t xU, int64_t yU,
                            int64_t *outL, int64_t *outR) {
    int64_t xmyL = xL - yL;
    int64_t xmyR = xU - yU;
    if (!(xay >= xL + yU) || !(xay <= xU + yL)) return -1;
    xmyL = band_avoid_off_by_one(xay, xmyL);
    xmyR = band_avoid_off_by_one(xay, xmyR);
    band_set_p(&xmyL, x_coord(xay, xmyL), xL, 1);
    band_set_p(&xmyL, yL, y_coord(xay, xmyL), 1);
    band_set_p(&xmyR, xU, x_coord(xay, xmyR), -1);
    band_set_p(&xmyR, y_coord(xay, xmyR), yU, -1);
    if ((xay + xmyL) % 2 != 0 || (xay + xmyR) % 2 != 0 || xmyL > xmyR) return -1;
    *outL = xmyL;
    *outR = xmyR;
    return 0;
}

static int64_t band_bound(int64_t z, int64_t lZ) { return z < 0 ? 0 : (z > lZ ? lZ : z); }

int or_band(const int64_t *anchors, int64_t n_anchors, int64_t lX, int64_t lY, int64_t expansion,
            int64_t *xmyL, int64_t *xmyR) {
    if (lX < 0 || lY < 0 || expansion % 2 != 0) return -1;
    int64_t lXalY = lX + lY;
    int64_t ai = 0, xay = 0, pxay = 0, pxmy = 0, nxay = 0, nxmy = 0;
    int64_t xL = 0, yL = 0, xU = 0, yU = 0;
    while (xay <= lXalY) {
        if (band_set_current(xay, xL, yL, xU, yU, &xmyL[xay], &xmyR[xay]) != 0) return -2;
        if (nxay == xay++) {
            pxay = nxay;
            pxmy = nxmy;
            int64_t x = lX, y = lY;
            if (ai < n_anchors) {
                x = anchors[2 * ai] + 1;
                y = anchors[2 * ai + 1] + 1;
                ai++;
                if (!(x > x_coord(pxay, pxmy)) || !(y > y_coord(pxay, pxmy)) || x > lX || y > lY ||
                    x <= 0 || y <= 0)
                    return -3;
            }
            nxay = x + y;
            nxmy = x - y;
            xL = band_bound(x_coord(pxay, pxmy - expansion), lX);
            yL = band_bound(y_coord(nxay, nxmy - expansion), lY);
            xU = band_bound(x_coord(nxay, nxmy + expansion), lX);
            yU = band_bound(y_coord(pxay, pxmy + expansion), lY);
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Split points: impl/pairwiseAligner.c:1706-1757                             */
/* ------------------------------------------------------------------------- */

static int split_p(int64_t *x1, int64_t *y1, int64_t x2, int64_t y2, int64_t x3, int64_t y3,
                   int64_t *out, int64_t cap, int64_t *n, int64_t split_matrix, int skip_block) {
    int64_t lX2 = x3 - x2;
    int64_t lY2 = y3 - y2;
    int64_t matrix_size = lX2 * lY2;
    if (matrix_size > split_matrix) {
        int64_t max_len = (int64_t)sqrt((double)split_matrix);
        int64_t hX = lX2 / 2 > max_len ? max_len : lX2 / 2;
        int64_t hY = lY2 / 2 > max_len ? max_len : lY2 / 2;
        if (!skip_block) {
            if (*n < cap) {
                out[4 * *n + 0] = *x1; out[4 * *n + 1] = *y1;
                out[4 * *n + 2] = x2 + hX; out[4 * *n + 3] = y2 + hY;
            }
            (*n)++;
        }
        *x1 = x3 - hX;
        *y1 = y3 - hY;
        return 1;
    }
    return 0;
}

int64_t or_split_points(const int64_t *anchors, int64_t n_anchors, int64_t lX, int64_t lY,
                        int64_t split_matrix, int ragged_left, int ragged_right,
                        int64_t *out, int64_t cap) {
    int64_t x1 = 0, y1 = 0, x2 = 0, y2 = 0, n = 0;
    for (int64_t i = 0; i < n_anchors; i++) {
        int64_t x3 = anchors[2 * i], y3 = anchors[2 * i + 1];
        split_p(&x1, &y1, x2, y2, x3, y3, out, cap, &n, split_matrix, ragged_left && i == 0);
        if (x3 < x2 || y3 < y2 || x3 >= lX || y3 >= lY) return -1;
        x2 = x3 + 1;
        y2 = y3 + 1;
    }
    if (!split_p(&x1, &y1, x2, y2, lX, lY, out, cap, &n, split_matrix, ragged_left && n_anchors == 0) ||
        !ragged_right) {
        if (n < cap) {
            out[4 * n + 0] = x1; out[4 * n + 1] = y1; out[4 * n + 2] = lX; out[4 * n + 3] = lY;
        }
        n++;
    }
    return n;
}

/* ------------------------------------------------------------------------- */
/* DP diagonals and matrices: impl/pairwiseAligner.c:837-1106                 */
/* ------------------------------------------------------------------------- */

typedef struct {
    int64_t xay, xmyL, xmyR;
    const int64_t *pp; /* first path entry of column x; column x has pp[x+1] - pp[x] paths */
    double *cells;     /* 3 per path entry; column x from 3 * (pp[x] - pp[x of xmyL]) */
} or_diag;

typedef struct {
    or_diag **diags;
    int64_t n;       /* diagonalNumber: valid xay in [0, n] */
    int64_t active;
    const int64_t *pp;
} or_matrix;

static int64_t diag_entries(const or_diag *d) {
    return d->pp[x_coord(d->xay, d->xmyR) + 1] - d->pp[x_coord(d->xay, d->xmyL)];
}

static or_matrix *matrix_new(int64_t n, const int64_t *pp) {
    or_matrix *m = (or_matrix *)calloc(1, sizeof(or_matrix));
    m->diags = (or_diag **)calloc((size_t)(n + 1), sizeof(or_diag *));
    m->n = n;
    m->pp = pp;
    return m;
}

static void matrix_free(or_matrix *m) {
    for (int64_t i = 0; i <= m->n; i++)
        if (m->diags[i]) { free(m->diags[i]->cells); free(m->diags[i]); }
    free(m->diags);
    free(m);
}

static or_diag *matrix_get(const or_matrix *m, int64_t xay) {
    if (xay < 0 || xay > m->n) return NULL;
    return m->diags[xay];
}

/* dpMatrix_createDiagonal + dpDiagonal_zeroValues */
static or_diag *matrix_create(or_matrix *m, int64_t xay, int64_t xmyL, int64_t xmyR) {
    or_diag *d = (or_diag *)malloc(sizeof(or_diag));
    d->xay = xay; d->xmyL = xmyL; d->xmyR = xmyR;
    d->pp = m->pp;
    int64_t w = diag_entries(d);
    d->cells = (double *)malloc(sizeof(double) * 3 * (size_t)w);
    for (int64_t i = 0; i < 3 * w; i++) d->cells[i] = OR_LOG_ZERO;
    m->diags[xay] = d;
    m->active++;
    return d;
}

static void matrix_delete(or_matrix *m, int64_t xay) {
    if (xay < 0 || xay > m->n) return;
    if (m->diags[xay]) {
        m->active--;
        free(m->diags[xay]->cells);
        free(m->diags[xay]);
        m->diags[xay] = NULL;
    }
}

/* dpDiagonal_getCell: NULL outside [xmyL, xmyR]; the cell's first path entry */
static double *diag_cell(const or_diag *d, int64_t xmy) {
    if (d == NULL || xmy < d->xmyL || xmy > d->xmyR) return NULL;
    return d->cells + 3 * (d->pp[x_coord(d->xay, xmy)] - d->pp[x_coord(d->xay, d->xmyL)]);
}

/* stateMachine3_startStateProb / raggedStart / endStateProb / raggedEnd
 * (impl/stateMachine.c:1022-1061) */
static void state_values(const or_model *m, int which, double out[3]) {
    const double *t = m->trans;
    switch (which) {
    case 0: out[0] = 0; out[1] = OR_LOG_ZERO; out[2] = OR_LOG_ZERO; break;          /* start */
    case 1: out[0] = OR_LOG_ZERO; out[1] = 0; out[2] = 0; break;                    /* ragged start */
    case 2: out[0] = t[OR_T_MATCH_CONTINUE]; out[1] = t[OR_T_MATCH_FROM_GAP_X];
            out[2] = t[OR_T_MATCH_FROM_GAP_Y]; break;                               /* end */
    default: out[0] = (t[OR_T_GAP_OPEN_X] + t[OR_T_GAP_OPEN_Y]) / 2.0;
             out[1] = t[OR_T_GAP_EXTEND_X]; out[2] = t[OR_T_GAP_EXTEND_Y]; break;    /* ragged end */
    }
}

static void diag_init_values(or_diag *d, const double v[3]) {
    int64_t w = diag_entries(d);
    for (int64_t i = 0; i < w; i++)
        for (int s = 0; s < 3; s++) d->cells[3 * i + s] = v[s];
}

/* doTransitionForward / doTransitionBackward (impl/pairwiseAligner.c:701-726) */
static inline void do_transition(int backward, double *from, double *to, int f, int t, double eP, double tP) {
    if (!backward) to[t] = or_logadd(to[t], from[f] + (eP + tP));
    else from[f] = or_logadd(from[f], to[t] + (eP + tP));
}

static const double k_null_event[2] = {-INFINITY, 0};

typedef struct {
    const or_model *m;
    const char *ref;   /* k-mer i = ref[i .. i+k) */
    int64_t lX;
    const double *ev;  /* 4 doubles per event */
    int64_t lY;
    const int64_t *pp; /* lX + 2 entries: first path entry of column x (column x, x >= 1, is k-mer x-1) */
    const char *pk;    /* kmer_length characters per path entry: the substituted k-mer */
} or_seqs;

static int64_t n_paths(const or_seqs *s, int64_t x) { return s->pp[x + 1] - s->pp[x]; }

/* Path->kmer: NULL in column 0 (no k-mer) */
static const char *path_kmer(const or_seqs *s, int64_t x, int64_t p) {
    return x >= 1 ? s->pk + (s->pp[x] + p) * s->m->kmer_length : NULL;
}

/* path_checkLegal (impl/pairwiseAligner.c:540-552): consecutive k-mers agree on
 * their overlap; a NULL k-mer (column 0) joins anything */
static int path_legal(const char *from, const char *to, int k) {
    if (from == NULL || to == NULL) return 1;
    for (int i = 0; i < k - 1; i++)
        if (from[i + 1] != to[i]) return 0;
    return 1;
}

/* stString_eq on path k-mers (two NULL k-mers are equal) */
static int kmer_eq(const char *a, const char *b, int k) {
    if (a == NULL || b == NULL) return a == b;
    return memcmp(a, b, (size_t)k) == 0;
}

/* stateMachine3_cellCalculate (impl/stateMachine.c:1194-1257) and the HDP variant
 * (:1259-1325) over the cells' paths: lower and middle transitions between legal
 * path pairs, upper transitions within the same path, each block looping over the
 * current cell's paths and then the neighbour's, as the reference does. */
static void cell_calculate(const or_seqs *s, int64_t x, double *cur, double *lower, double *middle,
                           double *upper, const double *event, int backward) {
    const or_model *m = s->m;
    const double *t = m->trans;
    const int k = m->kmer_length;
    const int64_t pc = n_paths(s, x), pn = x >= 1 ? n_paths(s, x - 1) : 0;
    if (lower) {
        for (int64_t p = 0; p < pc; p++) {
            const char *kc = path_kmer(s, x, p);
            double eP = gapx_emission(m, kc);
            for (int64_t q = 0; q < pn; q++) {
                if (!path_legal(path_kmer(s, x - 1, q), kc, k)) continue;
                double *lo = lower + 3 * q, *c = cur + 3 * p;
                do_transition(backward, lo, c, M_ST, X_ST, eP, t[OR_T_GAP_OPEN_X]);
                do_transition(backward, lo, c, X_ST, X_ST, eP, t[OR_T_GAP_EXTEND_X]);
                do_transition(backward, lo, c, Y_ST, X_ST, eP, t[OR_T_GAP_SWITCH_TO_X]);
            }
        }
    }
    if (middle) {
        for (int64_t p = 0; p < pc; p++) {
            const char *kc = path_kmer(s, x, p);
            double eP = or_emission(m, kc, event, 1);
            for (int64_t q = 0; q < pn; q++) {
                if (!path_legal(path_kmer(s, x - 1, q), kc, k)) continue;
                double *mi = middle + 3 * q, *c = cur + 3 * p;
                do_transition(backward, mi, c, M_ST, M_ST, eP, t[OR_T_MATCH_CONTINUE]);
                do_transition(backward, mi, c, X_ST, M_ST, eP, t[OR_T_MATCH_FROM_GAP_X]);
                do_transition(backward, mi, c, Y_ST, M_ST, eP, t[OR_T_MATCH_FROM_GAP_Y]);
            }
        }
    }
    if (upper) {
        for (int64_t p = 0; p < pc; p++) {
            const char *kc = path_kmer(s, x, p);
            double eP = or_emission(m, kc, event, 0);
            for (int64_t q = 0; q < pc; q++) {
                if (!kmer_eq(kc, path_kmer(s, x, q), k)) continue;
                double *up = upper + 3 * q, *c = cur + 3 * p;
                do_transition(backward, up, c, M_ST, Y_ST, eP, t[OR_T_GAP_OPEN_Y]);
                do_transition(backward, up, c, Y_ST, Y_ST, eP, t[OR_T_GAP_EXTEND_Y]);
            }
        }
    }
}

/* diagonalCalculation (impl/pairwiseAligner.c:1127-1158): cells in ascending xmy,
 * lower = (d-1, xmy-1), middle = (d-2, xmy), upper = (d-1, xmy+1). */
static void diagonal_calculation(const or_seqs *s, or_diag *cur, const or_diag *m1, const or_diag *m2,
                                 int backward) {
    for (int64_t xmy = cur->xmyL; xmy <= cur->xmyR; xmy += 2) {
        int64_t x = x_coord(cur->xay, xmy);
        int64_t y = y_coord(cur->xay, xmy);
        const double *event = (y - 1 >= 0) ? s->ev + 4 * (y - 1) : k_null_event;
        double *c = diag_cell(cur, xmy);
        double *lo = diag_cell(m1, xmy - 1);
        double *mi = diag_cell(m2, xmy);
        double *up = diag_cell(m1, xmy + 1);
        cell_calculate(s, x, c, lo, mi, up, event, backward);
    }
}

/* cell_dotProduct + hdCell_totalProbability + dpDiagonal_dotProduct
 * (impl/pairwiseAligner.c:728-734, :677-692, :1014-1027): per cell the paths with
 * equal k-mers, per path the three states */
static double diag_dot(const or_seqs *s, const or_diag *a, const or_diag *b) {
    const int k = s->m->kmer_length;
    double total = OR_LOG_ZERO;
    for (int64_t xmy = a->xmyL; xmy <= a->xmyR; xmy += 2) {
        const int64_t x = x_coord(a->xay, xmy), np = n_paths(s, x);
        const double *c1 = diag_cell(a, xmy), *c2 = diag_cell(b, xmy);
        double hd = OR_LOG_ZERO;
        for (int64_t p = 0; p < np; p++)
            for (int64_t q = 0; q < np; q++) {
                if (!kmer_eq(path_kmer(s, x, p), path_kmer(s, x, q), k)) continue;
                const double *u = c1 + 3 * p, *v = c2 + 3 * q;
                double cell = u[0] + v[0];
                cell = or_logadd(cell, u[1] + v[1]);
                cell = or_logadd(cell, u[2] + v[2]);
                hd = or_logadd(hd, cell);
            }
        total = or_logadd(total, hd);
    }
    return total;
}

/* diagonalCalculationTotalProbability (impl/pairwiseAligner.c:1182-1200) */
static double total_probability(const or_seqs *s, int64_t xay, const or_matrix *fw, const or_matrix *bw) {
    double total = diag_dot(s, matrix_get(fw, xay), matrix_get(bw, xay));
    const or_diag *f1 = matrix_get(fw, xay - 1);
    const or_diag *b1 = matrix_get(bw, xay + 1);
    if (b1 != NULL && f1 != NULL) {
        or_diag md;
        md.xay = b1->xay; md.xmyL = b1->xmyL; md.xmyR = b1->xmyR;
        md.pp = b1->pp;
        int64_t w = diag_entries(b1);
        md.cells = (double *)malloc(sizeof(double) * 3 * (size_t)w);
        for (int64_t i = 0; i < 3 * w; i++) md.cells[i] = OR_LOG_ZERO;
        diagonal_calculation(s, &md, NULL, f1, 0);
        total = or_logadd(total, diag_dot(s, &md, b1));
        free(md.cells);
    }
    return total;
}

/* hdCell_construct (impl/pairwiseAligner.c:599-656) for every column: column 0 has
 * one path with no k-mer; column x >= 1 one path per combination of the degenerate
 * base options at the 'X' positions of k-mer x-1, the first 'X' varying slowest
 * (path_listPotentialKmers, :569-585; hdCell_getSubstitutedKmer, :667-676).
 * Returns 0, or -30 for an 'X' without options. */
static int build_paths(const char *ref, int64_t lX, int k, const char *options, int64_t **pp_out, char **pk_out) {
    const int64_t nopt = options ? (int64_t)strlen(options) : 0;
    int64_t *pp = (int64_t *)malloc(sizeof(int64_t) * (size_t)(lX + 2));
    pp[0] = 0;
    pp[1] = 1;
    for (int64_t x = 1; x <= lX; x++) {
        int64_t n = 0;
        for (int j = 0; j < k; j++) n += ref[x - 1 + j] == 'X';
        if (n > 0 && nopt == 0) { free(pp); return -30; }
        int64_t np = 1;
        for (int64_t j = 0; j < n; j++) np *= nopt;
        pp[x + 1] = pp[x] + np;
    }
    char *pk = (char *)malloc((size_t)(pp[lX + 1] + 1) * (size_t)k);
    for (int64_t x = 1; x <= lX; x++) {
        for (int64_t p = 0; p < pp[x + 1] - pp[x]; p++) {
            char *dst = pk + (pp[x] + p) * k;
            memcpy(dst, ref + x - 1, (size_t)k);
            int64_t rem = p;
            for (int j = k - 1; j >= 0; j--)
                if (dst[j] == 'X') {
                    dst[j] = options[rem % nopt];
                    rem /= nopt;
                }
        }
    }
    *pp_out = pp;
    *pk_out = pk;
    return 0;
}

int64_t or_path_kmers(const char *kmer, int k, const char *options, char *out, int64_t cap) {
    int64_t *pp = NULL;
    char *pk = NULL;
    if (build_paths(kmer, 1, k, options, &pp, &pk) != 0) return -1;
    const int64_t n = pp[2] - pp[1];
    if (out && n <= cap) memcpy(out, pk + pp[1] * k, (size_t)(n * k));
    free(pp);
    free(pk);
    return n;
}

int or_path_legal(const char *from, const char *to, int k) { return path_legal(from, to, k); }

typedef struct {
    or_pair *buf;
    int64_t n, cap;
} or_pairlist;

static void pairlist_push(or_pairlist *l, int64_t p, int64_t x, int64_t y, int64_t path) {
    if (l->n >= l->cap) {
        int64_t nc = l->cap ? 2 * l->cap : 1024;
        l->buf = (or_pair *)realloc(l->buf, sizeof(or_pair) * (size_t)nc);
        l->cap = nc;
    }
    l->buf[l->n].p = p; l->buf[l->n].x = x; l->buf[l->n].y = y; l->buf[l->n].path = path;
    l->n++;
}

/* diagonalCalculationPosteriorMatchProbs (impl/pairwiseAligner.c:1202-1253) */
static void posterior_match_probs(const or_seqs *s, int64_t xay, const or_matrix *fw, const or_matrix *bw,
                                  double total, double threshold, or_pairlist *out) {
    const int k = s->m->kmer_length;
    const or_diag *fd = matrix_get(fw, xay), *bd = matrix_get(bw, xay);
    for (int64_t xmy = fd->xmyL; xmy <= fd->xmyR; xmy += 2) {
        int64_t x = x_coord(xay, xmy), y = y_coord(xay, xmy);
        if (x > 0 && y > 0) {
            const double *cf = diag_cell(fd, xmy), *cb = diag_cell(bd, xmy);
            const int64_t np = n_paths(s, x);
            for (int64_t fp = 0; fp < np; fp++)
                for (int64_t bp = 0; bp < np; bp++) {
                    if (!kmer_eq(path_kmer(s, x, fp), path_kmer(s, x, bp), k)) continue;
                    double p = exp((cf[3 * fp + M_ST] + cb[3 * bp + M_ST]) - total);
                    if (p >= threshold) {
                        if (p > 1.0) p = 1.0;
                        p = floor(p * OR_PAIR_PROB_1);
                        pairlist_push(out, (int64_t)p, x - 1, y - 1, fp);
                    }
                }
        }
    }
}

/* cell_signal_updateExpectations (impl/pairwiseAligner.c:763-790): the posterior
 * of one transition, added to the transition expectations; for transitions into
 * the match state also to the k-mer's event expectations
 * (continuousPairHmm_addToEmissionExpectation, impl/continuousHmm.c:160-169). */
static void expect_transition(or_expect *ex, const double *from, const double *to, int f, int t, double eP,
                              double tP, double total, int64_t kidx, double xd) {
    double p = exp(from[f] + to[t] + (eP + tP) - total);
    ex->transitions[f * 3 + t] += p;
    if (t == M_ST) {
        ex->event_exp[2 * kidx] += p * xd;
        ex->posteriors[kidx] += p;
        double u = ex->event_exp[2 * kidx] / ex->posteriors[kidx];
        ex->event_exp[2 * kidx + 1] += p * (xd - u) * (xd - u);
        ex->observed[kidx] = 1;
    }
}

/* diagonalCalculation_Expectations (impl/pairwiseAligner.c:1255-1275): the cells
 * of the backward diagonal xay against the forward diagonals xay-1 and xay-2
 * (NULL once deleted: the first posterior diagonal of every window after the
 * first sees no xay-2), in stateMachine3_cellCalculate's transition order. */
static void expectations_diagonal(const or_seqs *s, int64_t xay, const or_matrix *fw, const or_matrix *bw,
                                  double total, or_expect *ex) {
    const or_model *m = s->m;
    const double *t = m->trans;
    ex->likelihood += total;
    const or_diag *cur = matrix_get(bw, xay);
    const or_diag *m1 = xay >= 1 ? matrix_get(fw, xay - 1) : NULL;
    const or_diag *m2 = xay >= 2 ? matrix_get(fw, xay - 2) : NULL;
    for (int64_t xmy = cur->xmyL; xmy <= cur->xmyR; xmy += 2) {
        int64_t x = x_coord(xay, xmy), y = y_coord(xay, xmy);
        const char *kmer = path_kmer(s, x, 0); /* single-path cells (the driver checks) */
        const double *event = (y - 1 >= 0) ? s->ev + 4 * (y - 1) : k_null_event;
        double *c = diag_cell(cur, xmy);
        double *lo = diag_cell(m1, xmy - 1);
        double *mi = diag_cell(m2, xmy);
        double *up = diag_cell(m1, xmy + 1);
        if (lo) {
            double eP = gapx_emission(m, kmer);
            expect_transition(ex, lo, c, M_ST, X_ST, eP, t[OR_T_GAP_OPEN_X], total, 0, 0);
            expect_transition(ex, lo, c, X_ST, X_ST, eP, t[OR_T_GAP_EXTEND_X], total, 0, 0);
            expect_transition(ex, lo, c, Y_ST, X_ST, eP, t[OR_T_GAP_SWITCH_TO_X], total, 0, 0);
        }
        if (mi) {
            /* a middle neighbour means x >= 1 and y >= 1 */
            double eP = or_emission(m, kmer, event, 1);
            int64_t kidx = or_kmer_id(kmer, m->alphabet, m->alphabet_size, m->kmer_length);
            double mu = m->match[5 * kidx];
            /* emissions_signal_descaleEventMean_JordanStyle (impl/stateMachine.c:285-289) */
            double xd = (event[0] + m->var * mu - m->scale * mu - m->shift) / m->var;
            double pr[3];
            for (int f = 0; f < 3; f++) pr[f] = exp(mi[f] + c[M_ST] + (eP + t[f == 0 ? OR_T_MATCH_CONTINUE
                                                                         : f == 1 ? OR_T_MATCH_FROM_GAP_X
                                                                                  : OR_T_MATCH_FROM_GAP_Y]) - total);
            expect_transition(ex, mi, c, M_ST, M_ST, eP, t[OR_T_MATCH_CONTINUE], total, kidx, xd);
            expect_transition(ex, mi, c, X_ST, M_ST, eP, t[OR_T_MATCH_FROM_GAP_X], total, kidx, xd);
            expect_transition(ex, mi, c, Y_ST, M_ST, eP, t[OR_T_MATCH_FROM_GAP_Y], total, kidx, xd);
            double mx = pr[0] > pr[1] ? pr[0] : pr[1];
            mx = mx > pr[2] ? mx : pr[2];
            if (ex->rec_cap > 0 && mx >= ex->rec_threshold && mx > 0.0) {
                if (ex->n_rec < ex->rec_cap) {
                    double *r = ex->rec + 5 * ex->n_rec;
                    r[0] = (double)(x - 1 + ex->x1);
                    r[1] = (double)(y - 1 + ex->y1);
                    r[2] = pr[0]; r[3] = pr[1]; r[4] = pr[2];
                }
                ex->n_rec++;
            }
        }
        if (up) {
            double eP = or_emission(m, kmer, event, 0);
            expect_transition(ex, up, c, M_ST, Y_ST, eP, t[OR_T_GAP_OPEN_Y], total, 0, 0);
            expect_transition(ex, up, c, Y_ST, Y_ST, eP, t[OR_T_GAP_EXTEND_Y], total, 0, 0);
        }
    }
}

/* getPosteriorProbsWithBanding (impl/pairwiseAligner.c:1282-1420); with `ex` the
 * per-diagonal callback is diagonalCalculation_Expectations instead of
 * diagonalCalculationPosteriorMatchProbs. */
static int posterior_probs_with_banding(const or_seqs *s, const int64_t *anchors, int64_t n_anchors,
                                        const or_params *p, int ragged_left, int ragged_right,
                                        or_pairlist *out, or_stats *stats, or_expect *ex) {
    if (!(p->traceback_diagonals >= 1) || !(p->diagonal_expansion >= 0) || p->diagonal_expansion % 2 != 0 ||
        !(p->min_diags_between_traceback >= 2) ||
        !(p->traceback_diagonals + 1 < p->min_diags_between_traceback))
        return -10;
    int64_t N = s->lX + s->lY;
    if (N == 0) return 0;
    int64_t *bL = (int64_t *)malloc(sizeof(int64_t) * (size_t)(N + 1));
    int64_t *bR = (int64_t *)malloc(sizeof(int64_t) * (size_t)(N + 1));
    int rc = or_band(anchors, n_anchors, s->lX, s->lY, p->diagonal_expansion, bL, bR);
    if (rc != 0) { free(bL); free(bR); return -11; }
    for (int64_t d = 0; d <= N; d++) stats->cells += (bR[d] - bL[d]) / 2 + 1;

    or_matrix *fw = matrix_new(N, s->pp), *bw = matrix_new(N, s->pp);
    double v[3];
    state_values(s->m, ragged_left ? 1 : 0, v);
    diag_init_values(matrix_create(fw, 0, bL[0], bR[0]), v);
    int64_t it = 1; /* band iterator index: next diagonal to hand out */
    int64_t traced_back_to = 0, total_posterior = 0;
    for (;;) {
        int64_t d = it <= N ? it : N;
        if (it <= N) it++;
        or_diag *cur = matrix_create(fw, d, bL[d], bR[d]);
        diagonal_calculation(s, cur, matrix_get(fw, d - 1), matrix_get(fw, d - 2), 0);
        int at_end = d == N;
        int traceback_point = d >= traced_back_to + p->min_diags_between_traceback &&
                              (bR[d] - bL[d]) / 2 + 1 <= p->diagonal_expansion * 2 + 1;
        if (at_end || traceback_point) {
            state_values(s->m, (at_end && ragged_right) ? 3 : 2, v);
            diag_init_values(matrix_create(bw, d, bL[d], bR[d]), v);
            if (d > traced_back_to + 1) matrix_create(bw, d - 1, bL[d - 1], bR[d - 1]);
            int64_t traced_back_from = d - (at_end ? 0 : p->traceback_diagonals + 1);
            if (g_win_buf) {
                if (g_win_n < g_win_cap) {
                    g_win_buf[4 * g_win_n + 0] = traced_back_to;
                    g_win_buf[4 * g_win_n + 1] = d;
                    g_win_buf[4 * g_win_n + 2] = traced_back_from;
                    g_win_buf[4 * g_win_n + 3] = at_end;
                }
                g_win_n++;
            }
            double total = OR_LOG_ZERO;
            int64_t this_tb = 0;
            int64_t d2 = d;
            while (d2 > traced_back_to) {
                if (d2 > traced_back_to + 2) matrix_create(bw, d2 - 2, bL[d2 - 2], bR[d2 - 2]);
                if (d2 > traced_back_to + 1)
                    diagonal_calculation(s, matrix_get(bw, d2), matrix_get(bw, d2 - 1), matrix_get(bw, d2 - 2), 1);
                if (d2 <= traced_back_from) {
                    if (this_tb++ % 10 == 0) {
                        total = total_probability(s, d2, fw, bw);
                        if (d2 == N) stats->log_likelihood += total; /* summed over sub-alignments */
                    }
                    if (ex) expectations_diagonal(s, d2, fw, bw, total, ex);
                    else posterior_match_probs(s, d2, fw, bw, total, p->threshold, out);
                    if (d2 < traced_back_from || at_end) matrix_delete(fw, d2);
                }
                if (d2 + 1 <= N) matrix_delete(bw, d2 + 1);
                d2--;
            }
            traced_back_to = traced_back_from;
            matrix_delete(bw, d2 + 1);
            matrix_delete(fw, d2);
            total_posterior += this_tb;
            stats->n_windows++;
            if (bw->active != 0 || (!at_end && fw->active != p->traceback_diagonals + 2)) {
                matrix_free(fw); matrix_free(bw); free(bL); free(bR);
                return -12;
            }
        }
        if (at_end) break;
    }
    int ok = total_posterior == N && traced_back_to == N && bw->active == 0 && fw->active == 0;
    matrix_free(fw);
    matrix_free(bw);
    free(bL);
    free(bR);
    return ok ? 0 : -13;
}

/* getAlignedPairsUsingAnchors (impl/pairwiseAligner.c:1871-1899) via
 * getPosteriorProbsWithBandingSplittingAlignmentsByLargeGaps (:1773-1837) and
 * alignedPairCoordinateCorrectionFn (:1862-1869): each sub-alignment's pairs are
 * shifted by (x1, y1) and moved to the output by popping from the end, so the
 * final list holds every sub-alignment's pairs in reverse emission order. */
int or_get_aligned_pairs(const or_model *m, const char *ref, int64_t lX,
                         const double *events, int64_t lY,
                         const int64_t *anchors, int64_t n_anchors,
                         const or_params *p, int ragged_left, int ragged_right,
                         or_pair *out, int64_t cap, int64_t *n_out, or_stats *stats) {
    return or_get_aligned_pairs_degenerate(m, ref, lX, events, lY, anchors, n_anchors, p, ragged_left, ragged_right,
                                           NULL, out, cap, n_out, stats);
}

int or_get_aligned_pairs_degenerate(const or_model *m, const char *ref, int64_t lX,
                                    const double *events, int64_t lY,
                                    const int64_t *anchors, int64_t n_anchors,
                                    const or_params *p, int ragged_left, int ragged_right, const char *options,
                                    or_pair *out, int64_t cap, int64_t *n_out, or_stats *stats) {
    or_stats st = {0, 0, 0.0};
    int64_t nsplit = or_split_points(anchors, n_anchors, lX, lY, p->split_matrix_bigger_than_this,
                                     ragged_left, ragged_right, NULL, 0);
    if (nsplit < 0) return -1;
    int64_t *splits = (int64_t *)malloc(sizeof(int64_t) * 4 * (size_t)(nsplit + 1));
    or_split_points(anchors, n_anchors, lX, lY, p->split_matrix_bigger_than_this, ragged_left, ragged_right,
                    splits, nsplit);
    int64_t *sub = (int64_t *)malloc(sizeof(int64_t) * 2 * (size_t)(n_anchors + 1));
    or_pairlist pl = {NULL, 0, 0};
    int64_t n = 0, j = 0;
    int rc = 0;
    for (int64_t i = 0; i < nsplit && rc == 0; i++) {
        int64_t x1 = splits[4 * i], y1 = splits[4 * i + 1], x2 = splits[4 * i + 2], y2 = splits[4 * i + 3];
        int64_t ns = 0;
        while (j < n_anchors) {
            int64_t x = anchors[2 * j], y = anchors[2 * j + 1];
            if (x + y >= x2 + y2) break;
            if (!(x >= x1 && x < x2 && y >= y1 && y < y2)) { rc = -2; break; }
            sub[2 * ns] = x - x1;
            sub[2 * ns + 1] = y - y1;
            ns++;
            j++;
        }
        if (rc) break;
        int64_t *pp = NULL;
        char *pk = NULL;
        rc = build_paths(ref + x1, x2 - x1, m->kmer_length, options, &pp, &pk);
        if (rc) break;
        or_seqs s = {m, ref + x1, x2 - x1, events + 4 * y1, y2 - y1, pp, pk};
        pl.n = 0;
        rc = posterior_probs_with_banding(&s, sub, ns, p, ragged_left || i > 0,
                                          ragged_right || i < nsplit - 1, &pl, &st, NULL);
        free(pp);
        free(pk);
        for (int64_t k = pl.n - 1; k >= 0 && rc == 0; k--) {
            if (n < cap) {
                out[n].p = pl.buf[k].p;
                out[n].x = pl.buf[k].x + x1;
                out[n].y = pl.buf[k].y + y1;
                out[n].path = pl.buf[k].path;
            }
            n++;
        }
    }
    if (rc == 0 && j != n_anchors) rc = -3;
    free(pl.buf);
    free(sub);
    free(splits);
    *n_out = n;
    if (stats) *stats = st;
    return rc;
}

/* getExpectationsUsingAnchors (impl/pairwiseAligner.c:1983-2003) through
 * getPosteriorProbsWithBandingSplittingAlignmentsByLargeGaps (:1773-1837): the
 * sub-alignments in order, each accumulating into the same expectations. */
int or_get_expectations(const or_model *m, const char *ref, int64_t lX, const double *events, int64_t lY,
                        const int64_t *anchors, int64_t n_anchors, const or_params *p, int ragged_left,
                        int ragged_right, or_expect *ex) {
    or_stats st = {0, 0, 0.0};
    int64_t nsplit = or_split_points(anchors, n_anchors, lX, lY, p->split_matrix_bigger_than_this,
                                     ragged_left, ragged_right, NULL, 0);
    if (nsplit < 0) return -1;
    int64_t *splits = (int64_t *)malloc(sizeof(int64_t) * 4 * (size_t)(nsplit + 1));
    or_split_points(anchors, n_anchors, lX, lY, p->split_matrix_bigger_than_this, ragged_left, ragged_right,
                    splits, nsplit);
    int64_t *sub = (int64_t *)malloc(sizeof(int64_t) * 2 * (size_t)(n_anchors + 1));
    int64_t j = 0;
    int rc = 0;
    ex->n_rec = 0;
    for (int64_t i = 0; i < nsplit && rc == 0; i++) {
        int64_t x1 = splits[4 * i], y1 = splits[4 * i + 1], x2 = splits[4 * i + 2], y2 = splits[4 * i + 3];
        int64_t ns = 0;
        while (j < n_anchors) {
            int64_t x = anchors[2 * j], y = anchors[2 * j + 1];
            if (x + y >= x2 + y2) break;
            if (!(x >= x1 && x < x2 && y >= y1 && y < y2)) { rc = -2; break; }
            sub[2 * ns] = x - x1;
            sub[2 * ns + 1] = y - y1;
            ns++;
            j++;
        }
        if (rc) break;
        int64_t *pp = NULL;
        char *pk = NULL;
        rc = build_paths(ref + x1, x2 - x1, m->kmer_length, NULL, &pp, &pk);
        if (rc) break;
        or_seqs s = {m, ref + x1, x2 - x1, events + 4 * y1, y2 - y1, pp, pk};
        ex->x1 = x1;
        ex->y1 = y1;
        rc = posterior_probs_with_banding(&s, sub, ns, p, ragged_left || i > 0, ragged_right || i < nsplit - 1,
                                          NULL, &st, ex);
        free(pp);
        free(pk);
    }
    if (rc == 0 && j != n_anchors) rc = -3;
    free(sub);
    free(splits);
    return rc;
}

/* The reference's tiny full-matrix test procedure (tests/stateMachineTests.c:
 * 565-692): every band diagonal allocated, forward from the start state over all
 * diagonals, backward from the end state over all diagonals, forward and
 * backward totals, each diagonal's total probability, and posterior pairs with a
 * fixed total (the forward total) in ascending diagonal order. */
int or_full_forward_backward(const or_model *m, const char *ref, int64_t lX,
                             const double *events, int64_t lY,
                             const int64_t *anchors, int64_t n_anchors, int64_t expansion,
                             double threshold,
                             or_pair *out, int64_t cap, int64_t *n_out,
                             double *total_forward, double *total_backward,
                             double *diagonal_totals) {
    int64_t N = lX + lY;
    int64_t *bL = (int64_t *)malloc(sizeof(int64_t) * (size_t)(N + 1));
    int64_t *bR = (int64_t *)malloc(sizeof(int64_t) * (size_t)(N + 1));
    if (or_band(anchors, n_anchors, lX, lY, expansion, bL, bR) != 0) { free(bL); free(bR); return -1; }
    int64_t *pp = NULL;
    char *pk = NULL;
    if (build_paths(ref, lX, m->kmer_length, NULL, &pp, &pk) != 0) { free(bL); free(bR); return -1; }
    or_seqs s = {m, ref, lX, events, lY, pp, pk};
    or_matrix *fw = matrix_new(N, pp), *bw = matrix_new(N, pp);
    for (int64_t d = 0; d <= N; d++) {
        matrix_create(fw, d, bL[d], bR[d]);
        matrix_create(bw, d, bL[d], bR[d]);
    }
    double v[3];
    state_values(m, 0, v);
    diag_init_values(matrix_get(fw, 0), v);
    state_values(m, 2, v);
    diag_init_values(matrix_get(bw, N), v);
    for (int64_t d = 1; d <= N; d++)
        diagonal_calculation(&s, matrix_get(fw, d), matrix_get(fw, d - 1), matrix_get(fw, d - 2), 0);
    for (int64_t d = N; d > 0; d--)
        diagonal_calculation(&s, matrix_get(bw, d), matrix_get(bw, d - 1), matrix_get(bw, d - 2), 1);
    /* cell_dotProduct2 of the final forward cell with the end state values */
    double endv[3], startv[3];
    state_values(m, 2, endv);
    state_values(m, 0, startv);
    const double *cf = diag_cell(matrix_get(fw, N), lX - lY);
    double tf = cf[0] + endv[0];
    tf = or_logadd(tf, cf[1] + endv[1]);
    tf = or_logadd(tf, cf[2] + endv[2]);
    tf = or_logadd(OR_LOG_ZERO, tf);
    const double *cb = diag_cell(matrix_get(bw, 0), 0);
    double tb = cb[0] + startv[0];
    tb = or_logadd(tb, cb[1] + startv[1]);
    tb = or_logadd(tb, cb[2] + startv[2]);
    tb = or_logadd(OR_LOG_ZERO, tb);
    *total_forward = tf;
    *total_backward = tb;
    if (diagonal_totals)
        for (int64_t d = 0; d <= N; d++) diagonal_totals[d] = total_probability(&s, d, fw, bw);
    or_pairlist pl = {NULL, 0, 0};
    for (int64_t d = 1; d <= N; d++) posterior_match_probs(&s, d, fw, bw, tf, threshold, &pl);
    for (int64_t k = 0; k < pl.n; k++)
        if (k < cap) out[k] = pl.buf[k];
    *n_out = pl.n;
    free(pl.buf);
    matrix_free(fw);
    matrix_free(bw);
    free(bL);
    free(bR);
    free(pp);
    free(pk);
    return 0;
}
