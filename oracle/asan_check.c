/*
 * asan_check.c -- runs the oracle's entry points on synthetic inputs, for the
 * address- and undefined-behaviour-sanitised build of the CPU-side C code (test
 * infrastructure only; see the `asan` target of oracle/Makefile).
 *
 * Inputs come from a fixed linear congruential generator: a 6-mer ACGT Gaussian
 * table, a reference, events that walk it (stay 0.3, skip 0.05) and anchors every
 * 100 events; an HDP grid for the same k-mers; a reference with degenerate 'X'
 * bases.  Prints one line per entry point with its pair count, windows and
 * log-likelihood; any memory error aborts the run under the sanitisers.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint64_t g_state = 0x9e3779b97f4a7c15ull;
static double uni(void) {
    g_state = g_state * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(g_state >> 11) / 9007199254740992.0;
}
static double normal(void) {
    const double u = uni() + 1e-12, v = uni();
    return sqrt(-2.0 * log(u)) * cos(6.283185307179586 * v);
}

#define K 6
#define NK 4096

int main(void) {
    static double match[5 * NK], gapy[5 * NK], gapx[15625];
    for (int i = 0; i < 15625; i++) gapx[i] = log(0.1);
    for (int i = 0; i < NK; i++) {
        const double mu = 60.0 + 50.0 * uni(), sd = 1.0 + 2.0 * uni(), nmu = 0.8 + 1.2 * uni(), lam = 1.4 + 2.6 * uni();
        double *m = match + 5 * i, *g = gapy + 5 * i;
        m[0] = mu; m[1] = sd; m[2] = nmu; m[3] = sqrt(nmu * nmu * nmu / lam); m[4] = lam;
        memcpy(g, m, sizeof(double) * 5);
        g[1] = sd * 1.75;
    }
    const int nt = 900, lX = nt - K + 1;
    char *ref = malloc((size_t)nt + 1);
    for (int i = 0; i < nt; i++) ref[i] = "ACGT"[(int)(4.0 * uni()) & 3];
    ref[nt] = 0;
    /* events walking the reference */
    int cap = 3 * lX, lY = 0;
    double *ev = malloc(sizeof(double) * 4 * (size_t)cap);
    int *path = malloc(sizeof(int) * (size_t)cap);
    for (int x = 0; x < lX && lY < cap;) {
        const char *km = ref + x;
        const int64_t id = or_kmer_id(km, "ACGT", 4, K);
        const double *m = match + 5 * id;
        ev[4 * lY + 0] = m[0] + m[1] * normal();
        ev[4 * lY + 1] = m[2] * (0.8 + 0.4 * uni());
        ev[4 * lY + 2] = 0.01;
        ev[4 * lY + 3] = 0.01 * lY;
        path[lY++] = x;
        const double u = uni();
        x += u < 0.3 ? 0 : u < 0.35 ? 2 : 1;
    }
    int64_t anchors[2 * 64];
    int na = 0;
    for (int j = 100; j < lY && na < 64; j += 100) {
        if (na > 0 && (path[j] <= anchors[2 * na - 2] || j <= anchors[2 * na - 1])) continue;
        anchors[2 * na] = path[j];
        anchors[2 * na + 1] = j;
        na++;
    }
    or_model gm;
    memset(&gm, 0, sizeof(gm));
    gm.kind = OR_GAUSS_DESCALED;
    gm.kmer_length = K;
    gm.alphabet_size = 4;
    strcpy(gm.alphabet, "ACGT");
    const double tr[9] = {log(0.79), log(0.19), log(0.013), log(0.8), log(0.19), log(0.01), log(0.98), log(1e-9),
                          log(0.013)};
    memcpy(gm.trans, tr, sizeof(tr));
    gm.match = match;
    gm.gapy = gapy;
    gm.gapx = gapx;
    gm.scale = 1.0;
    gm.shift = 0.0;
    gm.var = 1.0;
    or_params p = {0.01, 1000, 40, 20, 3000LL * 3000};
    or_pair *pairs = malloc(sizeof(or_pair) * 16 * (size_t)(lX + lY));
    int64_t n = 0;
    or_stats st;
    memset(&st, 0, sizeof(st));
    int rc = or_get_aligned_pairs(&gm, ref, lX, ev, lY, anchors, na, &p, 1, 1, pairs, 16 * (lX + lY), &n, &st);
    printf("gauss rc %d pairs %lld windows %lld ll %.6f\n", rc, (long long)n, (long long)st.n_windows,
           st.log_likelihood);
    if (rc || n == 0) return 1;
    /* several windows, split blocks */
    or_params p2 = {0.01, 200, 20, 20, 60 * 60};
    memset(&st, 0, sizeof(st));
    rc = or_get_aligned_pairs(&gm, ref, lX, ev, lY, anchors, na, &p2, 0, 0, pairs, 16 * (lX + lY), &n, &st);
    printf("split rc %d pairs %lld windows %lld ll %.6f\n", rc, (long long)n, (long long)st.n_windows,
           st.log_likelihood);
    if (rc || n == 0) return 1;
    /* HDP grid: one row per k-mer, a Gaussian bump tabulated on 300 knots */
    const int64_t G = 300;
    double *grid = malloc(sizeof(double) * (size_t)G);
    double *dens = malloc(sizeof(double) * (size_t)(G * NK)), *slopes = malloc(sizeof(double) * (size_t)(G * NK));
    int64_t *rows = malloc(sizeof(int64_t) * NK);
    static double hm[5 * NK];
    or_linspace(30.0, 130.0, G, grid);
    for (int i = 0; i < NK; i++) {
        const double mu = match[5 * i], sd = match[5 * i + 1];
        for (int64_t g = 0; g < G; g++) {
            const double z = (grid[g] - mu) / sd;
            dens[i * G + g] = exp(-0.5 * z * z) / (sd * 2.5066282746310002);
        }
        or_spline_knot_slopes(grid, dens + i * G, G, slopes + i * G);
        rows[i] = i;
        hm[5 * i] = mu;
    }
    or_model hmod = gm;
    hmod.kind = OR_HDP;
    hmod.match = hm;
    hmod.gapy = hm;
    hmod.grid_length = G;
    hmod.grid = grid;
    hmod.density = dens;
    hmod.slopes = slopes;
    hmod.kmer_row = rows;
    memset(&st, 0, sizeof(st));
    rc = or_get_aligned_pairs(&hmod, ref, lX, ev, lY, anchors, na, &p, 1, 1, pairs, 16 * (lX + lY), &n, &st);
    printf("hdp rc %d pairs %lld windows %lld ll %.6f\n", rc, (long long)n, (long long)st.n_windows,
           st.log_likelihood);
    if (rc || n == 0) return 1;
    /* degenerate references: every CpG cytosine 'X' over "CE" */
    char *xref = malloc((size_t)nt + 1);
    memcpy(xref, ref, (size_t)nt + 1);
    for (int i = 0; i + 1 < nt; i++)
        if (xref[i] == 'C' && xref[i + 1] == 'G') xref[i] = 'X';
    or_model xm = gm;
    xm.alphabet_size = 5;
    strcpy(xm.alphabet, "ACEGT");
    static double xmatch[5 * 15625], xgapy[5 * 15625];
    for (int i = 0; i < 15625; i++) {
        memcpy(xmatch + 5 * i, match + 5 * (i % NK), sizeof(double) * 5);
        memcpy(xgapy + 5 * i, gapy + 5 * (i % NK), sizeof(double) * 5);
    }
    xm.match = xmatch;
    xm.gapy = xgapy;
    memset(&st, 0, sizeof(st));
    rc = or_get_aligned_pairs_degenerate(&xm, xref, lX, ev, lY, anchors, na, &p, 1, 1, "CE", pairs, 16 * (lX + lY),
                                         &n, &st);
    printf("degenerate rc %d pairs %lld windows %lld ll %.6f\n", rc, (long long)n, (long long)st.n_windows,
           st.log_likelihood);
    if (rc || n == 0) return 1;
    /* EM expectations with records */
    double trans[9], evx[2 * NK], post[NK], rec[5 * 4096];
    unsigned char obs[NK];
    for (int i = 0; i < 9; i++) trans[i] = 0.0;
    for (int i = 0; i < NK; i++) { post[i] = 0.0; obs[i] = 0; evx[2 * i] = evx[2 * i + 1] = 0.0; }
    or_expect ex;
    memset(&ex, 0, sizeof(ex));
    ex.transitions = trans;
    ex.event_exp = evx;
    ex.posteriors = post;
    ex.observed = obs;
    ex.rec = rec;
    ex.rec_cap = 4096;
    ex.rec_threshold = 0.5;
    rc = or_get_expectations(&gm, ref, lX, ev, lY, anchors, na, &p, 1, 1, &ex);
    printf("expectations rc %d likelihood %.6f records %lld\n", rc, ex.likelihood, (long long)ex.n_rec);
    if (rc) return 1;
    /* band and split geometry */
    int64_t *L = malloc(sizeof(int64_t) * (size_t)(lX + lY + 1)), *R = malloc(sizeof(int64_t) * (size_t)(lX + lY + 1));
    rc = or_band(anchors, na, lX, lY, 20, L, R);
    int64_t sp[4 * 64];
    const int64_t ns = or_split_points(anchors, na, lX, lY, 60 * 60, 0, 0, sp, 64);
    printf("band rc %d split blocks %lld\n", rc, (long long)ns);
    free(L); free(R); free(xref); free(rows); free(slopes); free(dens); free(grid); free(pairs); free(path);
    free(ev); free(ref);
    return rc != 0;
}
