/*
 * machine_check.c -- exercises signalalign_machine.h (host-only calls, no GPU).
 *
 *   machine_check params <.npRead> <template .model> <complement .model>
 *       signalUtils_estimateNanoporeParams on each strand of the read the way
 *       signalMachine runs it (descaling machine, one-D assignments from the
 *       strand read, drift correction); prints per strand
 *       "<t|c> n shift scale var drift shift_sd scale_sd var_sd".
 *   machine_check rewrite <.npRead> <template .model> <complement .model> <reference> <in.tsv> <out.tsv>
 *       Re-estimates the parameters as above, then writes the aligned pairs of a
 *       "full" TSV (reference position, event, posterior, path k-mer of each row)
 *       through writePosteriorProbsFull, with the whole reference as the aligned
 *       span (template rows: forward target, offset 0; complement rows: reverse
 *       complement target, offset = reference length).
 *   machine_check anchors <trim> [<.npRead> <2D: 0|1> <t|c>]
 *       Reads one exonerate cigar line from stdin, prints the rebased alignment
 *       ("pa start1 end1 strand1 start2 end2 strand2"), the anchors ("a x y") and,
 *       with a read, the anchors remapped to that strand's events ("r x y").
 *   machine_check nhdp <.nhdp> <.model> <k-mer id>...
 *       deserialize_nhdp; prints "h num_dps n_rows grid_length grid[0] grid[last]",
 *       then per k-mer id "k id observed row E sd" (E, sd of observed DPs from
 *       nhdp_dp_expected_value / nhdp_dp_variance), then "m id mean sd" of the
 *       .model machine after stateMachine3_setModelToHdpExpectedValues.
 *   machine_check solve <n> <A row-major...> <b...>
 *       nanopore_lineq_solve; prints x.
 *   machine_check filter <x y>...
 *       filterToRemoveOverlap of the (sorted) pairs; prints "x y" per kept pair.
 *   machine_check sort <p x y>...
 *       stList_sort with sortByXPlusYCoordinate2; prints "p x y" in the new order.
 */
#define _GNU_SOURCE
#include "signalalign_machine.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static void print_params(const char *s, int64_t n, const NanoporeReadAdjustmentParameters *q) {
    printf("%s %lld %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", s, (long long)n, q->shift, q->scale, q->var,
           q->drift, q->shift_sd, q->scale_sd, q->var_sd);
}

/* the estimate signalMachine makes for each strand (signalMachine.c:675-716) */
static void estimate(NanoporeRead *r, const char *tmodel, const char *cmodel, StateMachine **sMt,
                     StateMachine **sMc, int print) {
    *sMt = getStateMachine3_descaled(tmodel, r->templateParams, false);
    stList *m = signalUtils_templateOneDAssignmentsFromRead(r, *sMt, 0.1);
    const int64_t nt = stList_length(m);
    stList_destruct(m);
    signalUtils_estimateNanoporeParams(*sMt, r, &r->templateParams, 0.1, signalUtils_templateOneDAssignmentsFromRead,
                                       nanopore_adjustTemplateEventsForDrift);
    if (print) print_params("t", nt, &r->templateParams);
    *sMc = NULL;
    if (r->twoD) {
        *sMc = getStateMachine3_descaled(cmodel, r->complementParams, false);
        m = signalUtils_complementOneDAssignmentsFromRead(r, *sMc, 0.1);
        const int64_t nc = stList_length(m);
        stList_destruct(m);
        signalUtils_estimateNanoporeParams(*sMc, r, &r->complementParams, 0.1,
                                           signalUtils_complementOneDAssignmentsFromRead,
                                           nanopore_adjustComplementEventsForDrift);
        if (print) print_params("c", nc, &r->complementParams);
    }
}

static char *read_first_line(const char *path) {
    FILE *f = fopen(path, "r");
    if (!f) return NULL;
    char *buf = NULL;
    size_t cap = 0;
    ssize_t n = getline(&buf, &cap, f);
    fclose(f);
    if (n < 0) return NULL;
    while (n > 0 && (buf[n - 1] == '\n' || buf[n - 1] == '\r')) buf[--n] = '\0';
    return buf;
}

static char revcomp_base(char c) {
    return c == 'A' ? 'T' : c == 'T' ? 'A' : c == 'C' ? 'G' : c == 'G' ? 'C' : c;
}

static int rewrite(char **argv) {
    NanoporeRead *r = nanopore_loadNanoporeReadFromFile(argv[2]);
    StateMachine *sMt, *sMc;
    estimate(r, argv[3], argv[4], &sMt, &sMc, 0);
    char *ref = read_first_line(argv[5]);
    if (!ref) return 2;
    const int64_t L = (int64_t)strlen(ref);
    char *rc = (char *)malloc((size_t)L + 1);
    for (int64_t i = 0; i < L; i++) rc[i] = revcomp_base(ref[L - 1 - i]);
    rc[L] = '\0';
    const int64_t k = sMt->kmerLength;

    FILE *in = fopen(argv[6], "r");
    if (!in) return 2;
    stList *pairs[2] = {stList_construct3(0, (void (*)(void *))stIntTuple_destruct),
                        stList_construct3(0, (void (*)(void *))stIntTuple_destruct)};
    stList *kmers = stList_construct3(0, free);
    char contig[256] = {0}, label[1024] = {0};
    char line[4096];
    while (fgets(line, sizeof line, in)) {
        char c0[256], rk[64], lab[1024], st[8], ki[64], pk[64];
        long long x_adj, y;
        double m, n, d, se, sn, p, dm, em;
        if (sscanf(line, "%255s %lld %63s %1023s %7s %lld %lf %lf %lf %63s %lf %lf %lf %lf %lf %63s", c0, &x_adj, rk,
                   lab, st, &y, &m, &n, &d, ki, &se, &sn, &p, &dm, &em, pk) != 16)
            return 3;
        strcpy(contig, c0);
        strcpy(label, lab);
        const int s = st[0] == 'c';
        /* template: x = x_adj; complement: x_adj = lenKmers - (x + (L - L)) */
        const int64_t x = s ? (L - k) - x_adj : x_adj;
        char *pkc = strdup(pk);
        stList_append(kmers, pkc);
        stList_append(pairs[s], stIntTuple_construct4((int64_t)llround(p * PAIR_ALIGNMENT_PROB_1), x, y,
                                                      (int64_t)(intptr_t)pkc));
    }
    fclose(in);
    remove(argv[7]);
    writePosteriorProbsFull(argv[7], label, sMt, r->templateParams, r->templateEvents, ref, true, contig, 0, 0,
                            pairs[0], template);
    if (sMc)
        writePosteriorProbsFull(argv[7], label, sMc, r->complementParams, r->complementEvents, rc, true, contig, 0, L,
                                pairs[1], complement);
    stList_destruct(pairs[0]);
    stList_destruct(pairs[1]);
    stList_destruct(kmers);
    stateMachine_destruct(sMt);
    stateMachine_destruct(sMc);
    nanopore_nanoporeReadDestruct(r);
    free(ref);
    free(rc);
    return 0;
}

static int anchors(int argc, char **argv) {
    struct PairwiseAlignment *pA = cigarRead(stdin);
    if (!pA) return 2;
    const int64_t start2 = pA->start2;
    PairwiseAlignmentParameters *p = pairwiseAlignmentBandingParameters_construct();
    p->constraintDiagonalTrim = atoll(argv[2]);
    stList *a = signalUtils_guideAlignmentToRebasedAnchorPairs(pA, p);
    printf("pa %lld %lld %lld %lld %lld %lld\n", (long long)pA->start1, (long long)pA->end1, (long long)pA->strand1,
           (long long)pA->start2, (long long)pA->end2, (long long)pA->strand2);
    for (int64_t i = 0; i < stList_length(a); i++) {
        stIntTuple *t = (stIntTuple *)stList_get(a, i);
        printf("a %lld %lld\n", (long long)stIntTuple_get(t, 0), (long long)stIntTuple_get(t, 1));
    }
    if (argc >= 6) {
        NanoporeRead *r = nanopore_loadNanoporeReadFromFile(argv[3]);
        const int two_d = atoi(argv[4]);
        const int comp = argv[5][0] == 'c';
        int64_t *map = comp ? r->complementEventMap : (two_d ? r->templateEventMap : r->templateStrandEventMap);
        stList *m = signalUtils_getRemappedAnchorPairs(a, map, start2);
        for (int64_t i = 0; i < stList_length(m); i++) {
            stIntTuple *t = (stIntTuple *)stList_get(m, i);
            printf("r %lld %lld\n", (long long)stIntTuple_get(t, 0), (long long)stIntTuple_get(t, 1));
        }
        stList_destruct(m);
        nanopore_nanoporeReadDestruct(r);
    }
    stList_destruct(a);
    pairwiseAlignmentBandingParameters_destruct(p);
    destructPairwiseAlignment(pA);
    return 0;
}

static int nhdp(int argc, char **argv) {
    NanoporeHDP *h = deserialize_nhdp(argv[2]);
    printf("h %lld %lld %lld %.17g %.17g\n", (long long)h->num_dps, (long long)h->n_rows, (long long)h->grid_length,
           h->grid[0], h->grid[h->grid_length - 1]);
    NanoporeReadAdjustmentParameters npp = {1.0, 0.0, 1.0, 0.0, 1.0, 1.0, 0.0};
    StateMachine *sM = getHdpStateMachine(h, argv[3], npp);
    stateMachine3_setModelToHdpExpectedValues(sM, h);
    for (int i = 4; i < argc; i++) {
        const int64_t k = atoll(argv[i]);
        const int obs = h->observed[k];
        printf("k %lld %d %lld %.17g %.17g\n", (long long)k, obs, (long long)h->row[k],
               obs ? nhdp_dp_expected_value(h, k) : 0.0, obs ? sqrt(nhdp_dp_variance(h, k)) : 0.0);
        printf("m %lld %.17g %.17g\n", (long long)k, sM->EMISSION_MATCH_MATRIX[k * MODEL_PARAMS],
               sM->EMISSION_MATCH_MATRIX[k * MODEL_PARAMS + 1]);
    }
    stateMachine_destruct(sM);
    destroy_nanopore_hdp(h);
    return 0;
}

static int solve(int argc, char **argv) {
    const int64_t n = atoll(argv[2]);
    if (argc != 3 + n * n + n) return 2;
    double *A = (double *)malloc(sizeof(double) * (size_t)(n * n)), *b = (double *)malloc(sizeof(double) * (size_t)n),
           *x = (double *)malloc(sizeof(double) * (size_t)n);
    for (int64_t i = 0; i < n * n; i++) A[i] = atof(argv[3 + i]);
    for (int64_t i = 0; i < n; i++) b[i] = atof(argv[3 + n * n + i]);
    nanopore_lineq_solve(A, b, x, n);
    for (int64_t i = 0; i < n; i++) printf("%.17g\n", x[i]);
    free(A);
    free(b);
    free(x);
    return 0;
}

static int filter(int argc, char **argv) {
    stList *l = stList_construct3(0, (void (*)(void *))stIntTuple_destruct);
    for (int i = 2; i + 1 < argc; i += 2) stList_append(l, stIntTuple_construct2(atoll(argv[i]), atoll(argv[i + 1])));
    stList *f = filterToRemoveOverlap(l);
    for (int64_t i = 0; i < stList_length(f); i++) {
        stIntTuple *t = (stIntTuple *)stList_get(f, i);
        printf("%lld %lld\n", (long long)stIntTuple_get(t, 0), (long long)stIntTuple_get(t, 1));
    }
    stList_destruct(f);
    stList_destruct(l);
    return 0;
}

static int sort(int argc, char **argv) {
    stList *l = stList_construct3(0, (void (*)(void *))stIntTuple_destruct);
    for (int i = 2; i + 2 < argc; i += 3)
        stList_append(l, stIntTuple_construct4(atoll(argv[i]), atoll(argv[i + 1]), atoll(argv[i + 2]), 0));
    stList_sort(l, sortByXPlusYCoordinate2);
    for (int64_t i = 0; i < stList_length(l); i++) {
        stIntTuple *t = (stIntTuple *)stList_get(l, i);
        printf("%lld %lld %lld\n", (long long)stIntTuple_get(t, 0), (long long)stIntTuple_get(t, 1),
               (long long)stIntTuple_get(t, 2));
    }
    stList_destruct(l);
    return 0;
}

int main(int argc, char **argv) {
    if (argc >= 5 && strcmp(argv[1], "params") == 0) {
        NanoporeRead *r = nanopore_loadNanoporeReadFromFile(argv[2]);
        StateMachine *sMt, *sMc;
        estimate(r, argv[3], argv[4], &sMt, &sMc, 1);
        stateMachine_destruct(sMt);
        stateMachine_destruct(sMc);
        nanopore_nanoporeReadDestruct(r);
        return 0;
    }
    if (argc == 8 && strcmp(argv[1], "rewrite") == 0) return rewrite(argv);
    if (argc >= 3 && strcmp(argv[1], "anchors") == 0) return anchors(argc, argv);
    if (argc >= 3 && strcmp(argv[1], "solve") == 0) return solve(argc, argv);
    if (argc >= 4 && strcmp(argv[1], "nhdp") == 0) return nhdp(argc, argv);
    if (argc >= 2 && strcmp(argv[1], "filter") == 0) return filter(argc, argv);
    if (argc >= 2 && strcmp(argv[1], "sort") == 0) return sort(argc, argv);
    fprintf(stderr, "usage: machine_check params|rewrite|anchors|nhdp|solve|filter|sort ...\n");
    return 1;
}
