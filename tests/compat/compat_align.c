/*
 * compat_align.c -- exercises signalalign_compat.h the way the reference's own
 * tests call the path (tests/stateMachineTests.c:565-692, :695-716).
 *
 *   compat_align kat5 <5-mer .model>
 *       The 5-mer inputs of test_sm3_5merDiagonalDPCalculations through
 *       getStateMachine3 + getAlignedPairsUsingAnchors (no anchors, expansion 2,
 *       threshold 0.2, not ragged): exactly the 7 pairs (0,0) (1,1) (2,2) (3,3)
 *       (5,4) (6,5) (8,6), each path k-mer equal to the reference k-mer.
 *   compat_align align <.model> <input.txt> <output.tsv> [degenerate type]
 *       One read through getStateMachine3_descaled; writes the total
 *       log-likelihood and the pairs (p, x, y, path k-mer) in list order.  The
 *       reference Sequence takes the DegenerateType given (its numeric value,
 *       default cytosineMethylation2), whose options an 'X' base stands for.
 *
 * Exit status 0 on success; the compat layer aborts on library errors.
 */
#include "signalalign_compat.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int kat5(const char *model) {
    char sX[] = "ACGATATGGACAT";
    double sY[28] = {
        70.0423375640843, 2.1070814631739,  0.0571, 0.0,  73.7087073662952, 1.90162684687837, 0.0571, 0.1,
        105.375581864011, 2.87252862011704, 0.0571, 0.2,  82.9620934477158, 2.38320603353748, 0.0571, 0.3,
        84.6977645711335, 3.08486975249442, 0.0571, 0.4,  58.0551144225027, 2.52297561817531, 0.0571, 0.5,
        94.337668063878,  1.9731952395105,  0.0571, 0.67,
    };
    StateMachine *sM = getStateMachine3(model);
    int64_t lX = sequence_correctSeqLength((int64_t)strlen(sX), kmer, sM->kmerLength);
    Sequence *SsX = sequence_constructKmerSequence(lX, sX, sequence_getKmer, sequence_sliceNucleotideSequence,
                                                   "CEO", 3, kmer);
    Sequence *SsY = sequence_construct(7, sY, sequence_getEvent, event);
    PairwiseAlignmentParameters *p = pairwiseAlignmentBandingParameters_construct();
    p->threshold = 0.2;
    p->diagonalExpansion = 2;
    stList *anchors = stList_construct();
    stList *pairs = getAlignedPairsUsingAnchors(sM, SsX, SsY, anchors, p, diagonalCalculationPosteriorMatchProbs,
                                                0, 0);
    static const int64_t want[7][2] = {{0, 0}, {1, 1}, {2, 2}, {3, 3}, {5, 4}, {6, 5}, {8, 6}};
    int ok = stList_length(pairs) == 7;
    for (int64_t i = 0; ok && i < stList_length(pairs); i++) {
        stIntTuple *t = stList_get(pairs, i);
        int64_t x = stIntTuple_get(t, 1), y = stIntTuple_get(t, 2);
        int hit = 0;
        for (int j = 0; j < 7; j++) hit |= want[j][0] == x && want[j][1] == y;
        const char *km = (const char *)(intptr_t)stIntTuple_get(t, 3);
        ok = hit && stIntTuple_length(t) == 4 && strncmp(km, sX + x, 5) == 0 && km[5] == '\0' &&
             stIntTuple_get(t, 0) >= 2000000 && stIntTuple_get(t, 0) <= PAIR_ALIGNMENT_PROB_1;
    }
    printf("kat5: %lld pairs, log-likelihood %.10f: %s\n", (long long)stList_length(pairs),
           sa_compat_last_log_likelihood(), ok ? "ok" : "FAILED");
    stList_destruct(pairs);
    stList_destruct(anchors);
    pairwiseAlignmentBandingParameters_destruct(p);
    sequence_destruct(SsX);
    sequence_destruct(SsY);
    stateMachine_destruct(sM);
    return ok ? 0 : 1;
}

static int align(const char *model, const char *in, const char *out, DegenerateType dtype) {
    FILE *f = fopen(in, "r");
    if (!f) return 2;
    long long nref = 0, nev = 0, nanc = 0, rl = 0, rr = 0, exp_ = 0, mind = 0, tb = 0, split = 0;
    double thr = 0;
    NanoporeReadAdjustmentParameters npp;
    memset(&npp, 0, sizeof(npp));
    npp.scale_sd = npp.var_sd = 1.0;
    if (fscanf(f, "%lld", &nref) != 1) return 2;
    char *ref = malloc((size_t)nref + 1);
    if (fscanf(f, "%s", ref) != 1) return 2;
    if (fscanf(f, "%lf %lf %lf %lld %lld", &npp.scale, &npp.shift, &npp.var, &rl, &rr) != 5) return 2;
    if (fscanf(f, "%lld %lf %lld %lld %lld", &exp_, &thr, &mind, &tb, &split) != 5) return 2;
    if (fscanf(f, "%lld", &nev) != 1) return 2;
    double *ev = malloc(sizeof(double) * 4 * (size_t)(nev ? nev : 1));
    for (long long i = 0; i < 4 * nev; i++)
        if (fscanf(f, "%lf", &ev[i]) != 1) return 2;
    if (fscanf(f, "%lld", &nanc) != 1) return 2;
    stList *anchors = stList_construct3(0, (void (*)(void *))stIntTuple_destruct);
    for (long long i = 0; i < nanc; i++) {
        long long x, y;
        if (fscanf(f, "%lld %lld", &x, &y) != 2) return 2;
        stList_append(anchors, stIntTuple_construct2(x, y));
    }
    fclose(f);

    StateMachine *sM = getStateMachine3_descaled(model, npp, 0);
    int64_t lX = sequence_correctSeqLength(nref, kmer, sM->kmerLength);
    Sequence *SsX = sequence_constructReferenceKmerSequence(lX, ref, sequence_getKmer, sequence_sliceNucleotideSequence,
                                                            dtype, kmer);
    Sequence *SsY = sequence_constructEventSequence(nev, ev);
    PairwiseAlignmentParameters *p = pairwiseAlignmentBandingParameters_construct();
    p->threshold = thr;
    p->diagonalExpansion = exp_;
    p->minDiagsBetweenTraceBack = mind;
    p->traceBackDiagonals = tb;
    p->splitMatrixBiggerThanThis = split;
    stList *pairs = getAlignedPairsUsingAnchors(sM, SsX, SsY, anchors, p, diagonalCalculationPosteriorMatchProbs,
                                                rl != 0, rr != 0);
    FILE *o = fopen(out, "w");
    if (!o) return 2;
    fprintf(o, "%.17g\n", sa_compat_last_log_likelihood());
    for (int64_t i = 0; i < stList_length(pairs); i++) {
        stIntTuple *t = stList_get(pairs, i);
        fprintf(o, "%lld\t%lld\t%lld\t%s\n", (long long)stIntTuple_get(t, 0), (long long)stIntTuple_get(t, 1),
                (long long)stIntTuple_get(t, 2), (const char *)(intptr_t)stIntTuple_get(t, 3));
    }
    fclose(o);
    printf("align: %lld pairs\n", (long long)stList_length(pairs));
    stList_destruct(pairs);
    stList_destruct(anchors);
    pairwiseAlignmentBandingParameters_destruct(p);
    sequence_destruct(SsX);
    sequence_destruct(SsY);
    stateMachine_destruct(sM);
    free(ref);
    free(ev);
    return 0;
}

int main(int argc, char **argv) {
    if (argc == 3 && strcmp(argv[1], "kat5") == 0) return kat5(argv[2]);
    if (argc == 5 && strcmp(argv[1], "align") == 0) return align(argv[2], argv[3], argv[4], cytosineMethylation2);
    if (argc == 6 && strcmp(argv[1], "align") == 0) /* the degenerate type of an 'X' reference */
        return align(argv[2], argv[3], argv[4], (DegenerateType)atoi(argv[5]));
    fprintf(stderr, "usage: compat_align kat5 <model> | align <model> <input> <output> [degenerate type]\n");
    return 2;
}
