/*
 * signalMachine -- the reference's per-read aligner executable (signalMachine.c:369-789)
 * over the MI355X library: the guide alignment (exonerate CIGAR) on stdin, the
 * read (.npRead), the models and the two reference files as options, aligned
 * pairs appended to the posteriors file as TSV.  Same options and outputs as the
 * reference, so its pipeline (runSignalAlign) can run this executable in its place.
 *
 * Served: three-state Gaussian machines (--templateModel / --complementModel) and
 * HDP machines (--sm3Hdp with --templateHdp / --complementHdp), template-only and
 * 2D reads, output formats full / variantCaller / assignments, the expectations
 * routine (--templateExpectations / --complementExpectations: the EM expectations
 * files instead of the alignment), read scaling re-estimated as the reference
 * does.  Refused with a message: error-correction runs (--error_correct_path).
 */
#define _GNU_SOURCE
#include "signalalign_machine.h"

#include <getopt.h>
#include <inttypes.h>
#include <stdlib.h>
#include <string.h>

#define ASSIGNMENT_THRESHOLD 0.1

static void usage(void) {
    fprintf(stderr, "signalMachine binary, meant to be used through the signalAlign program.\n");
    fprintf(stderr, "See doc for runSignalAlign for help\n");
}

static void refuse(const char *what) {
    fprintf(stderr, "signalMachine (signalalign_amd): %s is not served by this build\n", what);
    exit(1);
}

static int64_t arg_i64(const char *s, const char *name, int64_t min) {
    char *end;
    long long v = strtoll(s, &end, 10);
    if (end == s || v < min) {
        fprintf(stderr, "signalMachine: bad value for --%s: %s\n", name, s);
        exit(1);
    }
    return (int64_t)v;
}

static StateMachine *build_machine(const char *modelFile, NanoporeReadAdjustmentParameters npp, NanoporeHDP *nHdp) {
    /* buildStateMachine (signalMachine.c:262-283): descaling machine, noise not
       scaled until the parameters are re-estimated */
    FILE *f = fopen(modelFile, "r");
    if (!f) {
        fprintf(stderr, "signalAlign - ERROR: couldn't find model file here: %s\n", modelFile);
        exit(1);
    }
    fclose(f);
    return nHdp ? getHdpStateMachine(nHdp, modelFile, npp) : getStateMachine3_descaled(modelFile, npp, false);
}

static Sequence *event_slice(NanoporeRead *r, double *events, int64_t nEvents, int64_t *map, int64_t mapLength,
                             struct PairwiseAlignment *pA) {
    if (pA->start2 < 0 || pA->end2 < pA->start2 || pA->end2 >= mapLength) {
        fprintf(stderr, "signalMachine: read coordinates [%" PRId64 ", %" PRId64 ") are outside the event map (%" PRId64
                        " bases)\n", pA->start2, pA->end2, mapLength);
        exit(1);
    }
    if (map[pA->end2] > nEvents) {
        fprintf(stderr, "signalMachine: the event map points past the read's events\n");
        exit(1);
    }
    (void)r;
    return makeEventSequenceFromPairwiseAlignment(events, pA->start2, pA->end2, map);
}

int main(int argc, char *argv[]) {
    int64_t diagExpansion = 50;
    double threshold = 0.01;
    int64_t constraintTrim = 14;
    int64_t degenerate = canonicalVariants;
    int64_t outFmt = full;
    bool twoD = false, hdp = false;
    char *templateModelFile = NULL, *complementModelFile = NULL, *readLabel = NULL, *npReadFile = NULL;
    char *forwardReference = NULL, *backwardReference = NULL, *errorCorrectPath = NULL;
    char *posteriorProbsFile = NULL, *templateExpectationsFile = NULL, *complementExpectationsFile = NULL;
    char *templateHdp = NULL, *complementHdp = NULL;

    static struct option long_options[] = {
        {"help", no_argument, 0, 'h'},
        {"sm3Hdp", no_argument, 0, 'd'},
        {"sparse_output", required_argument, 0, 's'},
        {"twoD", no_argument, 0, 'e'},
        {"degenerate", required_argument, 0, 'o'},
        {"templateModel", required_argument, 0, 'T'},
        {"complementModel", required_argument, 0, 'C'},
        {"readLabel", required_argument, 0, 'L'},
        {"npRead", required_argument, 0, 'q'},
        {"forward_reference", required_argument, 0, 'f'},
        {"backward_reference", required_argument, 0, 'b'},
        {"error_correct_path", required_argument, 0, 'p'},
        {"posteriors", required_argument, 0, 'u'},
        {"templateHdp", required_argument, 0, 'v'},
        {"complementHdp", required_argument, 0, 'w'},
        {"templateExpectations", required_argument, 0, 't'},
        {"complementExpectations", required_argument, 0, 'c'},
        {"diagonalExpansion", required_argument, 0, 'x'},
        {"threshold", required_argument, 0, 'D'},
        {"constraintTrim", required_argument, 0, 'm'},
        {0, 0, 0, 0}};
    for (;;) {
        int idx = 0;
        int key = getopt_long(argc, argv, "hdes:o:T:C:L:q:f:b:p:u:v:w:t:c:x:D:m:", long_options, &idx);
        if (key == -1) break;
        switch (key) {
            case 'h': usage(); return 1;
            case 's': outFmt = arg_i64(optarg, "sparse_output", 0); break;
            case 'e': twoD = true; break;
            case 'o': degenerate = arg_i64(optarg, "degenerate", 0); break;
            case 'd': hdp = true; break;
            case 'T': templateModelFile = optarg; break;
            case 'C': complementModelFile = optarg; break;
            case 'L': readLabel = optarg; break;
            case 'q': npReadFile = optarg; break;
            case 'f': forwardReference = optarg; break;
            case 'b': backwardReference = optarg; break;
            case 'p': errorCorrectPath = optarg; break;
            case 'u': posteriorProbsFile = optarg; break;
            case 't': templateExpectationsFile = optarg; break;
            case 'c': complementExpectationsFile = optarg; break;
            case 'v': templateHdp = optarg; break;
            case 'w': complementHdp = optarg; break;
            case 'x': diagExpansion = arg_i64(optarg, "diagonalExpansion", 0); break;
            case 'D': {
                char *end;
                threshold = strtod(optarg, &end);
                if (end == optarg || threshold < 0) {
                    fprintf(stderr, "signalMachine: bad value for --threshold: %s\n", optarg);
                    return 1;
                }
                break;
            }
            case 'm': constraintTrim = arg_i64(optarg, "constraintTrim", 0); break;
            default: usage(); return 1;
        }
    }
    if (!templateModelFile || (twoD && !complementModelFile)) {
        fprintf(stderr, "Missing model files, exiting\n");
        return 1;
    }
    if (templateHdp || complementHdp) {
        if (!templateHdp || (twoD && !complementHdp)) {
            fprintf(stderr, "Need to have template and complement HDPs\n");
            return 1;
        }
        if (!hdp) fprintf(stderr, "[signalAlign] - Warning: this kind of stateMachine does not use the HDPs you gave\n");
    } else if (hdp) {
        fprintf(stderr, "signalMachine: --sm3Hdp needs --templateHdp (and --complementHdp with --twoD)\n");
        return 1;
    }
    if (errorCorrectPath) refuse("error correction (--error_correct_path)");
    if (!forwardReference || !backwardReference) {
        fprintf(stderr, "[signalAlign] - ERROR: did not get reference files %s %s\n",
                forwardReference ? forwardReference : "(null)", backwardReference ? backwardReference : "(null)");
        return 1;
    }
    if (!npReadFile) {
        fprintf(stderr, "signalMachine: missing --npRead\n");
        return 1;
    }
    if (diagExpansion % 2 != 0) {
        fprintf(stderr, "signalMachine: --diagonalExpansion must be even\n");
        return 1;
    }
    if (!readLabel) readLabel = "read";

    /* guide alignment from stdin */
    struct PairwiseAlignment *pA = cigarRead(stdin);
    if (!pA) {
        fprintf(stderr, "signalMachine: no exonerate cigar line on stdin\n");
        return 1;
    }
    PairwiseAlignmentParameters *p = pairwiseAlignmentBandingParameters_construct();
    p->threshold = threshold;
    p->constraintDiagonalTrim = constraintTrim;
    p->diagonalExpansion = diagExpansion;

    NanoporeHDP *nHdpT = hdp ? deserialize_nhdp(templateHdp) : NULL;
    NanoporeHDP *nHdpC = hdp && twoD ? deserialize_nhdp(complementHdp) : NULL;
    ReferenceSequence *R = signalUtils_ReferenceSequenceConstructFull(forwardReference, backwardReference, pA);
    NanoporeRead *npRead = nanopore_loadNanoporeReadFromFile(npReadFile);

    int64_t *tMap = twoD ? npRead->templateEventMap : npRead->templateStrandEventMap;
    const int64_t tMapLength = twoD ? npRead->readLength : npRead->templateReadLength;
    Sequence *tEvents = event_slice(npRead, npRead->templateEvents, npRead->nbTemplateEvents, tMap, tMapLength, pA);
    Sequence *cEvents = twoD ? event_slice(npRead, npRead->complementEvents, npRead->nbComplementEvents,
                                           npRead->complementEventMap, npRead->readLength, pA)
                             : NULL;
    /* pairs come back from (0, 0) of the slices: their offsets in the whole read
       and the whole reference */
    const int64_t tShift = tMap[pA->start2];
    const int64_t cShift = twoD ? npRead->complementEventMap[pA->start2] : 0;
    const int64_t rShiftT = pA->start1;
    const int64_t rShiftC = twoD ? pA->end1 : 0;
    const bool forward = pA->strand1;
    const int64_t start2 = pA->start2;

    stList *anchorPairs = signalUtils_guideAlignmentToRebasedAnchorPairs(pA, p); /* rebases pA */

    if (templateExpectationsFile || complementExpectationsFile) {
        /* expectations routine (signalMachine.c:618-700) */
        const StateMachineType type = hdp ? threeStateHdp : threeState;
        StateMachine *sMt = build_machine(templateModelFile, npRead->templateParams, nHdpT);
        signalUtils_estimateNanoporeParams(sMt, npRead, &npRead->templateParams, ASSIGNMENT_THRESHOLD,
                                           signalUtils_templateOneDAssignmentsFromRead,
                                           nanopore_adjustTemplateEventsForDrift);
        Hmm *tExp = hmmContinuous_getExpectationsHmm(sMt, p->threshold, 0.001, 0.001);
        fprintf(stderr, "signalAlign - getting expectations for template\n");
        getSignalExpectations(sMt, tExp, tEvents, tMap, start2, R->getTemplateTargetSequence(R), p, anchorPairs,
                              (DegenerateType)degenerate);
        if (hdp)
            fprintf(stderr, "signalAlign - got %" PRId64 "template HDP assignments\n",
                    hmmContinuous_howManyAssignments(tExp));
        if (templateExpectationsFile) {
            fprintf(stderr, "signalAlign - writing expectations to file: %s\n", templateExpectationsFile);
            hmmContinuous_writeToFile(templateExpectationsFile, tExp, type);
        }
        StateMachine *sMc = NULL;
        Hmm *cExp = NULL;
        if (twoD) {
            fprintf(stderr, "signalAlign - getting expectations for complement\n");
            sMc = build_machine(complementModelFile, npRead->complementParams, nHdpC);
            signalUtils_estimateNanoporeParams(sMc, npRead, &npRead->complementParams, ASSIGNMENT_THRESHOLD,
                                               signalUtils_complementOneDAssignmentsFromRead,
                                               nanopore_adjustComplementEventsForDrift);
            cExp = hmmContinuous_getExpectationsHmm(sMc, p->threshold, 0.001, 0.001);
            getSignalExpectations(sMc, cExp, cEvents, npRead->complementEventMap, start2,
                                  R->getComplementTargetSequence(R), p, anchorPairs, (DegenerateType)degenerate);
            if (hdp)
                fprintf(stderr, "signalAlign - got %" PRId64 "complement HDP assignments\n",
                        hmmContinuous_howManyAssignments(cExp));
            if (complementExpectationsFile) {
                fprintf(stderr, "signalAlign - writing expectations to file: %s\n", complementExpectationsFile);
                hmmContinuous_writeToFile(complementExpectationsFile, cExp, type);
            }
        }
        hmmContinuous_destruct(tExp, type);
        hmmContinuous_destruct(cExp, type);
        stateMachine_destruct(sMt);
        stateMachine_destruct(sMc);
        stList_destruct(anchorPairs);
        sequence_destruct(tEvents);
        sequence_destruct(cEvents);
        signalUtils_ReferenceSequenceDestruct(R);
        nanopore_nanoporeReadDestruct(npRead);
        pairwiseAlignmentBandingParameters_destruct(p);
        destructPairwiseAlignment(pA);
        destroy_nanopore_hdp(nHdpT);
        destroy_nanopore_hdp(nHdpC);
        return 0;
    }

    fprintf(stderr, "signalAlign - starting template alignment\n");
    StateMachine *sMt = build_machine(templateModelFile, npRead->templateParams, nHdpT);
    signalUtils_estimateNanoporeParams(sMt, npRead, &npRead->templateParams, ASSIGNMENT_THRESHOLD,
                                       signalUtils_templateOneDAssignmentsFromRead,
                                       nanopore_adjustTemplateEventsForDrift);
    if (nHdpT) stateMachine3_setModelToHdpExpectedValues(sMt, nHdpT);
    stList *tPairs = performSignalAlignment(sMt, tEvents, tMap, start2, R->getTemplateTargetSequence(R), p,
                                            anchorPairs, (DegenerateType)degenerate);
    const double tScore = scoreByPosteriorProbabilityIgnoringGaps(tPairs);
    stList_sort(tPairs, sortByXPlusYCoordinate2);
    if (posteriorProbsFile)
        outputAlignment((OutputFormat)outFmt, posteriorProbsFile, readLabel, sMt, npRead->templateParams,
                        npRead->templateEvents, R->getTemplateTargetSequence(R), forward, pA->contig1, tShift, rShiftT,
                        tPairs, template);

    stList *cPairs = NULL;
    double cScore = 0.0;
    StateMachine *sMc = NULL;
    if (twoD) {
        fprintf(stderr, "signalAlign - starting complement alignment\n");
        sMc = build_machine(complementModelFile, npRead->complementParams, nHdpC);
        signalUtils_estimateNanoporeParams(sMc, npRead, &npRead->complementParams, ASSIGNMENT_THRESHOLD,
                                           signalUtils_complementOneDAssignmentsFromRead,
                                           nanopore_adjustComplementEventsForDrift);
        if (nHdpC) stateMachine3_setModelToHdpExpectedValues(sMc, nHdpC);
        cPairs = performSignalAlignment(sMc, cEvents, npRead->complementEventMap, start2,
                                        R->getComplementTargetSequence(R), p, anchorPairs,
                                        (DegenerateType)degenerate);
        cScore = scoreByPosteriorProbabilityIgnoringGaps(cPairs);
        stList_sort(cPairs, sortByXPlusYCoordinate2);
        if (posteriorProbsFile)
            outputAlignment((OutputFormat)outFmt, posteriorProbsFile, readLabel, sMc, npRead->complementParams,
                            npRead->complementEvents, R->getComplementTargetSequence(R), forward, pA->contig1, cShift,
                            rShiftC, cPairs, complement);
    }

    fprintf(stdout, "%s %" PRId64 "\t%" PRId64 "(%f)\t", readLabel, stList_length(anchorPairs), stList_length(tPairs),
            tScore);
    if (twoD)
        fprintf(stdout, "%" PRId64 "(%f)\n", stList_length(cPairs), cScore);
    else
        fprintf(stdout, "\n");

    stList_destruct(tPairs);
    stList_destruct(cPairs);
    stList_destruct(anchorPairs);
    stateMachine_destruct(sMt);
    stateMachine_destruct(sMc);
    sequence_destruct(tEvents);
    sequence_destruct(cEvents);
    signalUtils_ReferenceSequenceDestruct(R);
    nanopore_nanoporeReadDestruct(npRead);
    pairwiseAlignmentBandingParameters_destruct(p);
    destructPairwiseAlignment(pA);
    destroy_nanopore_hdp(nHdpT);
    destroy_nanopore_hdp(nHdpC);
    fprintf(stderr, "signalAlign - SUCCESS: finished alignment of query %s, exiting\n", readLabel);
    return 0;
}
