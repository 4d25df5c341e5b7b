/*
 * signalalign_compat.h -- signalAlign's pairwiseAligner C API for the banded
 * forward-backward path, served by the MI355X library (libsignalalign_amd.so).
 *
 * A C caller written against the reference keeps its calls:
 *   getStateMachine3_descaled / getStateMachine3 / stateMachine_destruct
 *   sequence_construct* / sequence_destruct
 *   pairwiseAlignmentBandingParameters_construct / _destruct
 *   getAlignedPairsUsingAnchors(sM, SsX, SsY, anchors, p,
 *                               diagonalCalculationPosteriorMatchProbs, raggedL, raggedR)
 * and gets the same stList of stIntTuple (floor(p * 1e7), x, y, path k-mer).
 *
 * Types keep the reference's field layout so that code reading sM->kmerLength,
 * sM->EMISSION_MATCH_MATRIX, p->threshold or SsY->elements is unchanged:
 *   Sequence, SequenceType, DegenerateType      inc/pairwiseAligner.h:43-66
 *   PairwiseAlignmentParameters                 inc/pairwiseAligner.h:123-134
 *   StateMachine, StateMachineType              inc/stateMachine.h:28-37, 76-104
 *   StateMachine3                               inc/stateMachine.h:139-160
 *   NanoporeReadAdjustmentParameters            inc/nanopore.h:11-19
 *
 * The reference links sonLib for stList / stIntTuple.  libsignalalign_compat.so
 * carries a minimal implementation of the calls this API trades in; a build that
 * links sonLib compiles sa_compat.c with SA_COMPAT_USE_SONLIB and drops them.
 *
 * Errors: the reference aborts (st_errAbort, exit) on bad input; so does this API,
 * after printing the library's message.  Calls are serialised on one device
 * context per process (device 0, or sa_compat_set_device before the first call).
 */
#ifndef SIGNALALIGN_COMPAT_H
#define SIGNALALIGN_COMPAT_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PAIR_ALIGNMENT_PROB_1 10000000
#define NB_EVENT_PARAMS 4
#define MODEL_PARAMS 5

/* ---- sonLib containers (the subset this API trades in) ------------------- */
typedef struct _stList stList;
typedef struct _stIntTuple stIntTuple;

stList *stList_construct(void);
stList *stList_construct3(int64_t length, void (*destructElement)(void *));
void stList_append(stList *list, void *item);
void *stList_get(stList *list, int64_t index);
int64_t stList_length(stList *list);
void stList_destruct(stList *list);

stIntTuple *stIntTuple_construct2(int64_t v0, int64_t v1);
stIntTuple *stIntTuple_construct3(int64_t v0, int64_t v1, int64_t v2);
stIntTuple *stIntTuple_construct4(int64_t v0, int64_t v1, int64_t v2, int64_t v3);
int64_t stIntTuple_get(stIntTuple *tuple, int64_t index);
int64_t stIntTuple_length(stIntTuple *tuple);
void stIntTuple_destruct(stIntTuple *tuple);

/* ---- sequences (inc/pairwiseAligner.h:43-66) ------------------------------ */
typedef enum { nucleotide = 0, kmer = 1, event = 2 } SequenceType;

typedef enum {
    cytosineMethylation2 = 0,
    cytosineMethylation3 = 1,
    adenosineMethylation = 2,
    canonicalVariants = 3
} DegenerateType;

typedef struct _sequence Sequence;
struct _sequence {
    int64_t length;
    SequenceType type;
    void *elements;
    char *degenerateBases;
    int64_t nbDegenerateBases;
    void *(*get)(void *elements, int64_t index);
    Sequence *(*sliceFcn)(Sequence *, int64_t, int64_t);
};

int64_t sequence_correctSeqLength(int64_t length, SequenceType type, int64_t kmerLength);
void *sequence_getKmer(void *elements, int64_t index);
void *sequence_getEvent(void *elements, int64_t index);
Sequence *sequence_sliceNucleotideSequence(Sequence *inputSequence, int64_t start, int64_t sliceLength);
Sequence *sequence_sliceEventSequence(Sequence *inputSequence, int64_t start, int64_t sliceLength);
Sequence *sequence_construct(int64_t length, void *elements, void *(*getFcn)(void *, int64_t), SequenceType type);
Sequence *sequence_construct2(int64_t length, void *elements, void *(*getFcn)(void *, int64_t),
                              Sequence *(*sliceFcn)(Sequence *, int64_t, int64_t), SequenceType type);
Sequence *sequence_constructKmerSequence(int64_t length, void *elements, void *(*getFcn)(void *, int64_t),
                                         Sequence *(*sliceFcn)(Sequence *, int64_t, int64_t),
                                         char *nucleotideOptions, int64_t nbOptions, SequenceType type);
Sequence *sequence_constructReferenceKmerSequence(int64_t length, void *elements,
                                                  void *(*getFcn)(void *, int64_t),
                                                  Sequence *(*sliceFcn)(Sequence *, int64_t, int64_t),
                                                  DegenerateType dType, SequenceType type);
Sequence *sequence_constructEventSequence(int64_t length, void *events);
void sequence_destruct(Sequence *seq);

/* ---- read adjustment (inc/nanopore.h:11-19) -------------------------------- */
typedef struct _nanoporeReadAdjustmentParameters {
    double scale;
    double shift;
    double var;
    double shift_sd;
    double scale_sd;
    double var_sd;
    double drift;
} NanoporeReadAdjustmentParameters;

/* ---- state machines (inc/stateMachine.h) ----------------------------------- */
typedef enum {
    fiveState = 0,
    fiveStateAsymmetric = 1,
    threeState = 2,
    threeStateAsymmetric = 3,
    vanilla = 4,
    echelon = 5,
    fourState = 6,
    threeStateHdp = 7,
} StateMachineType;

typedef struct _stateMachine StateMachine;
struct _stateMachine {
    StateMachineType type;
    int64_t stateNumber;
    int64_t matchState;
    int64_t parameterSetSize;
    double scale;
    double shift;
    double var;
    char *alphabet;
    int64_t alphabetSize;
    int64_t kmerLength;
    double *EMISSION_MATCH_MATRIX;  /* 5 values per k-mer: level mean, sd, noise mean, sd, lambda */
    double *EMISSION_GAP_X_PROBS;   /* log skip emission per k-mer */
    double *EMISSION_GAP_Y_MATRIX;  /* extra-event table, same layout as the match table */
    double (*startStateProb)(StateMachine *sM, int64_t state);
    double (*endStateProb)(StateMachine *sM, int64_t state);
    double (*raggedEndStateProb)(StateMachine *sM, int64_t state);
    double (*raggedStartStateProb)(StateMachine *sM, int64_t state);
    void (*cellCalculate)(StateMachine *sM, void *current, void *lower, void *middle, void *upper, void *cX,
                          void *cY, void (*doTransition)(double *, double *, int64_t, int64_t, double, double, void *),
                          void *extraArgs);
    void (*cellCalculateUpdateExpectations)(double *fromCells, double *toCells, int64_t from, int64_t to, double eP,
                                            double tP, void *extraArgs);
};

typedef struct _StateMachine3 StateMachine3;
struct _StateMachine3 {
    StateMachine model;
    double TRANSITION_MATCH_CONTINUE;
    double TRANSITION_MATCH_FROM_GAP_X;
    double TRANSITION_MATCH_FROM_GAP_Y;
    double TRANSITION_GAP_OPEN_X;
    double TRANSITION_GAP_OPEN_Y;
    double TRANSITION_GAP_EXTEND_X;
    double TRANSITION_GAP_EXTEND_Y;
    double TRANSITION_GAP_SWITCH_TO_X;
    double TRANSITION_GAP_SWITCH_TO_Y;
    double (*getXGapProbFcn)(StateMachine *self, const double *emissionXGapProbs, void *i);
    double (*getYGapProbFcn)(StateMachine *self, void *x, void *y, bool match);
    double (*getMatchProbFcn)(StateMachine *self, void *x, void *y, bool match);
};

/* impl/stateMachine.c:1627-1643: .model file, events descaled with npp (scale,
 * shift, var); scaleNoise applies emissions_signal_scaleNoise (:609-629). */
StateMachine *getStateMachine3_descaled(const char *modelFile, NanoporeReadAdjustmentParameters npp,
                                        bool scaleNoise);
/* impl/stateMachine.c:1645-1653: .model file, events compared undescaled. */
StateMachine *getStateMachine3(const char *modelFile);
/* HDP machine (impl/stateMachine.c:1655-1669) built from the density grid the
 * reference's NanoporeHDP evaluates: `grid` (grid_length uniform points),
 * `density` and `slopes` (n_rows x grid_length), `kmer_row` (row of the first
 * observed ancestor DP per k-mer, -1 for none) and the level means used to descale
 * (written into the match table's mean column, where the reference's HDP machine
 * reads them; NULL keeps the .model file's).  Takes the .model file for its
 * transitions, alphabet, k-mer length and Gaussian table (read scaling uses it). */
StateMachine *getHdpStateMachineFromGrid(const char *modelFile, const double *grid, int64_t grid_length,
                                         const double *density, const double *slopes, int64_t n_rows,
                                         const int64_t *kmer_row, const double *level_mean,
                                         NanoporeReadAdjustmentParameters npp);
void stateMachine_destruct(StateMachine *stateMachine);

/* ---- parameters (inc/pairwiseAligner.h:123-134) ---------------------------- */
typedef struct _pairwiseAlignmentBandingParameters {
    double threshold;
    int64_t minDiagsBetweenTraceBack;
    int64_t traceBackDiagonals;
    int64_t diagonalExpansion;
    int64_t constraintDiagonalTrim;
    int64_t anchorMatrixBiggerThanThis;
    int64_t repeatMaskMatrixBiggerThanThis;
    int64_t splitMatrixBiggerThanThis;
    bool alignAmbiguityCharacters;
    float gapGamma;
} PairwiseAlignmentParameters;

PairwiseAlignmentParameters *pairwiseAlignmentBandingParameters_construct(void);
void pairwiseAlignmentBandingParameters_destruct(PairwiseAlignmentParameters *p);

/* ---- the path (inc/pairwiseAligner.h:384-392) ------------------------------- */
typedef struct _dpMatrix DpMatrix;

/* Names the posterior callback of getPosteriorProbsWithBanding; the thresholded
 * posteriors run on the device, so calling it directly aborts. */
void diagonalCalculationPosteriorMatchProbs(StateMachine *sM, int64_t xay, DpMatrix *forwardDpMatrix,
                                            DpMatrix *backwardDpMatrix, Sequence *sX, Sequence *sY,
                                            double totalProbability, PairwiseAlignmentParameters *p,
                                            void *extraArgs);

/* Replaces impl/pairwiseAligner.c:1871-1899.  anchorPairs: stIntTuple (x, y),
 * 0-based k-mer and event indices, strictly increasing.  Returns a new list of
 * stIntTuple (floor(p * 1e7), x, y, (int64_t) path k-mer string) in the
 * reference's order; the k-mer strings live until the list is destructed.
 * Only diagonalCalculationPosteriorMatchProbs is accepted as the callback. */
stList *getAlignedPairsUsingAnchors(
    StateMachine *sM, Sequence *SsX, Sequence *SsY, stList *anchorPairs, PairwiseAlignmentParameters *p,
    void (*diagonalPosteriorProbFn)(StateMachine *, int64_t, DpMatrix *, DpMatrix *, Sequence *, Sequence *, double,
                                    PairwiseAlignmentParameters *, void *),
    bool alignmentHasRaggedLeftEnd, bool alignmentHasRaggedRightEnd);

/* ---- EM expectations (inc/stateMachine.h:49-74, inc/continuousHmm.h) ------- */
/* The reference's Hmm header; the expectations HMMs below extend it. */
typedef struct _hmm Hmm;
struct _hmm {
    double likelihood;
    StateMachineType type;
    int64_t stateNumber;
    int64_t parameterSetSize;
    int64_t matrixSize;
    char *alphabet;
    int64_t alphabetSize;
    int64_t kmerLength;
    double *transitions;  /* stateNumber^2: from * stateNumber + to */
    void (*addToTransitionExpectationFcn)(Hmm *hmm, int64_t from, int64_t to, double p);
    void (*setTransitionFcn)(Hmm *hmm, int64_t from, int64_t to, double p);
    double (*getTransitionsExpFcn)(Hmm *hmm, int64_t from, int64_t to);
};

/* impl/continuousHmm.c:842-855: a ContinuousPairHmm (threeState: transition and
 * per-k-mer event expectations, pseudocounts as given) or an HdpHmm
 * (threeStateHdp: transition expectations and event assignments whose match
 * posterior reaches `threshold`), with the machine's scale, shift, var and match
 * table. */
Hmm *hmmContinuous_getExpectationsHmm(StateMachine *sM, double threshold, double transitionsPseudocount,
                                      double emissionsPseudocount);
/* Names the expectations callback of getPosteriorProbsWithBanding; the
 * expectations run on the device, so calling it directly aborts. */
void diagonalCalculation_Expectations(StateMachine *sM, int64_t xay, DpMatrix *forwardDpMatrix,
                                      DpMatrix *backwardDpMatrix, Sequence *sX, Sequence *sY, double totalProbability,
                                      PairwiseAlignmentParameters *p, void *extraArgs);
/* Replaces impl/pairwiseAligner.c:1983-2003 (with diagonalCalculation_Expectations):
 * adds this read's expectations to hmmExpectations, in the reference's order. */
void getExpectationsUsingAnchors(StateMachine *sM, Hmm *hmmExpectations, Sequence *SsX, Sequence *SsY,
                                 stList *anchorPairs, PairwiseAlignmentParameters *p,
                                 void (*diagonalCalcExpectationFcn)(StateMachine *, int64_t, DpMatrix *, DpMatrix *,
                                                                    Sequence *, Sequence *, double,
                                                                    PairwiseAlignmentParameters *, void *),
                                 bool alignmentHasRaggedLeftEnd, bool alignmentHasRaggedRightEnd);
int64_t hmmContinuous_howManyAssignments(Hmm *hmm);
/* impl/continuousHmm.c:872-883 (continuousPairHmm_writeToFile, hdpHmm_writeToFile) */
void hmmContinuous_writeToFile(const char *outFile, Hmm *hmm, StateMachineType type);
void hmmContinuous_destruct(Hmm *hmm, StateMachineType type);

/* ---- additions with no counterpart in the reference ------------------------- */
/* Device for the process-wide context (before the first alignment; default 0). */
int sa_compat_set_device(int device);
/* FP64 (default, matches the reference) or FP32 recurrence for later calls. */
int sa_compat_set_precision(int precision);
/* Total log-likelihood of the last getAlignedPairsUsingAnchors call on this thread. */
double sa_compat_last_log_likelihood(void);

#ifdef __cplusplus
}
#endif

#endif
