/*
 * signalalign_machine.h -- the per-read driver around the banded forward-backward
 * path: signalMachine's inputs, read scaling and outputs, in the reference's C API.
 *
 * libsignalalign_compat.so exports these next to getAlignedPairsUsingAnchors, and
 * the signalMachine executable built beside the library is the reference's
 * signalMachine (same options, exonerate CIGAR on stdin, same TSV files), so the
 * reference's pipeline can call it unchanged.
 *
 *   .npRead reader            nanopore_loadNanoporeReadFromFile   impl/nanopore.c:98-424
 *   one-D assignments         nanopore_getOneDAssignmentsFromRead impl/nanopore.c:474-533
 *   read scaling              nanopore_compute_mean_scale_params  impl/nanopore.c:586-781
 *                             nanopore_compute_noise_scale_params impl/nanopore.c:783-848
 *                             signalUtils_estimateNanoporeParams  impl/signalMachineUtils.c:186-223
 *   drift                     nanopore_adjust*EventsForDrift      impl/nanopore.c:535-560
 *   guide alignment           cigarRead (sonLib), convertPairwiseForwardStrandAlignmentToAnchorPairs
 *                             (impl/pairwiseAligner.c:1454-1478), filterToRemoveOverlap (:1575-1616),
 *                             signalUtils_guideAlignmentToRebasedAnchorPairs (impl/signalMachineUtils.c:130-163)
 *   event remapping           nanopore_remapAnchorPairsWithOffset impl/nanopore.c:438-450
 *   reference slices          signalUtils_ReferenceSequenceConstructFull impl/signalMachineUtils.c:20-86
 *   HDP models (.nhdp)        deserialize_nhdp                    impl/nanopore_hdp.c:1071-1098
 *   TSV writers               writePosteriorProbsFull / VC / writeAssignments  signalMachine.c:68-260
 *   output order              sortByXPlusYCoordinate2             impl/pairwiseAligner.c:1434-1438
 *
 * Errors abort with a message, as the reference's st_errAbort does.
 */
#ifndef SIGNALALIGN_MACHINE_H
#define SIGNALALIGN_MACHINE_H

#include <stdio.h>

#include "signalalign_compat.h"

#ifdef __cplusplus
extern "C" {
#endif

#ifndef MACHEP
#define MACHEP 1.11022302462515654042E-16
#endif

/* ---- sonLib additions ------------------------------------------------------ */
/* Stable sort of the list's items (the comparator receives two items). */
void stList_sort(stList *list, int (*cmpFn)(const void *a, const void *b));
/* Lexicographic order of two tuples, the shorter first on a tie. */
int stIntTuple_cmpFn(stIntTuple *a, stIntTuple *b);

/* sonLib's pairwise alignment (its pairwiseAlignment.h): sequence 1 is the
 * reference, sequence 2 the read; strands are 1 for '+', 0 for '-'. */
#define PAIRWISE_MATCH 0
#define PAIRWISE_INDEL_X 1
#define PAIRWISE_INDEL_Y 2

struct List {
    int64_t length;
    int64_t maxLength;
    void **list;
    void (*destructElement)(void *);
};

struct AlignmentOperation {
    int64_t opType;
    int64_t length;
    double score;
};

struct PairwiseAlignment {
    char *contig1;
    int64_t start1;
    int64_t end1;
    int64_t strand1;
    char *contig2;
    int64_t start2;
    int64_t end2;
    int64_t strand2;
    double score;
    struct List *operationList;
};

/* One exonerate "cigar:" line: read name, start, end, strand, reference name,
 * start, end, strand, score, then operation/length pairs (M, I = read only,
 * D = reference only).  NULL at end of input. */
struct PairwiseAlignment *cigarRead(FILE *fileHandle);
/* The same from a string. */
struct PairwiseAlignment *cigarReadFromString(const char *line);
void checkPairwiseAlignment(struct PairwiseAlignment *pA);
void destructPairwiseAlignment(struct PairwiseAlignment *pA);

/* ---- reads (inc/nanopore.h:21-60) ------------------------------------------- */
typedef struct _nanoporeRead {
    int64_t readLength;
    int64_t templateReadLength;
    int64_t complementReadLength;
    int64_t nbTemplateEvents;
    int64_t nbComplementEvents;
    NanoporeReadAdjustmentParameters templateParams;
    NanoporeReadAdjustmentParameters complementParams;

    char *twoDread;
    char *templateRead;
    char *complementRead;

    int64_t *templateEventMap;
    int64_t *templateStrandEventMap;
    double *templateEvents; /* mean, stdev, length, start time */

    int64_t *complementEventMap;
    int64_t *complementStrandEventMap;
    double *complementEvents;

    int64_t *templateModelState;
    double *templatePModel;

    int64_t *complementModelState;
    double *complementPModel;

    bool scaled;
    int twoD;
} NanoporeRead;

typedef struct _eventKmerTuple {
    double eventMean;
    double eventSd;
    double deltaTime;
    int64_t kmerIndex;
} EventKmerTuple;

NanoporeRead *nanopore_loadNanoporeReadFromFile(const char *nanoporeReadFile);
void nanopore_nanoporeReadDestruct(NanoporeRead *npRead);
NanoporeReadAdjustmentParameters *nanopore_readAdjustmentParametersConstruct(void);
EventKmerTuple *nanopore_eventKmerTupleConstruct(double mean, double sd, double deltaTime, int64_t kmerIndex);

/* kmer_id over a sorted alphabet (impl/nanopore_hdp.c:360-393); aborts on a
 * character outside it. */
int64_t kmer_id(const char *kmer, const char *alphabet, int64_t alphabetSize, int64_t kmerLength);

/* ---- read scaling ----------------------------------------------------------- */
stList *nanopore_getOneDAssignmentsFromRead(int64_t *strandEventMap, double *events, char *strandRead,
                                            int64_t readLength, char *alphabet, int64_t alphabetSize,
                                            int64_t kmerLength);
stList *nanopore_getTemplateOneDAssignments(NanoporeRead *npRead, double threshold);
stList *nanopore_getComplementOneDAssignments(NanoporeRead *npRead, double threshold);
/* Solves A x = b (n x n, row-major) by Gauss-Jordan elimination; aborts when A
 * is singular. */
void nanopore_lineq_solve(double *A, double *b, double *x_out, int64_t n);
void nanopore_compute_mean_scale_params(double *model, stList *kmerToEventMap,
                                        NanoporeReadAdjustmentParameters *params, bool drift_out, bool var_out);
void nanopore_compute_noise_scale_params(double *model, stList *kmerToEventMap,
                                         NanoporeReadAdjustmentParameters *params);
void nanopore_adjustEventsForDrift(NanoporeRead *npRead);
void nanopore_adjustTemplateEventsForDrift(NanoporeRead *npRead);
void nanopore_adjustComplementEventsForDrift(NanoporeRead *npRead);
void nanopore_dontAdjustEvents(NanoporeRead *npRead);
/* emissions_signal_scaleNoise (impl/stateMachine.c:609-629) on a loaded machine. */
void emissions_signal_scaleNoise(StateMachine *sM, NanoporeReadAdjustmentParameters npp);

stList *signalUtils_templateOneDAssignmentsFromRead(NanoporeRead *npRead, StateMachine *sM, double ignore);
stList *signalUtils_complementOneDAssignmentsFromRead(NanoporeRead *npRead, StateMachine *sM, double ignore);
void signalUtils_estimateNanoporeParams(StateMachine *sM, NanoporeRead *npRead,
                                        NanoporeReadAdjustmentParameters *params, double assignmentThreshold,
                                        stList *(*assignmentFunction)(NanoporeRead *, StateMachine *, double),
                                        void (*driftAdjustmentFunction)(NanoporeRead *));

/* ---- anchors ---------------------------------------------------------------- */
stList *convertPairwiseForwardStrandAlignmentToAnchorPairs(struct PairwiseAlignment *pA, int64_t trim);
stList *filterToRemoveOverlap(stList *sortedOverlappingPairs);
/* Rebases pA to reference position 0 on the forward strand (pA is modified),
 * trims and filters its match blocks into (reference, read base) anchors. */
stList *signalUtils_guideAlignmentToRebasedAnchorPairs(struct PairwiseAlignment *pA, PairwiseAlignmentParameters *p);
stList *nanopore_remapAnchorPairsWithOffset(stList *unmappedPairs, int64_t *eventMap, int64_t mapOffset);
stList *signalUtils_getRemappedAnchorPairs(stList *unmappedAnchors, int64_t *eventMap, int64_t mapOffset);

/* ---- reference slices -------------------------------------------------------- */
typedef struct _referenceSequence ReferenceSequence;
struct _referenceSequence {
    char *reference;
    char *complementOfReference;
    char *trimmedForwardSequence;
    char *trimmedBackwardSequence;
    struct PairwiseAlignment *A;
    char *(*getTemplateTargetSequence)(ReferenceSequence *R);
    char *(*getComplementTargetSequence)(ReferenceSequence *R);
    bool initialized;
};

char *signalUtils_stringReverse(char *str);
char *signalUtils_getSubSequence(char *seq, int64_t start, int64_t end, bool strand);
ReferenceSequence *signalUtils_ReferenceSequenceConstructFull(char *forwardReferencePath, char *backwardReferencePath,
                                                              struct PairwiseAlignment *pA);
void signalUtils_ReferenceSequenceDestruct(ReferenceSequence *self);

/* ---- HDP models --------------------------------------------------------------- */
/* A serialised nanopore HDP (.nhdp), as far as alignment reads it: the sampling
 * grid, each observed Dirichlet process's posterior predictive density and spline
 * slopes on it, and the DP tree (deserialize_nhdp, impl/nanopore_hdp.c:1071-1098,
 * around deserialize_hdp, impl/hdp.c:3047-3317; the Gibbs-sampling state is
 * skipped).  DP ids of the leaves are kmer_id over the sorted alphabet. */
typedef struct _nanoporeHDP {
    char *alphabet; /* sorted */
    int64_t alphabet_size;
    int64_t kmer_length;
    int64_t num_dps;
    int64_t grid_length;
    double *grid;
    int64_t *parent;    /* -1 at the root */
    bool *observed;     /* data assigned to the DP or a descendant */
    int64_t *row;       /* row of an observed DP in density/slopes, else -1 */
    int64_t n_rows;
    double *density;    /* n_rows x grid_length */
    double *slopes;
} NanoporeHDP;

NanoporeHDP *deserialize_nhdp(const char *filepath);
void destroy_nanopore_hdp(NanoporeHDP *nhdp);
/* dir_proc_expected_val / dir_proc_variance (impl/hdp.c:2772-2801) of a DP. */
double nhdp_dp_expected_value(NanoporeHDP *nhdp, int64_t dp);
double nhdp_dp_variance(NanoporeHDP *nhdp, int64_t dp);
/* impl/stateMachine.c:1655-1669: the .model file's machine with the HDP's
 * densities for emissions (k-mers read their first observed ancestor's). */
StateMachine *getHdpStateMachine(NanoporeHDP *hdp, const char *modelFile, NanoporeReadAdjustmentParameters npp);
/* impl/stateMachine.c:1163-1192: level mean and sd of every observed k-mer set
 * to its DP's expected value and standard deviation. */
void stateMachine3_setModelToHdpExpectedValues(StateMachine *sM, NanoporeHDP *nhdp);

/* ---- alignment and output (signalMachine.c) ----------------------------------- */
/* inc/stateMachine.h:43-46 ("template" is a C++ keyword, hence the C++ names) */
#ifdef __cplusplus
typedef enum { SA_TEMPLATE_STRAND = 0, SA_COMPLEMENT_STRAND = 1 } Strand;
#else
typedef enum _strand { template = 0, complement = 1 } Strand;
#endif
typedef enum { full = 0, variantCaller = 1, assignments = 2 } OutputFormat;

Sequence *makeEventSequenceFromPairwiseAlignment(double *events, int64_t queryStart, int64_t queryEnd,
                                                 int64_t *eventMap);
stList *performSignalAlignment(StateMachine *sM, Sequence *eventSequence, int64_t *eventMap, int64_t mapOffset,
                               char *target, PairwiseAlignmentParameters *p, stList *unmappedAnchors,
                               DegenerateType degenerate);
/* signalMachine.c:371-394: the read's expectations added to hmmExpectations
 * (anchors remapped to the events, the target as a k-mer sequence). */
void getSignalExpectations(StateMachine *sM, Hmm *hmmExpectations, Sequence *eventSequence, int64_t *eventMap,
                           int64_t mapOffset, char *trainingTarget, PairwiseAlignmentParameters *p,
                           stList *unmappedAnchors, DegenerateType degenerate);
double scoreByPosteriorProbabilityIgnoringGaps(stList *alignedPairs);
int sortByXPlusYCoordinate2(const void *i, const void *j);
/* Positions of degenerate ('X') bases in a k-mer (impl/pairwiseAligner.c:510-524),
 * as a list of int64_t*. */
stList *path_findDegeneratePositions(char *kmer, int64_t kmerLength);

void writePosteriorProbsFull(char *posteriorProbsFile, char *readLabel, StateMachine *sM,
                             NanoporeReadAdjustmentParameters npp, double *events, char *target, bool forward,
                             char *contig, int64_t eventSequenceOffset, int64_t referenceSequenceOffset,
                             stList *alignedPairs, Strand strand);
void writePosteriorProbsVC(char *posteriorProbsFile, char *readLabel, StateMachine *sM, char *target, bool forward,
                           int64_t eventSequenceOffset, int64_t referenceSequenceOffset, stList *alignedPairs,
                           Strand strand);
void writeAssignments(char *posteriorProbsFile, StateMachine *sM, double *events, int64_t eventSequenceOffset,
                      NanoporeReadAdjustmentParameters npp, stList *alignedPairs, Strand strand);
void outputAlignment(OutputFormat fmt, char *posteriorProbsFile, char *readLabel, StateMachine *sM,
                     NanoporeReadAdjustmentParameters npp, double *events, char *target, bool forward, char *contig,
                     int64_t eventSequenceOffset, int64_t referenceSequenceOffset, stList *alignedPairs,
                     Strand strand);

#ifdef __cplusplus
}
#endif

#endif
