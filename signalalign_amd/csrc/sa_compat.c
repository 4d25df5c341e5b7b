/*
 * sa_compat.c -- signalAlign's pairwiseAligner C API (getAlignedPairsUsingAnchors
 * and the constructors its callers use) on top of the MI355X C-ABI.
 *
 * The state machine keeps the reference's tables on the host in the reference's
 * layout; on the first alignment (and whenever the tables change) they are
 * uploaded once as an sa_model and stay resident.  Each alignment is one sa_align
 * call; its pairs become the reference's stList of stIntTuple.
 */
#include "signalalign_compat.h"
#include "signalalign_machine.h"
#include "signalalign_amd.h"

#include <inttypes.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define LOG_ZERO_C (-INFINITY)
#define EXTRA_EVENT_NOISE_MULTIPLIER_C 1.75
#define LOG_GAP_X_C (-2.3025850929940455) /* log(0.1), impl/stateMachine.c:1472-1474 */

static void die(const char *where, const char *msg) {
    fprintf(stderr, "signalalign_amd: %s: %s\n", where, msg);
    abort();
}

static void *xmalloc(size_t n) {
    void *p = malloc(n ? n : 1);
    if (!p) die("malloc", "out of host memory");
    return p;
}

/* ------------------------------------------------------------------------ */
/* sonLib subset                                                             */
/* ------------------------------------------------------------------------ */
#ifndef SA_COMPAT_USE_SONLIB
struct _stList {
    void **items;
    int64_t length, capacity;
    void (*destructElement)(void *);
    void *arena; /* path k-mer strings of an alignment result */
};

struct _stIntTuple {
    int64_t length;
    int64_t v[4];
};

stList *stList_construct3(int64_t length, void (*destructElement)(void *)) {
    stList *l = (stList *)xmalloc(sizeof(stList));
    l->length = length;
    l->capacity = length > 8 ? length : 8;
    l->items = (void **)calloc((size_t)l->capacity, sizeof(void *));
    if (!l->items) die("stList_construct3", "out of host memory");
    l->destructElement = destructElement;
    l->arena = NULL;
    return l;
}

stList *stList_construct(void) { return stList_construct3(0, NULL); }

void stList_append(stList *l, void *item) {
    if (l->length == l->capacity) {
        l->capacity *= 2;
        void **n = (void **)realloc(l->items, sizeof(void *) * (size_t)l->capacity);
        if (!n) die("stList_append", "out of host memory");
        l->items = n;
    }
    l->items[l->length++] = item;
}

void *stList_get(stList *l, int64_t i) {
    if (i < 0 || i >= l->length) die("stList_get", "index out of range");
    return l->items[i];
}

int64_t stList_length(stList *l) { return l->length; }

void stList_destruct(stList *l) {
    if (!l) return;
    if (l->destructElement)
        for (int64_t i = 0; i < l->length; i++) l->destructElement(l->items[i]);
    free(l->items);
    free(l->arena);
    free(l);
}

static stIntTuple *tuple(int64_t n, int64_t a, int64_t b, int64_t c, int64_t d) {
    stIntTuple *t = (stIntTuple *)xmalloc(sizeof(stIntTuple));
    t->length = n;
    t->v[0] = a; t->v[1] = b; t->v[2] = c; t->v[3] = d;
    return t;
}
stIntTuple *stIntTuple_construct2(int64_t a, int64_t b) { return tuple(2, a, b, 0, 0); }
stIntTuple *stIntTuple_construct3(int64_t a, int64_t b, int64_t c) { return tuple(3, a, b, c, 0); }
stIntTuple *stIntTuple_construct4(int64_t a, int64_t b, int64_t c, int64_t d) { return tuple(4, a, b, c, d); }
int64_t stIntTuple_get(stIntTuple *t, int64_t i) {
    if (i < 0 || i >= t->length) die("stIntTuple_get", "index out of range");
    return t->v[i];
}
int64_t stIntTuple_length(stIntTuple *t) { return t->length; }
void stIntTuple_destruct(stIntTuple *t) { free(t); }

int stIntTuple_cmpFn(stIntTuple *a, stIntTuple *b) {
    const int64_t n = a->length < b->length ? a->length : b->length;
    for (int64_t i = 0; i < n; i++)
        if (a->v[i] != b->v[i]) return a->v[i] < b->v[i] ? -1 : 1;
    return a->length < b->length ? -1 : (a->length > b->length ? 1 : 0);
}

/* stable merge sort of the item pointers */
void stList_sort(stList *l, int (*cmpFn)(const void *a, const void *b)) {
    const int64_t n = l->length;
    if (n < 2) return;
    void **tmp = (void **)xmalloc(sizeof(void *) * (size_t)n);
    void **src = l->items, **dst = tmp;
    for (int64_t w = 1; w < n; w *= 2) {
        for (int64_t lo = 0; lo < n; lo += 2 * w) {
            const int64_t mid = lo + w < n ? lo + w : n, hi = lo + 2 * w < n ? lo + 2 * w : n;
            int64_t i = lo, j = mid, o = lo;
            while (i < mid && j < hi) dst[o++] = cmpFn(src[j], src[i]) < 0 ? src[j++] : src[i++];
            while (i < mid) dst[o++] = src[i++];
            while (j < hi) dst[o++] = src[j++];
        }
        void **t = src;
        src = dst;
        dst = t;
    }
    if (src != l->items) memcpy(l->items, src, sizeof(void *) * (size_t)n);
    free(tmp);
}

static void attach_arena(stList *l, void *arena) { l->arena = arena; }
#else
/* sonLib's stList has no slot for the strings: they are kept, as the reference
 * keeps its path k-mers, for the life of the process. */
static void attach_arena(stList *l, void *arena) { (void)l; (void)arena; }
#endif

/* ------------------------------------------------------------------------ */
/* sequences (impl/pairwiseAligner.c:325-471)                                */
/* ------------------------------------------------------------------------ */
static double g_null_event[NB_EVENT_PARAMS] = {0.0, 0.0, 0.0, 0.0};

int64_t sequence_correctSeqLength(int64_t length, SequenceType type, int64_t kmerLength) {
    if (length <= 0) return 0;
    return type == nucleotide ? length : length - (kmerLength - 1);
}

void *sequence_getKmer(void *elements, int64_t index) { return (char *)elements + (index >= 0 ? index : 0); }

void *sequence_getEvent(void *elements, int64_t index) {
    return index >= 0 ? (void *)((double *)elements + index * NB_EVENT_PARAMS) : (void *)g_null_event;
}

Sequence *sequence_construct(int64_t length, void *elements, void *(*getFcn)(void *, int64_t), SequenceType type) {
    Sequence *s = (Sequence *)xmalloc(sizeof(Sequence));
    s->length = length;
    s->type = type;
    s->elements = elements;
    s->degenerateBases = NULL;
    s->nbDegenerateBases = 0;
    s->get = getFcn;
    s->sliceFcn = NULL;
    return s;
}

Sequence *sequence_construct2(int64_t length, void *elements, void *(*getFcn)(void *, int64_t),
                              Sequence *(*sliceFcn)(Sequence *, int64_t, int64_t), SequenceType type) {
    Sequence *s = sequence_construct(length, elements, getFcn, type);
    s->sliceFcn = sliceFcn;
    return s;
}

Sequence *sequence_constructKmerSequence(int64_t length, void *elements, void *(*getFcn)(void *, int64_t),
                                         Sequence *(*sliceFcn)(Sequence *, int64_t, int64_t),
                                         char *nucleotideOptions, int64_t nbOptions, SequenceType type) {
    if (type != kmer) die("sequence_constructKmerSequence", "can only make a reference sequence of k-mers");
    Sequence *s = sequence_construct2(length, elements, getFcn, sliceFcn, type);
    s->degenerateBases = (char *)xmalloc((size_t)nbOptions + 1);
    memcpy(s->degenerateBases, nucleotideOptions, (size_t)nbOptions);
    s->degenerateBases[nbOptions] = '\0';
    s->nbDegenerateBases = nbOptions;
    return s;
}

Sequence *sequence_constructReferenceKmerSequence(int64_t length, void *elements,
                                                  void *(*getFcn)(void *, int64_t),
                                                  Sequence *(*sliceFcn)(Sequence *, int64_t, int64_t),
                                                  DegenerateType dType, SequenceType type) {
    /* sequence_getBaseOptions / sequence_nbBaseOptions (impl/pairwiseAligner.c:264-292) */
    const char *opts = dType == cytosineMethylation2 ? "CE"
                     : dType == cytosineMethylation3 ? "CEO"
                     : dType == adenosineMethylation ? "AI"
                                                     : "ACGT";
    return sequence_constructKmerSequence(length, elements, getFcn, sliceFcn, (char *)opts, (int64_t)strlen(opts),
                                          type);
}

Sequence *sequence_sliceNucleotideSequence(Sequence *in, int64_t start, int64_t sliceLength) {
    return sequence_constructKmerSequence(sliceLength, (char *)in->elements + start, in->get, in->sliceFcn,
                                          in->degenerateBases, in->nbDegenerateBases, in->type);
}

Sequence *sequence_sliceEventSequence(Sequence *in, int64_t start, int64_t sliceLength) {
    return sequence_construct2(sliceLength, (double *)in->elements + start * NB_EVENT_PARAMS, in->get, in->sliceFcn,
                               in->type);
}

Sequence *sequence_constructEventSequence(int64_t length, void *events) {
    return sequence_construct2(length, events, sequence_getEvent, sequence_sliceEventSequence, event);
}

void sequence_destruct(Sequence *seq) {
    if (!seq) return;
    free(seq->degenerateBases);
    free(seq);
}

/* ------------------------------------------------------------------------ */
/* parameters (impl/pairwiseAligner.c:1843-1860)                              */
/* ------------------------------------------------------------------------ */
PairwiseAlignmentParameters *pairwiseAlignmentBandingParameters_construct(void) {
    PairwiseAlignmentParameters *p = (PairwiseAlignmentParameters *)xmalloc(sizeof(*p));
    p->threshold = 0.01;
    p->minDiagsBetweenTraceBack = 1000;
    p->traceBackDiagonals = 40;
    p->diagonalExpansion = 20;
    p->constraintDiagonalTrim = 14;
    p->anchorMatrixBiggerThanThis = 500 * 500;
    p->repeatMaskMatrixBiggerThanThis = 500 * 500;
    p->splitMatrixBiggerThanThis = (int64_t)3000 * 3000;
    p->alignAmbiguityCharacters = false;
    p->gapGamma = 0.5f;
    return p;
}

void pairwiseAlignmentBandingParameters_destruct(PairwiseAlignmentParameters *p) { free(p); }

/* ------------------------------------------------------------------------ */
/* state machines                                                            */
/* ------------------------------------------------------------------------ */

/* The public prefix is the reference's StateMachine3; the rest is private. */
typedef struct {
    StateMachine3 sm3;
    int emission;           /* sa_emission_kind */
    sa_model *dev;          /* resident tables, built on first use */
    uint64_t dev_sum;       /* checksum of the host tables that were uploaded */
    /* HDP machine */
    double *grid, *density, *slopes, *level_mean;
    int64_t *kmer_row;
    int64_t grid_length, n_rows;
} CompatMachine;

/* impl/stateMachine.c:1022-1061 */
static double sm3_start(StateMachine *sM, int64_t state) { (void)sM; return state == 0 ? 0.0 : LOG_ZERO_C; }
static double sm3_ragged_start(StateMachine *sM, int64_t state) {
    (void)sM;
    return (state == 1 || state == 2) ? 0.0 : LOG_ZERO_C;
}
static double sm3_end(StateMachine *sM, int64_t state) {
    StateMachine3 *s = (StateMachine3 *)sM;
    return state == 0 ? s->TRANSITION_MATCH_CONTINUE
         : state == 1 ? s->TRANSITION_MATCH_FROM_GAP_X
         : state == 2 ? s->TRANSITION_MATCH_FROM_GAP_Y : 0.0;
}
static double sm3_ragged_end(StateMachine *sM, int64_t state) {
    StateMachine3 *s = (StateMachine3 *)sM;
    return state == 0 ? (s->TRANSITION_GAP_OPEN_X + s->TRANSITION_GAP_OPEN_Y) / 2.0
         : state == 1 ? s->TRANSITION_GAP_EXTEND_X
         : state == 2 ? s->TRANSITION_GAP_EXTEND_Y : 0.0;
}

static char *read_line(FILE *f) {
    size_t cap = 1 << 16, n = 0;
    char *buf = (char *)xmalloc(cap);
    int c;
    while ((c = fgetc(f)) != EOF && c != '\n') {
        if (n + 1 == cap) {
            cap *= 2;
            char *nb = (char *)realloc(buf, cap);
            if (!nb) die("read_line", "out of host memory");
            buf = nb;
        }
        buf[n++] = (char)c;
    }
    if (c == EOF && n == 0) {
        free(buf);
        return NULL;
    }
    buf[n] = '\0';
    return buf;
}

/* whitespace-separated doubles of one line; returns the count */
static int64_t parse_doubles(const char *line, double *out, int64_t cap) {
    int64_t n = 0;
    const char *p = line;
    for (;;) {
        while (*p == ' ' || *p == '\t' || *p == '\r') p++;
        if (!*p) break;
        char *end;
        double v = strtod(p, &end);
        if (end == p) return -1;
        if (n < cap) out[n] = v;
        n++;
        p = end;
    }
    return n;
}

static double log_or_zero(double v) { return v > 0.0 ? log(v) : LOG_ZERO_C; }

/* stateMachine3_loadFromFile (impl/stateMachine.c:1328-1426) with
 * stateMachine3_loadTransitionsFromFile (:1090-1146) and the defaults of
 * stateMachine3_construct (:1440-1480). */
static CompatMachine *load_machine(const char *modelFile, StateMachineType type) {
    FILE *f = fopen(modelFile, "r");
    if (!f) die("stateMachine3_loadFromFile", "cannot open the model file");
    char *l0 = read_line(f);
    if (!l0) die("stateMachine3_loadFromFile", "empty model file");
    long long stateNumber = 0, alphabetSize = 0, kmerLength = 0;
    char alphabet[64] = {0};
    if (sscanf(l0, "%lld %lld %63s %lld", &stateNumber, &alphabetSize, alphabet, &kmerLength) != 4)
        die("stateMachine3_loadFromFile", "model header must have 4 fields");
    free(l0);
    if (stateNumber != 3) die("stateMachine3_loadFromFile", "only three-state models are supported");
    if ((long long)strlen(alphabet) != alphabetSize) die("stateMachine3_loadFromFile", "alphabet size mismatch");
    if (kmerLength < 1 || kmerLength > 10) die("stateMachine3_loadFromFile", "bad k-mer length");

    CompatMachine *m = (CompatMachine *)calloc(1, sizeof(CompatMachine));
    if (!m) die("stateMachine3_loadFromFile", "out of host memory");
    StateMachine *sM = &m->sm3.model;
    sM->type = type;
    sM->stateNumber = 3;
    sM->matchState = 0;
    int64_t nk = 1;
    for (long long i = 0; i < kmerLength; i++) nk *= alphabetSize;
    sM->parameterSetSize = nk;
    sM->scale = 1.0;
    sM->shift = 0.0;
    sM->var = 1.0;
    /* sequence_prepareAlphabet: the alphabet sorted */
    sM->alphabet = (char *)xmalloc((size_t)alphabetSize + 1);
    memcpy(sM->alphabet, alphabet, (size_t)alphabetSize + 1);
    for (long long i = 1; i < alphabetSize; i++)
        for (long long j = i; j > 0 && sM->alphabet[j - 1] > sM->alphabet[j]; j--) {
            char t = sM->alphabet[j];
            sM->alphabet[j] = sM->alphabet[j - 1];
            sM->alphabet[j - 1] = t;
        }
    sM->alphabetSize = alphabetSize;
    sM->kmerLength = kmerLength;
    sM->EMISSION_MATCH_MATRIX = (double *)xmalloc(sizeof(double) * (size_t)(nk * MODEL_PARAMS));
    sM->EMISSION_GAP_Y_MATRIX = (double *)xmalloc(sizeof(double) * (size_t)(nk * MODEL_PARAMS));
    sM->EMISSION_GAP_X_PROBS = (double *)xmalloc(sizeof(double) * (size_t)nk);
    for (int64_t i = 0; i < nk; i++) sM->EMISSION_GAP_X_PROBS[i] = LOG_GAP_X_C;
    sM->startStateProb = sm3_start;
    sM->endStateProb = sm3_end;
    sM->raggedEndStateProb = sm3_ragged_end;
    sM->raggedStartStateProb = sm3_ragged_start;

    char *l1 = read_line(f);
    double tr[10];
    if (!l1 || parse_doubles(l1, tr, 10) != 10) die("stateMachine3_loadFromFile", "bad transitions line");
    free(l1);
    StateMachine3 *s3 = &m->sm3;
    s3->TRANSITION_MATCH_CONTINUE = log_or_zero(tr[0]);
    s3->TRANSITION_GAP_OPEN_X = log_or_zero(tr[1]);
    s3->TRANSITION_GAP_OPEN_Y = log_or_zero(tr[2]);
    s3->TRANSITION_MATCH_FROM_GAP_X = log_or_zero(tr[3]);
    s3->TRANSITION_GAP_EXTEND_X = log_or_zero(tr[4]);
    s3->TRANSITION_MATCH_FROM_GAP_Y = log_or_zero(tr[6]);
    s3->TRANSITION_GAP_SWITCH_TO_Y = log_or_zero(tr[7]);
    s3->TRANSITION_GAP_EXTEND_Y = log_or_zero(tr[8]);
    s3->TRANSITION_GAP_SWITCH_TO_X = LOG_ZERO_C;

    char *l2 = read_line(f);
    fclose(f);
    if (!l2 || parse_doubles(l2, sM->EMISSION_MATCH_MATRIX, nk * MODEL_PARAMS) != nk * MODEL_PARAMS)
        die("stateMachine3_loadFromFile", "match emissions line has the wrong number of values");
    free(l2);
    memcpy(sM->EMISSION_GAP_Y_MATRIX, sM->EMISSION_MATCH_MATRIX, sizeof(double) * (size_t)(nk * MODEL_PARAMS));
    for (int64_t i = 1; i < nk * MODEL_PARAMS; i += MODEL_PARAMS)
        sM->EMISSION_GAP_Y_MATRIX[i] *= EXTRA_EVENT_NOISE_MULTIPLIER_C;
    return m;
}

/* emissions_signal_scaleNoise (impl/stateMachine.c:609-629); the extra-event
 * noise sd uses the match table's lambda, as the reference does. */
void emissions_signal_scaleNoise(StateMachine *sM, NanoporeReadAdjustmentParameters npp) {
    double *M = sM->EMISSION_MATCH_MATRIX, *Y = sM->EMISSION_GAP_Y_MATRIX;
    for (int64_t i = 0; i < sM->parameterSetSize * MODEL_PARAMS; i += MODEL_PARAMS) {
        M[i + 2] *= npp.scale_sd;
        M[i + 4] *= npp.var_sd;
        M[i + 3] = sqrt(pow(M[i + 2], 3.0) / M[i + 4]);
        Y[i + 2] *= npp.scale_sd;
        Y[i + 4] *= npp.var_sd;
        Y[i + 3] = sqrt(pow(Y[i + 2], 3.0) / M[i + 4]);
    }
}

StateMachine *getStateMachine3_descaled(const char *modelFile, NanoporeReadAdjustmentParameters npp,
                                        bool scaleNoise) {
    CompatMachine *m = load_machine(modelFile, threeState);
    m->emission = SA_EMIT_GAUSS_DESCALED;
    m->sm3.model.scale = npp.scale;
    m->sm3.model.shift = npp.shift;
    m->sm3.model.var = npp.var;
    if (scaleNoise) emissions_signal_scaleNoise(&m->sm3.model, npp);
    return &m->sm3.model;
}

StateMachine *getStateMachine3(const char *modelFile) {
    CompatMachine *m = load_machine(modelFile, threeState);
    m->emission = SA_EMIT_GAUSS;
    return &m->sm3.model;
}

static void *dup_array(const void *src, size_t bytes) {
    void *p = xmalloc(bytes);
    memcpy(p, src, bytes);
    return p;
}

StateMachine *getHdpStateMachineFromGrid(const char *modelFile, const double *grid, int64_t grid_length,
                                         const double *density, const double *slopes, int64_t n_rows,
                                         const int64_t *kmer_row, const double *level_mean,
                                         NanoporeReadAdjustmentParameters npp) {
    if (grid_length < 2 || n_rows < 1) die("getHdpStateMachineFromGrid", "empty density grid");
    CompatMachine *m = load_machine(modelFile, threeStateHdp);
    const int64_t nk = m->sm3.model.parameterSetSize;
    m->emission = SA_EMIT_HDP;
    m->sm3.model.scale = npp.scale;
    m->sm3.model.shift = npp.shift;
    m->sm3.model.var = npp.var;
    m->grid_length = grid_length;
    m->n_rows = n_rows;
    m->grid = (double *)dup_array(grid, sizeof(double) * (size_t)grid_length);
    m->density = (double *)dup_array(density, sizeof(double) * (size_t)(grid_length * n_rows));
    m->slopes = (double *)dup_array(slopes, sizeof(double) * (size_t)(grid_length * n_rows));
    m->kmer_row = (int64_t *)dup_array(kmer_row, sizeof(int64_t) * (size_t)nk);
    /* the descaling level means live in the match table's mean column, as in the
       reference's HDP machine (emissions_signal_getModelLevelMean); NULL keeps the
       .model file's */
    if (level_mean)
        for (int64_t i = 0; i < nk; i++) m->sm3.model.EMISSION_MATCH_MATRIX[i * MODEL_PARAMS] = level_mean[i];
    m->level_mean = (double *)xmalloc(sizeof(double) * (size_t)nk);
    for (int64_t i = 0; i < nk; i++) m->sm3.model.EMISSION_GAP_X_PROBS[i] = LOG_GAP_X_C;
    return &m->sm3.model;
}

void stateMachine_destruct(StateMachine *sM) {
    if (!sM) return;
    CompatMachine *m = (CompatMachine *)sM;
    if (m->dev) sa_model_destroy(m->dev);
    free(sM->alphabet);
    free(sM->EMISSION_MATCH_MATRIX);
    free(sM->EMISSION_GAP_Y_MATRIX);
    free(sM->EMISSION_GAP_X_PROBS);
    free(m->grid);
    free(m->density);
    free(m->slopes);
    free(m->kmer_row);
    free(m->level_mean);
    free(m);
}

/* ------------------------------------------------------------------------ */
/* device context                                                            */
/* ------------------------------------------------------------------------ */
static pthread_mutex_t g_mu = PTHREAD_MUTEX_INITIALIZER;
static sa_context *g_ctx = NULL;
static int g_device = 0;
static sa_precision g_precision = SA_FP64;
static __thread double g_last_ll = 0.0;

int sa_compat_set_device(int device) {
    pthread_mutex_lock(&g_mu);
    int rc = g_ctx ? SA_ERR_INVALID_ARGUMENT : SA_OK;
    if (!g_ctx) g_device = device;
    pthread_mutex_unlock(&g_mu);
    return rc;
}

int sa_compat_set_precision(int precision) {
    if (precision != SA_FP64 && precision != SA_FP32) return SA_ERR_INVALID_ARGUMENT;
    pthread_mutex_lock(&g_mu);
    g_precision = (sa_precision)precision;
    pthread_mutex_unlock(&g_mu);
    return SA_OK;
}

double sa_compat_last_log_likelihood(void) { return g_last_ll; }

static sa_context *context(void) {
    if (!g_ctx) {
        int rc = sa_context_create(g_device, &g_ctx);
        if (rc != SA_OK) {
            g_ctx = NULL;
            die("getAlignedPairsUsingAnchors", "no usable GPU (sa_context_create failed)");
        }
    }
    return g_ctx;
}

static uint64_t fnv(uint64_t h, const void *p, size_t n) {
    const unsigned char *b = (const unsigned char *)p;
    for (size_t i = 0; i < n; i++) {
        h ^= b[i];
        h *= 1099511628211ull;
    }
    return h;
}

static uint64_t table_sum(const CompatMachine *m) {
    const StateMachine *sM = &m->sm3.model;
    const size_t nk = (size_t)sM->parameterSetSize;
    uint64_t h = 1469598103934665603ull;
    h = fnv(h, &m->sm3.TRANSITION_MATCH_CONTINUE, sizeof(double) * 9);
    h = fnv(h, sM->EMISSION_MATCH_MATRIX, sizeof(double) * nk * MODEL_PARAMS);
    h = fnv(h, sM->EMISSION_GAP_Y_MATRIX, sizeof(double) * nk * MODEL_PARAMS);
    h = fnv(h, sM->EMISSION_GAP_X_PROBS, sizeof(double) * nk);
    return h;
}

/* the machine's tables resident on the device (uploaded again after a change) */
static sa_model *device_model(sa_context *ctx, CompatMachine *m) {
    const uint64_t sum = table_sum(m);
    if (m->dev && m->dev_sum == sum) return m->dev;
    if (m->dev) sa_model_destroy(m->dev);
    m->dev = NULL;
    const StateMachine *sM = &m->sm3.model;
    sa_model_desc d;
    memset(&d, 0, sizeof(d));
    d.kind = (sa_emission_kind)m->emission;
    d.kmer_length = sM->kmerLength;
    d.alphabet = sM->alphabet;
    const double *t = &m->sm3.TRANSITION_MATCH_CONTINUE; /* the nine slots in StateMachine3 order */
    for (int i = 0; i < 9; i++) d.transitions[i] = t[i];
    if (m->emission == SA_EMIT_HDP) {
        d.grid_length = m->grid_length;
        d.grid = m->grid;
        d.n_rows = m->n_rows;
        d.density = m->density;
        d.slopes = m->slopes;
        d.kmer_row = m->kmer_row;
        for (size_t i = 0; i < (size_t)sM->parameterSetSize; i++)
            m->level_mean[i] = sM->EMISSION_MATCH_MATRIX[i * MODEL_PARAMS];
        d.level_mean = m->level_mean;
    } else {
        d.match = sM->EMISSION_MATCH_MATRIX;
        d.gapy = sM->EMISSION_GAP_Y_MATRIX;
        d.gapx = sM->EMISSION_GAP_X_PROBS;
    }
    int rc = sa_model_create(ctx, &d, &m->dev);
    if (rc != SA_OK) die("getAlignedPairsUsingAnchors", sa_last_error(ctx));
    m->dev_sum = sum;
    return m->dev;
}

/* ------------------------------------------------------------------------ */
/* the path                                                                  */
/* ------------------------------------------------------------------------ */
void diagonalCalculationPosteriorMatchProbs(StateMachine *sM, int64_t xay, DpMatrix *forwardDpMatrix,
                                            DpMatrix *backwardDpMatrix, Sequence *sX, Sequence *sY,
                                            double totalProbability, PairwiseAlignmentParameters *p,
                                            void *extraArgs) {
    (void)sM; (void)xay; (void)forwardDpMatrix; (void)backwardDpMatrix; (void)sX; (void)sY;
    (void)totalProbability; (void)p; (void)extraArgs;
    die("diagonalCalculationPosteriorMatchProbs", "posteriors are computed on the device; pass this function to "
                                                   "getAlignedPairsUsingAnchors instead of calling it");
}

/* the C-ABI read and parameters of one call (anchors copied into *anc, freed by
   the caller) */
static void make_read(const char *where, StateMachine *sM, Sequence *SsX, Sequence *SsY, stList *anchorPairs,
                      PairwiseAlignmentParameters *p, bool raggedLeft, bool raggedRight, sa_read *r,
                      sa_params *prm, int64_t **anc_out) {
    if (!sM || !SsX || !SsY || !p) die(where, "null argument");
    if (sM->type != threeState && sM->type != threeStateHdp) die(where, "needs a three-state machine");
    if (SsX->type != kmer) die(where, "SsX must be a k-mer sequence");
    if (p->diagonalExpansion % 2 != 0) die(where, "diagonalExpansion must be even");
    const int64_t k = sM->kmerLength;
    const int64_t lX = SsX->length, lY = SsY->length;
    const int64_t na = anchorPairs ? stList_length(anchorPairs) : 0;
    int64_t *anc = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(2 * na));
    for (int64_t i = 0; i < na; i++) {
        stIntTuple *t = (stIntTuple *)stList_get(anchorPairs, i);
        anc[2 * i] = stIntTuple_get(t, 0);
        anc[2 * i + 1] = stIntTuple_get(t, 1);
    }
    memset(r, 0, sizeof(*r));
    r->ref = (const char *)SsX->elements;
    r->ref_length = lX > 0 ? lX + k - 1 : 0;
    r->events = (const double *)SsY->elements;
    r->n_events = lY;
    r->event_stride = NB_EVENT_PARAMS;
    r->anchors = anc;
    r->n_anchors = na;
    r->scale = sM->scale;
    r->shift = sM->shift;
    r->var = sM->var;
    r->ragged_left = raggedLeft;
    r->ragged_right = raggedRight;
    /* a reference with 'X' takes the Sequence's degenerate base options */
    if (r->ref_length > 0 && memchr(r->ref, 'X', (size_t)r->ref_length)) {
        if (!SsX->degenerateBases || !SsX->degenerateBases[0]) die(where, "'X' in a sequence without degenerate bases");
        r->degenerate_options = SsX->degenerateBases;
    }
    prm->threshold = p->threshold;
    prm->min_diags_between_traceback = p->minDiagsBetweenTraceBack;
    prm->traceback_diagonals = p->traceBackDiagonals;
    prm->diagonal_expansion = p->diagonalExpansion;
    prm->split_matrix_bigger_than_this = p->splitMatrixBiggerThanThis;
    *anc_out = anc;
}

stList *getAlignedPairsUsingAnchors(
    StateMachine *sM, Sequence *SsX, Sequence *SsY, stList *anchorPairs, PairwiseAlignmentParameters *p,
    void (*diagonalPosteriorProbFn)(StateMachine *, int64_t, DpMatrix *, DpMatrix *, Sequence *, Sequence *, double,
                                    PairwiseAlignmentParameters *, void *),
    bool alignmentHasRaggedLeftEnd, bool alignmentHasRaggedRightEnd) {
    static const char *where = "getAlignedPairsUsingAnchors";
    if (diagonalPosteriorProbFn != diagonalCalculationPosteriorMatchProbs)
        die(where, "only diagonalCalculationPosteriorMatchProbs is supported");
    sa_read r;
    sa_params prm;
    int64_t *anc = NULL;
    make_read(where, sM, SsX, SsY, anchorPairs, p, alignmentHasRaggedLeftEnd, alignmentHasRaggedRightEnd, &r, &prm,
              &anc);
    CompatMachine *m = (CompatMachine *)sM;
    const int64_t k = sM->kmerLength;
    const int64_t lX = SsX->length;

    pthread_mutex_lock(&g_mu);
    sa_context *ctx = context();
    sa_model *dm = device_model(ctx, m);
    sa_result *res = NULL;
    /* degenerate references run in FP64 */
    int rc = sa_align(ctx, dm, &r, &prm, r.degenerate_options ? SA_FP64 : g_precision, &res);
    if (rc != SA_OK) {
        fprintf(stderr, "signalalign_amd: %s: status %d: %s\n", where, rc, sa_last_error(ctx));
        abort();
    }
    pthread_mutex_unlock(&g_mu);
    free(anc);

    sa_read_result rr;
    if (sa_result_read(res, 0, &rr) != SA_OK) die(where, "no result");
    g_last_ll = rr.log_likelihood;
    stList *out = stList_construct3(0, (void (*)(void *))stIntTuple_destruct);
    /* one NUL-terminated string per reference k-mer: the path k-mer of a
       single-path cell is the reference k-mer at x; a degenerate reference's pairs
       each get their path's k-mer (the first 'X' varying slowest) */
    char *arena = NULL;
    if (rr.n_pairs > 0 && !rr.paths) {
        arena = (char *)xmalloc((size_t)(lX * (k + 1)));
        for (int64_t x = 0; x < lX; x++) {
            memcpy(arena + x * (k + 1), (const char *)SsX->elements + x, (size_t)k);
            arena[x * (k + 1) + k] = '\0';
        }
    } else if (rr.n_pairs > 0) {
        arena = (char *)xmalloc((size_t)(rr.n_pairs * (k + 1)));
        const int64_t nopt = (int64_t)strlen(SsX->degenerateBases);
        for (int64_t i = 0; i < rr.n_pairs; i++) {
            char *dst = arena + i * (k + 1);
            memcpy(dst, (const char *)SsX->elements + rr.pairs[i].x, (size_t)k);
            dst[k] = '\0';
            int64_t rem = rr.paths[i];
            for (int64_t j = k - 1; j >= 0; j--)
                if (dst[j] == 'X') {
                    dst[j] = SsX->degenerateBases[rem % nopt];
                    rem /= nopt;
                }
        }
    }
    for (int64_t i = 0; i < rr.n_pairs; i++) {
        const sa_pair *q = &rr.pairs[i];
        const char *kmer = rr.paths ? arena + i * (k + 1) : arena + (int64_t)q->x * (k + 1);
        stList_append(out, stIntTuple_construct4(q->p, q->x, q->y, (int64_t)(intptr_t)kmer));
    }
    attach_arena(out, arena);
    sa_result_free(res);
    return out;
}

/* ------------------------------------------------------------------------ */
/* EM expectations (impl/pairwiseAligner.c:1983-2003, impl/continuousHmm.c)  */
/* ------------------------------------------------------------------------ */
/* The public prefix is the reference's Hmm; the rest holds a ContinuousPairHmm's
 * accumulators (threeState) or an HdpHmm's assignments (threeStateHdp). */
typedef struct {
    Hmm base;
    double scale, shift, var;
    double *eventModel;          /* MODEL_PARAMS per k-mer: the machine's match table */
    double *eventExpectations;   /* 2 per k-mer: sum p x~, running sum p (x~ - u)^2 */
    double *posteriors;          /* per k-mer */
    bool *observed;
    double threshold;            /* HDP: assignment threshold */
    char *assignKmers;           /* HDP: kmerLength characters per assignment */
    double *assignEvents;
    int64_t nAssign, capAssign;
} ExpectHmm;

static void add_transition_expectation(Hmm *hmm, int64_t from, int64_t to, double p) {
    hmm->transitions[from * hmm->stateNumber + to] += p;
}
static void set_transition_expectation(Hmm *hmm, int64_t from, int64_t to, double p) {
    hmm->transitions[from * hmm->stateNumber + to] = p;
}
static double get_transition_expectation(Hmm *hmm, int64_t from, int64_t to) {
    return hmm->transitions[from * hmm->stateNumber + to];
}

Hmm *hmmContinuous_getExpectationsHmm(StateMachine *sM, double threshold, double transitionsPseudocount,
                                      double emissionsPseudocount) {
    static const char *where = "hmmContinuous_getExpectationsHmm";
    if (!sM || (sM->type != threeState && sM->type != threeStateHdp)) die(where, "needs a three-state machine");
    ExpectHmm *h = (ExpectHmm *)calloc(1, sizeof(ExpectHmm));
    if (!h) die(where, "out of host memory");
    const int64_t nk = sM->parameterSetSize;
    h->base.type = sM->type;
    h->base.stateNumber = sM->stateNumber;
    h->base.parameterSetSize = nk;
    h->base.matrixSize = MODEL_PARAMS;
    h->base.likelihood = 0.0;
    h->base.alphabet = (char *)dup_array(sM->alphabet, (size_t)sM->alphabetSize + 1);
    h->base.alphabetSize = sM->alphabetSize;
    h->base.kmerLength = sM->kmerLength;
    h->base.transitions = (double *)xmalloc(sizeof(double) * (size_t)(sM->stateNumber * sM->stateNumber));
    for (int64_t i = 0; i < sM->stateNumber * sM->stateNumber; i++) h->base.transitions[i] = transitionsPseudocount;
    h->base.addToTransitionExpectationFcn = add_transition_expectation;
    h->base.setTransitionFcn = set_transition_expectation;
    h->base.getTransitionsExpFcn = get_transition_expectation;
    h->scale = sM->scale;
    h->shift = sM->shift;
    h->var = sM->var;
    h->eventModel = (double *)dup_array(sM->EMISSION_MATCH_MATRIX, sizeof(double) * (size_t)(nk * MODEL_PARAMS));
    h->threshold = threshold;
    if (sM->type == threeState) {
        h->eventExpectations = (double *)calloc((size_t)(2 * nk), sizeof(double));
        h->posteriors = (double *)xmalloc(sizeof(double) * (size_t)nk);
        h->observed = (bool *)calloc((size_t)nk, sizeof(bool));
        if (!h->eventExpectations || !h->observed) die(where, "out of host memory");
        for (int64_t i = 0; i < nk; i++) h->posteriors[i] = emissionsPseudocount;
    }
    return &h->base;
}

void diagonalCalculation_Expectations(StateMachine *sM, int64_t xay, DpMatrix *forwardDpMatrix,
                                      DpMatrix *backwardDpMatrix, Sequence *sX, Sequence *sY, double totalProbability,
                                      PairwiseAlignmentParameters *p, void *extraArgs) {
    (void)sM; (void)xay; (void)forwardDpMatrix; (void)backwardDpMatrix; (void)sX; (void)sY;
    (void)totalProbability; (void)p; (void)extraArgs;
    die("diagonalCalculation_Expectations", "expectations are computed on the device; pass this function to "
                                             "getExpectationsUsingAnchors instead of calling it");
}

static int64_t kmer_index(const char *kmer, const char *alphabet, int64_t A, int64_t k) {
    int64_t id = 0;
    for (int64_t i = 0; i < k; i++) {
        const char *hit = memchr(alphabet, kmer[i], (size_t)A);
        if (!hit) die("getExpectationsUsingAnchors", "k-mer character outside the alphabet");
        id = id * A + (int64_t)(hit - alphabet);
    }
    return id;
}

void getExpectationsUsingAnchors(StateMachine *sM, Hmm *hmmExpectations, Sequence *SsX, Sequence *SsY,
                                 stList *anchorPairs, PairwiseAlignmentParameters *p,
                                 void (*diagonalCalcExpectationFcn)(StateMachine *, int64_t, DpMatrix *, DpMatrix *,
                                                                    Sequence *, Sequence *, double,
                                                                    PairwiseAlignmentParameters *, void *),
                                 bool alignmentHasRaggedLeftEnd, bool alignmentHasRaggedRightEnd) {
    static const char *where = "getExpectationsUsingAnchors";
    if (diagonalCalcExpectationFcn != diagonalCalculation_Expectations)
        die(where, "only diagonalCalculation_Expectations is supported");
    if (!hmmExpectations || hmmExpectations->type != sM->type) die(where, "the HMM does not fit the machine");
    ExpectHmm *h = (ExpectHmm *)hmmExpectations;
    const bool hdp = sM->type == threeStateHdp;
    sa_read r;
    sa_params prm;
    int64_t *anc = NULL;
    make_read(where, sM, SsX, SsY, anchorPairs, p, alignmentHasRaggedLeftEnd, alignmentHasRaggedRightEnd, &r, &prm,
              &anc);
    pthread_mutex_lock(&g_mu);
    sa_context *ctx = context();
    sa_model *dm = device_model(ctx, (CompatMachine *)sM);
    sa_expect_result *res = NULL;
    /* the order-dependent accumulators replay every cell that adds to them; HDP
       assignments need the cells with a transition at or above the threshold */
    int rc = sa_expect(ctx, dm, &r, &prm, hdp ? h->threshold : 0.0, &res);
    if (rc != SA_OK) {
        fprintf(stderr, "signalalign_amd: %s: status %d: %s\n", where, rc, sa_last_error(ctx));
        abort();
    }
    pthread_mutex_unlock(&g_mu);
    free(anc);
    sa_expectations e;
    sa_expect_get(res, &e);
    h->base.likelihood += e.likelihood;
    for (int i = 0; i < 9; i++) h->base.transitions[i] += e.transitions[i];
    const int64_t k = sM->kmerLength, A = sM->alphabetSize;
    const char *ref = (const char *)SsX->elements;
    const double *ev = (const double *)SsY->elements;
    for (int64_t q = 0; q < e.n_records; q++) {
        const sa_match_record *rec = &e.records[q];
        const char *kmer = ref + rec->x;
        const double mean = ev[(int64_t)rec->y * NB_EVENT_PARAMS];
        if (hdp) {
            /* cell_signal_updateExpectationsAndAssignments (impl/pairwiseAligner.c:792-811) */
            for (int j = 0; j < 3; j++) {
                if (!(rec->p[j] >= h->threshold)) continue;
                if (h->nAssign == h->capAssign) {
                    h->capAssign = h->capAssign ? 2 * h->capAssign : 1024;
                    h->assignKmers = (char *)realloc(h->assignKmers, (size_t)(h->capAssign * k));
                    h->assignEvents = (double *)realloc(h->assignEvents, sizeof(double) * (size_t)h->capAssign);
                    if (!h->assignKmers || !h->assignEvents) die(where, "out of host memory");
                }
                memcpy(h->assignKmers + h->nAssign * k, kmer, (size_t)k);
                h->assignEvents[h->nAssign++] = mean;
            }
        } else {
            /* cell_signal_updateExpectations (:763-790) with
               continuousPairHmm_addToEmissionExpectation (impl/continuousHmm.c:160-169) */
            const int64_t kk = kmer_index(kmer, sM->alphabet, A, k);
            const double mu = h->eventModel[kk * MODEL_PARAMS];
            const double xd = (mean + h->var * mu - h->scale * mu - h->shift) / h->var;
            for (int j = 0; j < 3; j++) {
                const double pj = rec->p[j];
                h->eventExpectations[2 * kk] += pj * xd;
                h->posteriors[kk] += pj;
                const double u = h->eventExpectations[2 * kk] / h->posteriors[kk];
                h->eventExpectations[2 * kk + 1] += pj * (xd - u) * (xd - u);
                h->observed[kk] = true;
            }
        }
    }
    if (!hdp)
        for (int64_t kk = 0; kk < e.n_kmers; kk++)
            if (e.observed[kk]) h->observed[kk] = true;
    sa_expect_free(res);
}

int64_t hmmContinuous_howManyAssignments(Hmm *hmm) {
    if (!hmm || hmm->type != threeStateHdp) die("hmmContinuous_howManyAssignments", "not an HDP HMM");
    return ((ExpectHmm *)hmm)->nAssign;
}

/* continuousPairHmm_writeToFile / hdpHmm_writeToFile (impl/continuousHmm.c:354-408,
 * :572-629): header, transitions and likelihood, event model, then the
 * expectations (threeState) or the assignments (threeStateHdp). */
void hmmContinuous_writeToFile(const char *outFile, Hmm *hmm, StateMachineType type) {
    static const char *where = "hmmContinuous_writeToFile";
    if (!hmm || (type != threeState && type != threeStateHdp) || hmm->type != type) die(where, "unsupported HMM type");
    ExpectHmm *h = (ExpectHmm *)hmm;
    const int64_t nt = hmm->stateNumber * hmm->stateNumber, nk = hmm->parameterSetSize;
    for (int64_t i = 0; i < nt; i++)
        if (isnan(hmm->transitions[i])) {
            fprintf(stderr, "GOT NaN TRANS\n");
            return;
        }
    FILE *f = fopen(outFile, "w");
    if (!f) die(where, "cannot open the output file");
    fprintf(f, "%" PRId64 "\t%" PRId64 "\t%s\t%" PRId64 "\t\n", hmm->stateNumber, hmm->alphabetSize, hmm->alphabet,
            hmm->kmerLength);
    for (int64_t i = 0; i < nt; i++) fprintf(f, "%f\t", hmm->transitions[i]);
    fprintf(f, "%f\n", hmm->likelihood);
    for (int64_t i = 0; i < nk * MODEL_PARAMS; i++) fprintf(f, "%lf\t", h->eventModel[i]);
    fprintf(f, "\n");
    if (type == threeState) {
        for (int64_t i = 0; i < 2 * nk; i++) fprintf(f, "%lf\t", h->eventExpectations[i]);
        fprintf(f, "\n");
        for (int64_t i = 0; i < nk; i++) fprintf(f, "%lf\t", h->posteriors[i]);
        fprintf(f, "\n");
        for (int64_t i = 0; i < nk; i++) fprintf(f, "%d\t", h->observed[i] ? 1 : 0);
        fprintf(f, "\n");
    } else {
        for (int64_t i = 0; i < h->nAssign; i++) fprintf(f, "%lf\t", h->assignEvents[i]);
        fprintf(f, "\n");
        for (int64_t i = 0; i < h->nAssign; i++) {
            fwrite(h->assignKmers + i * hmm->kmerLength, 1, (size_t)hmm->kmerLength, f);
            fprintf(f, "\t");
        }
        fprintf(f, "\n");
    }
    fclose(f);
}

void hmmContinuous_destruct(Hmm *hmm, StateMachineType type) {
    (void)type;
    if (!hmm) return;
    ExpectHmm *h = (ExpectHmm *)hmm;
    free(h->base.alphabet);
    free(h->base.transitions);
    free(h->eventModel);
    free(h->eventExpectations);
    free(h->posteriors);
    free(h->observed);
    free(h->assignKmers);
    free(h->assignEvents);
    free(h);
}
