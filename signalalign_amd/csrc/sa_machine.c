/*
 * sa_machine.c -- signalMachine's per-read driver pieces in the reference's C API:
 * .npRead reader, read scaling, guide alignment -> anchors, reference slices,
 * TSV writers.  Host-only C; the alignment itself is getAlignedPairsUsingAnchors
 * (sa_compat.c), which runs on the GPU.
 */
#define _GNU_SOURCE
#include "signalalign_machine.h"

#include <ctype.h>
#include <inttypes.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

static void fail(const char *where, const char *fmt, const char *arg, long long a, long long b) {
    fprintf(stderr, "signalalign_amd: %s: ", where);
    fprintf(stderr, fmt, arg ? arg : "", a, b);
    fprintf(stderr, "\n");
    abort();
}
#define FAIL(where, msg) fail(where, "%s" msg, NULL, 0, 0)

static void *xalloc(size_t n) {
    void *p = calloc(1, n ? n : 1);
    if (!p) FAIL("malloc", "out of host memory");
    return p;
}

static char *xstrdup(const char *s) {
    size_t n = strlen(s);
    char *d = (char *)xalloc(n + 1);
    memcpy(d, s, n + 1);
    return d;
}

/* ------------------------------------------------------------------------ */
/* text                                                                      */
/* ------------------------------------------------------------------------ */
/* one line without its newline, or NULL at end of file (stFile_getLineFromFile) */
static char *line_of(FILE *f) {
    char *buf = NULL;
    size_t cap = 0;
    ssize_t n = getline(&buf, &cap, f);
    if (n < 0) {
        free(buf);
        return NULL;
    }
    while (n > 0 && (buf[n - 1] == '\n' || buf[n - 1] == '\r')) buf[--n] = '\0';
    return buf;
}

/* whitespace-separated fields of a line, in place; returns the count */
typedef struct {
    char **tok;
    int64_t n, cap;
} Fields;

static void split_fields(char *s, Fields *fl) {
    fl->n = 0;
    char *p = s;
    while (p && *p) {
        while (*p && isspace((unsigned char)*p)) p++;
        if (!*p) break;
        if (fl->n == fl->cap) {
            fl->cap = fl->cap ? 2 * fl->cap : 64;
            fl->tok = (char **)realloc(fl->tok, sizeof(char *) * (size_t)fl->cap);
            if (!fl->tok) FAIL("split", "out of host memory");
        }
        fl->tok[fl->n++] = p;
        while (*p && !isspace((unsigned char)*p)) p++;
        if (*p) *p++ = '\0';
    }
}

static void fields_free(Fields *fl) {
    free(fl->tok);
    fl->tok = NULL;
    fl->n = fl->cap = 0;
}

static int parse_i64(const char *s, int64_t *out) {
    char *end;
    long long v = strtoll(s, &end, 10);
    if (end == s) return 0;
    *out = (int64_t)v;
    return 1;
}

static int parse_f64(const char *s, double *out) {
    char *end;
    double v = strtod(s, &end);
    if (end == s) return 0;
    *out = v;
    return 1;
}

/* ------------------------------------------------------------------------ */
/* k-mer indices                                                             */
/* ------------------------------------------------------------------------ */
int64_t kmer_id(const char *kmer, const char *alphabet, int64_t alphabetSize, int64_t kmerLength) {
    int64_t id = 0;
    for (int64_t i = 0; i < kmerLength; i++) {
        const char *hit = memchr(alphabet, kmer[i], (size_t)alphabetSize);
        if (!kmer[i] || !hit) fail("kmer_id", "k-mer character '%s' is not in the alphabet", (char[]){kmer[i] ? kmer[i] : '?', 0}, 0, 0);
        id = id * alphabetSize + (int64_t)(hit - alphabet);
    }
    return id;
}

/* The .npRead's model-state k-mers are indexed over the fixed six-letter
 * alphabet ACEGOT at k = 6 (emissions_discrete_getKmerIndexFromKmer,
 * impl/stateMachine.c:85-123).  The reference maps any other character to an
 * index past the table; this reader refuses it. */
static int64_t model_state_index(const char *s) {
    if ((int64_t)strlen(s) < 6) fail("nanopore_loadNanoporeReadFromFile", "model state '%s' is shorter than 6", s, 0, 0);
    return kmer_id(s, "ACEGOT", 6, 6);
}

/* ------------------------------------------------------------------------ */
/* .npRead (impl/nanopore.c:8-49, 98-424)                                    */
/* ------------------------------------------------------------------------ */
static void read_ints(char *line, int64_t *dst, int64_t n, const char *what) {
    Fields fl = {0};
    split_fields(line ? line : (char *)"", &fl);
    if (fl.n != n) fail("nanopore_loadNanoporeReadFromFile", "%s has %lld values, should have %lld", what, fl.n, n);
    for (int64_t i = 0; i < n; i++)
        if (!parse_i64(fl.tok[i], &dst[i])) fail("nanopore_loadNanoporeReadFromFile", "error loading in %s", what, 0, 0);
    free(fl.tok);
}

static void read_doubles(char *line, double *dst, int64_t n, const char *what) {
    Fields fl = {0};
    split_fields(line ? line : (char *)"", &fl);
    if (fl.n != n) fail("nanopore_loadNanoporeReadFromFile", "%s has %lld values, should have %lld", what, fl.n, n);
    for (int64_t i = 0; i < n; i++)
        if (!parse_f64(fl.tok[i], &dst[i])) fail("nanopore_loadNanoporeReadFromFile", "error loading in %s", what, 0, 0);
    free(fl.tok);
}

static void read_states(char *line, int64_t *dst, int64_t n, const char *what) {
    Fields fl = {0};
    split_fields(line ? line : (char *)"", &fl);
    if (fl.n != n) fail("nanopore_loadNanoporeReadFromFile", "%s has %lld k-mers, should have %lld", what, fl.n, n);
    for (int64_t i = 0; i < n; i++) dst[i] = model_state_index(fl.tok[i]);
    free(fl.tok);
}

/* first word of a line into a buffer of cap+1 bytes, truncated to cap */
static void read_word(char *line, char *dst, int64_t cap, const char *what) {
    Fields fl = {0};
    split_fields(line ? line : (char *)"", &fl);
    if (fl.n < 1) fail("nanopore_loadNanoporeReadFromFile", "error parsing %s", what, 0, 0);
    size_t n = strlen(fl.tok[0]);
    if ((int64_t)n > cap) n = (size_t)cap;
    memcpy(dst, fl.tok[0], n);
    dst[n] = '\0';
    free(fl.tok);
}

NanoporeRead *nanopore_loadNanoporeReadFromFile(const char *path) {
    static const char *W = "nanopore_loadNanoporeReadFromFile";
    FILE *f = fopen(path, "r");
    if (!f) fail(W, "cannot open %s", path, 0, 0);
    char *lines[14];
    for (int i = 0; i < 14; i++) lines[i] = line_of(f);
    fclose(f);

    Fields h = {0};
    split_fields(lines[0] ? lines[0] : (char *)"", &h);
    if (h.n != 18) fail(W, "incorrect line 1 for file %s got %lld tokens should get 18", path, h.n, 0);
    int64_t len[5];
    for (int i = 0; i < 5; i++)
        if (!parse_i64(h.tok[i], &len[i]) || len[i] < 0) fail(W, "error parsing line 1 field %s%lld", "", i + 1, 0);
    double pv[12];
    for (int i = 0; i < 12; i++)
        if (!parse_f64(h.tok[5 + i], &pv[i])) fail(W, "error parsing line 1 field %s%lld", "", i + 6, 0);
    int64_t two_d;
    if (!parse_i64(h.tok[17], &two_d)) FAIL(W, "error parsing 'has twoD'");
    free(h.tok);

    NanoporeRead *r = (NanoporeRead *)xalloc(sizeof(NanoporeRead));
    r->readLength = len[0];
    r->nbTemplateEvents = len[1];
    r->nbComplementEvents = len[2];
    r->templateReadLength = len[3];
    r->complementReadLength = len[4];
    NanoporeReadAdjustmentParameters *ps[2] = {&r->templateParams, &r->complementParams};
    for (int s = 0; s < 2; s++) {
        ps[s]->scale = pv[6 * s + 0];
        ps[s]->shift = pv[6 * s + 1];
        ps[s]->var = pv[6 * s + 2];
        ps[s]->scale_sd = pv[6 * s + 3];
        ps[s]->var_sd = pv[6 * s + 4];
        ps[s]->drift = pv[6 * s + 5];
        ps[s]->shift_sd = 0.0;
    }
    r->twoD = (int)two_d;
    r->scaled = true;
    r->twoDread = (char *)xalloc((size_t)r->readLength + 1);
    r->templateRead = (char *)xalloc((size_t)r->templateReadLength + 1);
    r->complementRead = (char *)xalloc((size_t)r->complementReadLength + 1);
    r->templateEventMap = (int64_t *)xalloc(sizeof(int64_t) * (size_t)r->readLength);
    r->complementEventMap = (int64_t *)xalloc(sizeof(int64_t) * (size_t)r->readLength);
    r->templateStrandEventMap = (int64_t *)xalloc(sizeof(int64_t) * (size_t)r->templateReadLength);
    r->complementStrandEventMap = (int64_t *)xalloc(sizeof(int64_t) * (size_t)r->complementReadLength);
    r->templateEvents = (double *)xalloc(sizeof(double) * (size_t)(r->nbTemplateEvents * NB_EVENT_PARAMS));
    r->complementEvents = (double *)xalloc(sizeof(double) * (size_t)(r->nbComplementEvents * NB_EVENT_PARAMS));
    r->templateModelState = (int64_t *)xalloc(sizeof(int64_t) * (size_t)r->nbTemplateEvents);
    r->complementModelState = (int64_t *)xalloc(sizeof(int64_t) * (size_t)r->nbComplementEvents);
    r->templatePModel = (double *)xalloc(sizeof(double) * (size_t)r->nbTemplateEvents);
    r->complementPModel = (double *)xalloc(sizeof(double) * (size_t)r->nbComplementEvents);

    if (r->twoD) read_word(lines[1], r->twoDread, r->readLength, "2D read");
    read_word(lines[2], r->templateRead, r->templateReadLength, "template read");
    read_ints(lines[3], r->templateStrandEventMap, r->templateReadLength, "template strand event map");
    if (r->twoD) read_word(lines[4], r->complementRead, r->complementReadLength, "complement read");
    read_ints(lines[5], r->complementStrandEventMap, r->complementReadLength, "complement strand event map");
    read_ints(lines[6], r->templateEventMap, r->readLength, "template 2D event map");
    read_doubles(lines[7], r->templateEvents, r->nbTemplateEvents * NB_EVENT_PARAMS, "template events");
    read_ints(lines[8], r->complementEventMap, r->readLength, "complement 2D event map");
    read_doubles(lines[9], r->complementEvents, r->nbComplementEvents * NB_EVENT_PARAMS, "complement events");
    read_states(lines[10], r->templateModelState, r->nbTemplateEvents, "template model states");
    read_doubles(lines[11], r->templatePModel, r->nbTemplateEvents, "template P(model state)");
    read_states(lines[12], r->complementModelState, r->nbComplementEvents, "complement model states");
    read_doubles(lines[13], r->complementPModel, r->nbComplementEvents, "complement P(model state)");
    for (int i = 0; i < 14; i++) free(lines[i]);
    return r;
}

void nanopore_nanoporeReadDestruct(NanoporeRead *r) {
    if (!r) return;
    free(r->twoDread);
    free(r->templateRead);
    free(r->complementRead);
    free(r->templateEventMap);
    free(r->templateStrandEventMap);
    free(r->templateEvents);
    free(r->complementEventMap);
    free(r->complementStrandEventMap);
    free(r->complementEvents);
    free(r->templateModelState);
    free(r->templatePModel);
    free(r->complementModelState);
    free(r->complementPModel);
    free(r);
}

NanoporeReadAdjustmentParameters *nanopore_readAdjustmentParametersConstruct(void) {
    /* impl/nanopore.c:86-96: all zero */
    return (NanoporeReadAdjustmentParameters *)xalloc(sizeof(NanoporeReadAdjustmentParameters));
}

EventKmerTuple *nanopore_eventKmerTupleConstruct(double mean, double sd, double deltaTime, int64_t kmerIndex) {
    EventKmerTuple *t = (EventKmerTuple *)xalloc(sizeof(EventKmerTuple));
    t->eventMean = mean;
    t->eventSd = sd;
    t->deltaTime = deltaTime;
    t->kmerIndex = kmerIndex;
    return t;
}

/* ------------------------------------------------------------------------ */
/* one-D assignments (impl/nanopore.c:474-533)                               */
/* ------------------------------------------------------------------------ */
/* Event fields: mean [0], sd [1], duration [2], start time [3]; the start time
 * is what the reference calls the event's delta time (impl/nanopore.c:51-69). */
static stList *tuples_from_model_states(int64_t *states, double *events, double *probs, int64_t n, double threshold) {
    stList *map = stList_construct3(0, free);
    for (int64_t i = 0; i < n; i++) {
        if (probs[i] < threshold) continue;
        const double *e = events + i * NB_EVENT_PARAMS;
        stList_append(map, nanopore_eventKmerTupleConstruct(e[0], e[1], e[2], states[i]));
    }
    return map;
}

stList *nanopore_getTemplateOneDAssignments(NanoporeRead *r, double threshold) {
    return tuples_from_model_states(r->templateModelState, r->templateEvents, r->templatePModel,
                                    r->nbTemplateEvents, threshold);
}

stList *nanopore_getComplementOneDAssignments(NanoporeRead *r, double threshold) {
    return tuples_from_model_states(r->complementModelState, r->complementEvents, r->complementPModel,
                                    r->nbComplementEvents, threshold);
}

stList *nanopore_getOneDAssignmentsFromRead(int64_t *strandEventMap, double *events, char *strandRead,
                                            int64_t readLength, char *alphabet, int64_t alphabetSize,
                                            int64_t kmerLength) {
    stList *map = stList_construct3(0, free);
    const int64_t rows = readLength - (kmerLength - 1);
    int64_t last = -1;
    for (int64_t i = 0; i < rows; i++) {
        const int64_t e = strandEventMap[i];
        const int64_t id = kmer_id(strandRead + i, alphabet, alphabetSize, kmerLength);
        /* an event mapped to several read positions counts once, at its first
           (highest-probability) position */
        if (e > last) {
            const double *ev = events + e * NB_EVENT_PARAMS;
            stList_append(map, nanopore_eventKmerTupleConstruct(ev[0], ev[1], ev[3], id));
            last = e;
        }
    }
    return map;
}

/* ------------------------------------------------------------------------ */
/* read scaling: weighted least squares (impl/nanopore.c:586-848)            */
/* ------------------------------------------------------------------------ */
void nanopore_lineq_solve(double *A, double *b, double *x_out, int64_t n) {
    double *M = (double *)xalloc(sizeof(double) * (size_t)(n * n));
    memcpy(M, A, sizeof(double) * (size_t)(n * n));
    memcpy(x_out, b, sizeof(double) * (size_t)n);
    for (int64_t c = 0; c < n; c++) {
        /* a row with a usable pivot moves up when the diagonal one is zero */
        if (fabs(M[c * n + c]) < MACHEP) {
            int64_t r = c + 1;
            while (r < n && fabs(M[r * n + c]) < MACHEP) r++;
            if (r >= n) FAIL("nanopore_lineq_solve", "matrix is not invertible");
            for (int64_t j = 0; j < n; j++) {
                double t = M[c * n + j];
                M[c * n + j] = M[r * n + j];
                M[r * n + j] = t;
            }
            double t = x_out[c];
            x_out[c] = x_out[r];
            x_out[r] = t;
        }
        const double inv = 1.0 / M[c * n + c];
        x_out[c] *= inv;
        for (int64_t j = 0; j < n; j++) M[c * n + j] *= inv;
        for (int64_t r = 0; r < n; r++) {
            if (r == c) continue;
            const double f = M[r * n + c];
            if (f == 0.0) continue;
            x_out[r] -= f * x_out[c];
            for (int64_t j = 0; j < n; j++) M[r * n + j] -= f * M[c * n + j];
        }
    }
    free(M);
}

/* Fits event = shift + scale * level_mean (+ drift * time) with weights
 * 1 / level_sd^2; var is the root mean weighted squared residual. */
void nanopore_compute_mean_scale_params(double *model, stList *map, NanoporeReadAdjustmentParameters *params,
                                        bool drift_out, bool var_out) {
    const int64_t n = stList_length(map);
    if (n == 0) FAIL("nanopore_compute_mean_scale_params", "cannot get scale params with no assignments");
    const int64_t p = drift_out ? 3 : 2;
    double XWX[9] = {0}, XWy[3] = {0}, beta[3] = {0};
    for (int64_t i = 0; i < n; i++) {
        const EventKmerTuple *t = (const EventKmerTuple *)stList_get(map, i);
        const double mu = model[t->kmerIndex * MODEL_PARAMS];
        const double sd = model[t->kmerIndex * MODEL_PARAMS + 1];
        const double w = 1.0 / (sd * sd);
        const double x[3] = {1.0, mu, t->deltaTime};
        for (int64_t a = 0; a < p; a++) {
            XWy[a] += w * x[a] * t->eventMean;
            for (int64_t c = a; c < p; c++) XWX[a * p + c] += w * x[a] * x[c];
        }
    }
    for (int64_t a = 0; a < p; a++)
        for (int64_t c = 0; c < a; c++) XWX[a * p + c] = XWX[c * p + a];
    nanopore_lineq_solve(XWX, XWy, beta, p);
    params->shift = beta[0];
    params->scale = beta[1];
    if (drift_out) params->drift = beta[2];
    if (var_out) {
        double disp = 0.0;
        for (int64_t i = 0; i < n; i++) {
            const EventKmerTuple *t = (const EventKmerTuple *)stList_get(map, i);
            const double mu = model[t->kmerIndex * MODEL_PARAMS];
            const double sd = model[t->kmerIndex * MODEL_PARAMS + 1];
            double pred = beta[0] + beta[1] * mu;
            if (drift_out) pred += beta[2] * t->deltaTime;
            const double res = t->eventMean - pred;
            disp += res * res / (sd * sd);
        }
        params->var = sqrt(disp / (double)n);
    }
}

/* Fits event_sd = shift_sd + scale_sd * noise_mean with weights 1 / noise_sd^2;
 * var_sd is the root mean weighted squared residual. */
void nanopore_compute_noise_scale_params(double *model, stList *map, NanoporeReadAdjustmentParameters *params) {
    const int64_t n = stList_length(map);
    if (n == 0) FAIL("nanopore_compute_noise_scale_params", "cannot get scale params with no assignments");
    double XWX[4] = {0}, XWy[2] = {0}, beta[2] = {0};
    for (int64_t i = 0; i < n; i++) {
        const EventKmerTuple *t = (const EventKmerTuple *)stList_get(map, i);
        const double nm = model[t->kmerIndex * MODEL_PARAMS + 2];
        const double ns = model[t->kmerIndex * MODEL_PARAMS + 3];
        const double w = 1.0 / (ns * ns);
        XWX[0] += w;
        XWX[1] += w * nm;
        XWX[3] += w * nm * nm;
        XWy[0] += w * t->eventSd;
        XWy[1] += w * nm * t->eventSd;
    }
    XWX[2] = XWX[1];
    nanopore_lineq_solve(XWX, XWy, beta, 2);
    params->shift_sd = beta[0];
    params->scale_sd = beta[1];
    double disp = 0.0;
    for (int64_t i = 0; i < n; i++) {
        const EventKmerTuple *t = (const EventKmerTuple *)stList_get(map, i);
        const double nm = model[t->kmerIndex * MODEL_PARAMS + 2];
        const double ns = model[t->kmerIndex * MODEL_PARAMS + 3];
        const double res = t->eventSd - (beta[0] + beta[1] * nm);
        disp += res * res / (ns * ns);
    }
    params->var_sd = sqrt(disp / (double)n);
}

static void drift_correct(double *events, double drift, int64_t n) {
    for (int64_t i = 0; i < n; i++) events[i * NB_EVENT_PARAMS] -= events[i * NB_EVENT_PARAMS + 3] * drift;
}

void nanopore_adjustEventsForDrift(NanoporeRead *r) {
    drift_correct(r->templateEvents, r->templateParams.drift, r->nbTemplateEvents);
    drift_correct(r->complementEvents, r->complementParams.drift, r->nbComplementEvents);
}

void nanopore_adjustTemplateEventsForDrift(NanoporeRead *r) {
    drift_correct(r->templateEvents, r->templateParams.drift, r->nbTemplateEvents);
}

void nanopore_adjustComplementEventsForDrift(NanoporeRead *r) {
    drift_correct(r->complementEvents, r->complementParams.drift, r->nbComplementEvents);
}

void nanopore_dontAdjustEvents(NanoporeRead *r) { (void)r; }

stList *signalUtils_templateOneDAssignmentsFromRead(NanoporeRead *r, StateMachine *sM, double ignore) {
    (void)ignore;
    return nanopore_getOneDAssignmentsFromRead(r->templateStrandEventMap, r->templateEvents, r->templateRead,
                                               r->templateReadLength, sM->alphabet, sM->alphabetSize,
                                               sM->kmerLength);
}

stList *signalUtils_complementOneDAssignmentsFromRead(NanoporeRead *r, StateMachine *sM, double ignore) {
    (void)ignore;
    return nanopore_getOneDAssignmentsFromRead(r->complementStrandEventMap, r->complementEvents,
                                               r->complementRead, r->complementReadLength, sM->alphabet,
                                               sM->alphabetSize, sM->kmerLength);
}

void signalUtils_estimateNanoporeParams(StateMachine *sM, NanoporeRead *r, NanoporeReadAdjustmentParameters *params,
                                        double assignmentThreshold,
                                        stList *(*assignmentFunction)(NanoporeRead *, StateMachine *, double),
                                        void (*driftAdjustmentFunction)(NanoporeRead *)) {
    stList *map = assignmentFunction(r, sM, assignmentThreshold);
    nanopore_compute_mean_scale_params(sM->EMISSION_MATCH_MATRIX, map, params, true, true);
    nanopore_compute_noise_scale_params(sM->EMISSION_MATCH_MATRIX, map, params);
    sM->scale = params->scale;
    sM->shift = params->shift;
    sM->var = params->var;
    driftAdjustmentFunction(r);
    emissions_signal_scaleNoise(sM, *params);
    stList_destruct(map);
}

/* ------------------------------------------------------------------------ */
/* guide alignment (sonLib's cigarRead / checkPairwiseAlignment)             */
/* ------------------------------------------------------------------------ */
static struct List *op_list(void) {
    struct List *l = (struct List *)xalloc(sizeof(struct List));
    l->maxLength = 16;
    l->list = (void **)xalloc(sizeof(void *) * (size_t)l->maxLength);
    l->destructElement = free;
    return l;
}

static void op_append(struct List *l, int64_t type, int64_t length) {
    if (l->length == l->maxLength) {
        l->maxLength *= 2;
        l->list = (void **)realloc(l->list, sizeof(void *) * (size_t)l->maxLength);
        if (!l->list) FAIL("cigarRead", "out of host memory");
    }
    struct AlignmentOperation *op = (struct AlignmentOperation *)xalloc(sizeof(*op));
    op->opType = type;
    op->length = length;
    op->score = 0.0;
    l->list[l->length++] = op;
}

struct PairwiseAlignment *cigarReadFromString(const char *line) {
    static const char *W = "cigarRead";
    char *s = xstrdup(line);
    Fields fl = {0};
    split_fields(s, &fl);
    struct PairwiseAlignment *pA = NULL;
    if (fl.n == 0) goto done;
    if (fl.n < 10 || strcmp(fl.tok[0], "cigar:") != 0) fail(W, "not an exonerate cigar line: %s", line, 0, 0);
    if ((fl.n - 10) % 2 != 0) FAIL(W, "operations must come in (type, length) pairs");
    pA = (struct PairwiseAlignment *)xalloc(sizeof(*pA));
    pA->contig2 = xstrdup(fl.tok[1]);
    pA->contig1 = xstrdup(fl.tok[5]);
    if (!parse_i64(fl.tok[2], &pA->start2) || !parse_i64(fl.tok[3], &pA->end2) ||
        !parse_i64(fl.tok[6], &pA->start1) || !parse_i64(fl.tok[7], &pA->end1) || !parse_f64(fl.tok[9], &pA->score))
        FAIL(W, "bad coordinates or score");
    const char *st[2] = {fl.tok[4], fl.tok[8]};
    for (int i = 0; i < 2; i++)
        if (strcmp(st[i], "+") != 0 && strcmp(st[i], "-") != 0) fail(W, "strand must be + or -, got %s", st[i], 0, 0);
    pA->strand2 = strcmp(fl.tok[4], "+") == 0;
    pA->strand1 = strcmp(fl.tok[8], "+") == 0;
    pA->operationList = op_list();
    for (int64_t i = 10; i < fl.n; i += 2) {
        int64_t len;
        if (strlen(fl.tok[i]) != 1 || !parse_i64(fl.tok[i + 1], &len) || len < 0)
            fail(W, "bad operation %s", fl.tok[i], 0, 0);
        int64_t type;
        switch (fl.tok[i][0]) {
            case 'M': type = PAIRWISE_MATCH; break;
            case 'D': type = PAIRWISE_INDEL_X; break; /* reference only */
            case 'I': type = PAIRWISE_INDEL_Y; break; /* read only */
            default: fail(W, "unknown operation %s", fl.tok[i], 0, 0); return NULL;
        }
        op_append(pA->operationList, type, len);
    }
done:
    free(fl.tok);
    free(s);
    return pA;
}

struct PairwiseAlignment *cigarRead(FILE *f) {
    for (;;) {
        char *line = line_of(f);
        if (!line) return NULL;
        struct PairwiseAlignment *pA = cigarReadFromString(line);
        free(line);
        if (pA) return pA;
    }
}

void checkPairwiseAlignment(struct PairwiseAlignment *pA) {
    int64_t span1 = 0, span2 = 0;
    for (int64_t i = 0; i < pA->operationList->length; i++) {
        const struct AlignmentOperation *op = (const struct AlignmentOperation *)pA->operationList->list[i];
        if (op->opType != PAIRWISE_INDEL_Y) span1 += op->length;
        if (op->opType != PAIRWISE_INDEL_X) span2 += op->length;
    }
    const int64_t d1 = pA->strand1 ? pA->end1 - pA->start1 : pA->start1 - pA->end1;
    const int64_t d2 = pA->strand2 ? pA->end2 - pA->start2 : pA->start2 - pA->end2;
    if (d1 < 0 || d2 < 0) FAIL("checkPairwiseAlignment", "coordinates run against their strand");
    if (d1 != span1) fail("checkPairwiseAlignment", "%sreference span %lld but operations cover %lld", "", d1, span1);
    if (d2 != span2) fail("checkPairwiseAlignment", "%sread span %lld but operations cover %lld", "", d2, span2);
}

void destructPairwiseAlignment(struct PairwiseAlignment *pA) {
    if (!pA) return;
    if (pA->operationList) {
        for (int64_t i = 0; i < pA->operationList->length; i++) free(pA->operationList->list[i]);
        free(pA->operationList->list);
        free(pA->operationList);
    }
    free(pA->contig1);
    free(pA->contig2);
    free(pA);
}

/* ------------------------------------------------------------------------ */
/* anchors (impl/pairwiseAligner.c:1454-1478, 1575-1616; impl/nanopore.c:438-450) */
/* ------------------------------------------------------------------------ */
stList *convertPairwiseForwardStrandAlignmentToAnchorPairs(struct PairwiseAlignment *pA, int64_t trim) {
    if (!pA->strand1 || !pA->strand2) FAIL("convertPairwiseForwardStrandAlignmentToAnchorPairs", "needs a forward-strand alignment");
    stList *out = stList_construct3(0, (void (*)(void *))stIntTuple_destruct);
    int64_t x = pA->start1, y = pA->start2;
    for (int64_t i = 0; i < pA->operationList->length; i++) {
        const struct AlignmentOperation *op = (const struct AlignmentOperation *)pA->operationList->list[i];
        /* a match block gives its diagonal minus `trim` positions at each end */
        if (op->opType == PAIRWISE_MATCH)
            for (int64_t l = trim; l < op->length - trim; l++) stList_append(out, stIntTuple_construct2(x + l, y + l));
        if (op->opType != PAIRWISE_INDEL_Y) x += op->length;
        if (op->opType != PAIRWISE_INDEL_X) y += op->length;
    }
    if (x != pA->end1 || y != pA->end2)
        FAIL("convertPairwiseForwardStrandAlignmentToAnchorPairs", "operations do not span the alignment");
    return out;
}

/* Keeps the pairs that are strictly increasing in both coordinates against
 * every other kept pair: a pair survives when it is below-left of all later
 * pairs (backward pass) and above-right of all earlier kept pairs (forward
 * pass).  Input sorted by (x, y). */
stList *filterToRemoveOverlap(stList *sorted) {
    const int64_t n = stList_length(sorted);
    unsigned char *below = (unsigned char *)xalloc((size_t)n + 1);
    int64_t mx = INT64_MAX, my = INT64_MAX;
    for (int64_t i = n - 1; i >= 0; i--) {
        stIntTuple *t = (stIntTuple *)stList_get(sorted, i);
        const int64_t x = stIntTuple_get(t, 0), y = stIntTuple_get(t, 1);
        below[i] = x < mx && y < my;
        if (x < mx) mx = x;
        if (y < my) my = y;
    }
    stList *out = stList_construct3(0, (void (*)(void *))stIntTuple_destruct);
    int64_t px = INT64_MIN, py = INT64_MIN, last_x = INT64_MIN, last_y = INT64_MIN;
    for (int64_t i = 0; i < n; i++) {
        stIntTuple *t = (stIntTuple *)stList_get(sorted, i);
        const int64_t x = stIntTuple_get(t, 0), y = stIntTuple_get(t, 1);
        if (x < last_x || (x == last_x && y < last_y)) FAIL("filterToRemoveOverlap", "input pairs are not sorted");
        last_x = x;
        last_y = y;
        /* the reference tests membership of an equal tuple in the backward set:
           an exact duplicate of a kept pair counts as kept */
        int keep_b = below[i];
        for (int64_t j = i + 1; !keep_b && j < n; j++) {
            stIntTuple *u = (stIntTuple *)stList_get(sorted, j);
            if (stIntTuple_get(u, 0) != x || stIntTuple_get(u, 1) != y) break;
            keep_b = below[j];
        }
        for (int64_t j = i - 1; !keep_b && j >= 0; j--) {
            stIntTuple *u = (stIntTuple *)stList_get(sorted, j);
            if (stIntTuple_get(u, 0) != x || stIntTuple_get(u, 1) != y) break;
            keep_b = below[j];
        }
        if (x > px && y > py && keep_b) stList_append(out, stIntTuple_construct2(x, y));
        if (x > px) px = x;
        if (y > py) py = y;
    }
    free(below);
    return out;
}

static int tuple_cmp_items(const void *a, const void *b) {
    return stIntTuple_cmpFn((stIntTuple *)a, (stIntTuple *)b);
}

stList *signalUtils_guideAlignmentToRebasedAnchorPairs(struct PairwiseAlignment *pA, PairwiseAlignmentParameters *p) {
    /* reference coordinates rebased to 0 and turned onto the forward strand;
       read coordinates stay (the event map translates them later) */
    const bool flip = !pA->strand1;
    const int64_t shift = pA->strand1 ? pA->start1 : pA->end1;
    pA->start1 -= shift;
    pA->end1 -= shift;
    if (flip) {
        pA->strand1 = 1;
        const int64_t t = pA->end1;
        pA->end1 = pA->start1;
        pA->start1 = t;
    }
    checkPairwiseAlignment(pA);
    stList *raw = convertPairwiseForwardStrandAlignmentToAnchorPairs(pA, p->constraintDiagonalTrim);
    stList_sort(raw, tuple_cmp_items);
    stList *out = filterToRemoveOverlap(raw);
    stList_destruct(raw);
    return out;
}

stList *nanopore_remapAnchorPairsWithOffset(stList *pairs, int64_t *eventMap, int64_t mapOffset) {
    stList *out = stList_construct3(0, (void (*)(void *))stIntTuple_destruct);
    const int64_t base = eventMap[mapOffset];
    for (int64_t i = 0; i < stList_length(pairs); i++) {
        stIntTuple *t = (stIntTuple *)stList_get(pairs, i);
        stList_append(out, stIntTuple_construct2(stIntTuple_get(t, 0), eventMap[stIntTuple_get(t, 1)] - base));
    }
    return out;
}

stList *signalUtils_getRemappedAnchorPairs(stList *unmapped, int64_t *eventMap, int64_t mapOffset) {
    stList *mapped = nanopore_remapAnchorPairsWithOffset(unmapped, eventMap, mapOffset);
    stList *out = filterToRemoveOverlap(mapped);
    stList_destruct(mapped);
    return out;
}

/* ------------------------------------------------------------------------ */
/* reference slices (impl/signalMachineUtils.c:7-86)                         */
/* ------------------------------------------------------------------------ */
char *signalUtils_stringReverse(char *str) {
    if (!str) return str;
    size_t n = strlen(str);
    for (size_t i = 0; i + 1 < n - i; i++) {
        char t = str[i];
        str[i] = str[n - 1 - i];
        str[n - 1 - i] = t;
    }
    return str;
}

char *signalUtils_getSubSequence(char *seq, int64_t start, int64_t end, bool strand) {
    const int64_t from = strand ? start : end, len = strand ? end - start : start - end;
    if (from < 0 || len < 0 || from + len > (int64_t)strlen(seq))
        fail("signalUtils_getSubSequence", "%sslice [%lld, +%lld) is outside the reference", "", from, len);
    char *s = (char *)xalloc((size_t)len + 1);
    memcpy(s, seq + from, (size_t)len);
    s[len] = '\0';
    return s;
}

static char *first_line(const char *path) {
    FILE *f = fopen(path, "r");
    if (!f) fail("signalUtils_ReferenceSequenceConstructFull", "couldn't find reference file %s", path, 0, 0);
    char *l = line_of(f);
    fclose(f);
    if (!l) fail("signalUtils_ReferenceSequenceConstructFull", "empty reference file %s", path, 0, 0);
    return l;
}

static char *template_target(ReferenceSequence *R) {
    return R->A->strand1 ? R->trimmedForwardSequence : R->trimmedBackwardSequence;
}

static char *complement_target(ReferenceSequence *R) {
    return R->A->strand1 ? R->trimmedBackwardSequence : R->trimmedForwardSequence;
}

ReferenceSequence *signalUtils_ReferenceSequenceConstructFull(char *forwardReferencePath, char *backwardReferencePath,
                                                              struct PairwiseAlignment *pA) {
    ReferenceSequence *R = (ReferenceSequence *)xalloc(sizeof(ReferenceSequence));
    R->reference = first_line(forwardReferencePath);
    R->complementOfReference = first_line(backwardReferencePath);
    R->A = (struct PairwiseAlignment *)xalloc(sizeof(struct PairwiseAlignment));
    *R->A = *pA;
    R->A->contig1 = xstrdup(pA->contig1);
    R->A->contig2 = xstrdup(pA->contig2);
    R->A->operationList = NULL;
    R->trimmedForwardSequence = signalUtils_getSubSequence(R->reference, pA->start1, pA->end1, pA->strand1);
    /* the backward file holds the complement, read here in reverse: the
       reverse complement of the aligned span */
    R->trimmedBackwardSequence =
        signalUtils_stringReverse(signalUtils_getSubSequence(R->complementOfReference, pA->start1, pA->end1, pA->strand1));
    R->getTemplateTargetSequence = template_target;
    R->getComplementTargetSequence = complement_target;
    R->initialized = true;
    return R;
}

void signalUtils_ReferenceSequenceDestruct(ReferenceSequence *R) {
    if (!R) return;
    destructPairwiseAlignment(R->A);
    free(R->reference);
    free(R->complementOfReference);
    free(R->trimmedForwardSequence);
    free(R->trimmedBackwardSequence);
    free(R);
}

/* ------------------------------------------------------------------------ */
/* HDP models (impl/nanopore_hdp.c:1071-1098, impl/hdp.c:3047-3317)           */
/* ------------------------------------------------------------------------ */
typedef struct {
    char **lines;
    int64_t n, pos;
    const char *path;
} LineCursor;

static char *next_line(LineCursor *c) {
    if (c->pos >= c->n) fail("deserialize_nhdp", "%s: unexpected end of file", c->path, 0, 0);
    return c->lines[c->pos++];
}

static int64_t first_i64(LineCursor *c, const char *what) {
    int64_t v;
    if (!parse_i64(next_line(c), &v)) fail("deserialize_nhdp", "bad %s", what, 0, 0);
    return v;
}

/* the values of a line into dst (n expected; an empty line leaves dst alone and
   returns 0) */
static int row_of_doubles(char *line, double *dst, int64_t n, const char *path) {
    Fields fl = {0};
    split_fields(line, &fl);
    if (fl.n == 0) {
        free(fl.tok);
        return 0;
    }
    if (fl.n != n) fail("deserialize_nhdp", "%s: a grid row has %lld values, should have %lld", path, fl.n, n);
    for (int64_t i = 0; i < n; i++)
        if (!parse_f64(fl.tok[i], &dst[i])) fail("deserialize_nhdp", "%s: bad grid value", path, 0, 0);
    free(fl.tok);
    return 1;
}

NanoporeHDP *deserialize_nhdp(const char *path) {
    static const char *W = "deserialize_nhdp";
    FILE *f = fopen(path, "r");
    if (!f) fail(W, "cannot open %s", path, 0, 0);
    LineCursor c = {NULL, 0, 0, path};
    int64_t cap = 0;
    for (char *l; (l = line_of(f));) {
        if (c.n == cap) {
            cap = cap ? 2 * cap : 1024;
            c.lines = (char **)realloc(c.lines, sizeof(char *) * (size_t)cap);
            if (!c.lines) FAIL(W, "out of host memory");
        }
        c.lines[c.n++] = l;
    }
    fclose(f);

    NanoporeHDP *h = (NanoporeHDP *)xalloc(sizeof(NanoporeHDP));
    h->alphabet_size = first_i64(&c, "alphabet size");
    Fields fl = {0};
    char *al = next_line(&c);
    split_fields(al, &fl);
    if (fl.n < 1 || (int64_t)strlen(fl.tok[0]) < h->alphabet_size || h->alphabet_size < 1)
        fail(W, "%s: bad alphabet", path, 0, 0);
    h->alphabet = (char *)xalloc((size_t)h->alphabet_size + 1);
    memcpy(h->alphabet, fl.tok[0], (size_t)h->alphabet_size);
    fields_free(&fl);
    for (int64_t i = 1; i < h->alphabet_size; i++) /* sorted, as the machine's */
        for (int64_t j = i; j > 0 && h->alphabet[j - 1] > h->alphabet[j]; j--) {
            char t = h->alphabet[j];
            h->alphabet[j] = h->alphabet[j - 1];
            h->alphabet[j - 1] = t;
        }
    h->kmer_length = first_i64(&c, "k-mer length");
    const int64_t finalized = first_i64(&c, "splines-finalized flag");
    const int64_t has_data = first_i64(&c, "has-data flag");
    const int64_t sample_gamma = first_i64(&c, "sample-gamma flag");
    h->num_dps = first_i64(&c, "number of DPs");
    if (h->num_dps < 1 || h->kmer_length < 1) fail(W, "%s: bad header", path, 0, 0);
    if (!finalized) fail(W, "%s: spline slopes are not finalised", path, 0, 0);
    int64_t *dp_ids = NULL, n_data = 0;
    if (has_data) {
        next_line(&c); /* the data points */
        split_fields(next_line(&c), &fl);
        n_data = fl.n;
        dp_ids = (int64_t *)xalloc(sizeof(int64_t) * (size_t)(n_data ? n_data : 1));
        for (int64_t i = 0; i < n_data; i++)
            if (!parse_i64(fl.tok[i], &dp_ids[i]) || dp_ids[i] < 0 || dp_ids[i] >= h->num_dps)
                fail(W, "%s: bad data-point DP id", path, 0, 0);
        fields_free(&fl);
    }
    next_line(&c); /* base distribution parameters */
    split_fields(next_line(&c), &fl);
    double g0, g1;
    int64_t gl;
    if (fl.n < 3 || !parse_f64(fl.tok[0], &g0) || !parse_f64(fl.tok[1], &g1) || !parse_i64(fl.tok[2], &gl) || gl < 2)
        fail(W, "%s: bad grid line", path, 0, 0);
    fields_free(&fl);
    h->grid_length = gl;
    h->grid = (double *)xalloc(sizeof(double) * (size_t)gl);
    const double dx = (g1 - g0) / (double)(gl - 1); /* linspace, impl/hdp_math_utils.c:498-510 */
    for (int64_t i = 0; i < gl; i++) h->grid[i] = g0 + (double)i * dx;
    h->grid[gl - 1] = g1;
    next_line(&c); /* gamma */
    if (sample_gamma)
        for (int i = 0; i < 4; i++) next_line(&c);
    h->parent = (int64_t *)xalloc(sizeof(int64_t) * (size_t)h->num_dps);
    for (int64_t i = 0; i < h->num_dps; i++) {
        split_fields(next_line(&c), &fl);
        if (fl.n < 1) fail(W, "%s: missing parent of a DP", path, 0, 0);
        if (strcmp(fl.tok[0], "-") == 0)
            h->parent[i] = -1;
        else if (!parse_i64(fl.tok[0], &h->parent[i]) || h->parent[i] < 0 || h->parent[i] >= h->num_dps)
            fail(W, "%s: bad parent id", path, 0, 0);
        fields_free(&fl);
    }
    /* observed: a data point sits in the DP or below it (mark_observed_dps,
       impl/hdp.c:1127-1155) */
    h->observed = (bool *)xalloc(sizeof(bool) * (size_t)h->num_dps);
    for (int64_t i = 0; i < n_data; i++)
        for (int64_t j = dp_ids[i]; j >= 0 && !h->observed[j]; j = h->parent[j]) h->observed[j] = true;
    free(dp_ids);
    h->row = (int64_t *)xalloc(sizeof(int64_t) * (size_t)h->num_dps);
    h->n_rows = 0;
    for (int64_t i = 0; i < h->num_dps; i++) h->row[i] = h->observed[i] ? h->n_rows++ : -1;
    if (h->n_rows == 0) fail(W, "%s: no observed DP", path, 0, 0);
    h->density = (double *)xalloc(sizeof(double) * (size_t)(h->n_rows * gl));
    h->slopes = (double *)xalloc(sizeof(double) * (size_t)(h->n_rows * gl));
    double *scratch = (double *)xalloc(sizeof(double) * (size_t)gl);
    if (has_data)
        for (int64_t i = 0; i < h->num_dps; i++) {
            double *dst = h->row[i] >= 0 ? h->density + h->row[i] * gl : scratch;
            if (!row_of_doubles(next_line(&c), dst, gl, path) && h->row[i] >= 0)
                fail(W, "%s: an observed DP has no posterior predictive", path, 0, 0);
        }
    for (int64_t i = 0; i < h->num_dps; i++) {
        double *dst = h->row[i] >= 0 ? h->slopes + h->row[i] * gl : scratch;
        row_of_doubles(next_line(&c), dst, gl, path);
    }
    free(scratch);
    for (int64_t i = 0; i < c.n; i++) free(c.lines[i]);
    free(c.lines);
    return h;
}

void destroy_nanopore_hdp(NanoporeHDP *h) {
    if (!h) return;
    free(h->alphabet);
    free(h->grid);
    free(h->parent);
    free(h->observed);
    free(h->row);
    free(h->density);
    free(h->slopes);
    free(h);
}

double nhdp_dp_expected_value(NanoporeHDP *h, int64_t dp) {
    if (dp < 0 || dp >= h->num_dps || h->row[dp] < 0) FAIL("nhdp_dp_expected_value", "DP is not observed");
    const double *g = h->grid, *p = h->density + h->row[dp] * h->grid_length;
    double e = 0.0;
    for (int64_t i = 1; i < h->grid_length; i++) e += g[i] * p[i] * (g[i] - g[i - 1]);
    return e;
}

double nhdp_dp_variance(NanoporeHDP *h, int64_t dp) {
    const double e = nhdp_dp_expected_value(h, dp);
    const double *g = h->grid, *p = h->density + h->row[dp] * h->grid_length;
    double v = 0.0;
    for (int64_t i = 1; i < h->grid_length; i++) {
        const double d = g[i] - e;
        v += d * d * p[i] * (g[i] - g[i - 1]);
    }
    return v;
}

StateMachine *getHdpStateMachine(NanoporeHDP *h, const char *modelFile, NanoporeReadAdjustmentParameters npp) {
    int64_t nk = 1;
    for (int64_t i = 0; i < h->kmer_length; i++) nk *= h->alphabet_size;
    if (h->num_dps < nk) FAIL("getHdpStateMachine", "the HDP has fewer DPs than k-mers");
    /* every k-mer reads its first observed ancestor (dir_proc_density,
       impl/hdp.c:2583-2607) */
    int64_t *kmer_row = (int64_t *)xalloc(sizeof(int64_t) * (size_t)nk);
    for (int64_t k = 0; k < nk; k++) {
        int64_t j = k;
        while (j >= 0 && !h->observed[j]) j = h->parent[j];
        if (j < 0) FAIL("getHdpStateMachine", "a k-mer has no observed ancestor DP");
        kmer_row[k] = h->row[j];
    }
    StateMachine *sM = getHdpStateMachineFromGrid(modelFile, h->grid, h->grid_length, h->density, h->slopes,
                                                  h->n_rows, kmer_row, NULL, npp);
    free(kmer_row);
    if (sM->kmerLength != h->kmer_length || sM->alphabetSize != h->alphabet_size ||
        strcmp(sM->alphabet, h->alphabet) != 0)
        FAIL("getHdpStateMachine", "the HDP and the model file have different alphabets or k-mer lengths");
    return sM;
}

void stateMachine3_setModelToHdpExpectedValues(StateMachine *sM, NanoporeHDP *h) {
    if (strcmp(sM->alphabet, h->alphabet) != 0 || sM->alphabetSize != h->alphabet_size ||
        sM->kmerLength != h->kmer_length)
        FAIL("stateMachine3_setModelToHdpExpectedValues", "the HDP and the machine have different alphabets");
    for (int64_t k = 0; k < sM->parameterSetSize; k++) {
        if (!h->observed[k]) continue;
        sM->EMISSION_MATCH_MATRIX[k * MODEL_PARAMS] = nhdp_dp_expected_value(h, k);
        sM->EMISSION_MATCH_MATRIX[k * MODEL_PARAMS + 1] = sqrt(nhdp_dp_variance(h, k));
    }
}

/* ------------------------------------------------------------------------ */
/* alignment (signalMachine.c:333-369)                                       */
/* ------------------------------------------------------------------------ */
Sequence *makeEventSequenceFromPairwiseAlignment(double *events, int64_t queryStart, int64_t queryEnd,
                                                 int64_t *eventMap) {
    const int64_t a = eventMap[queryStart], b = eventMap[queryEnd];
    if (b < a) FAIL("makeEventSequenceFromPairwiseAlignment", "the event map runs backwards over the alignment");
    return sequence_constructEventSequence(b - a, events + a * NB_EVENT_PARAMS);
}

stList *performSignalAlignment(StateMachine *sM, Sequence *eventSequence, int64_t *eventMap, int64_t mapOffset,
                               char *target, PairwiseAlignmentParameters *p, stList *unmappedAnchors,
                               DegenerateType degenerate) {
    if (sM->type != threeState && sM->type != threeStateHdp)
        FAIL("performSignalAlignment", "needs a three-state machine");
    const int64_t lX = sequence_correctSeqLength((int64_t)strlen(target), kmer, sM->kmerLength);
    stList *anchors = signalUtils_getRemappedAnchorPairs(unmappedAnchors, eventMap, mapOffset);
    Sequence *sX = sequence_constructReferenceKmerSequence(lX, target, sequence_getKmer,
                                                           sequence_sliceNucleotideSequence, degenerate, kmer);
    stList *pairs = getAlignedPairsUsingAnchors(sM, sX, eventSequence, anchors, p,
                                                diagonalCalculationPosteriorMatchProbs, 1, 1);
    sequence_destruct(sX);
    stList_destruct(anchors);
    return pairs;
}

void getSignalExpectations(StateMachine *sM, Hmm *hmmExpectations, Sequence *eventSequence, int64_t *eventMap,
                           int64_t mapOffset, char *trainingTarget, PairwiseAlignmentParameters *p,
                           stList *unmappedAnchors, DegenerateType degenerate) {
    const int64_t lX = sequence_correctSeqLength((int64_t)strlen(trainingTarget), event, sM->kmerLength);
    stList *anchors = signalUtils_getRemappedAnchorPairs(unmappedAnchors, eventMap, mapOffset);
    Sequence *target = sequence_constructReferenceKmerSequence(lX, trainingTarget, sequence_getKmer,
                                                               sequence_sliceNucleotideSequence, degenerate, kmer);
    getExpectationsUsingAnchors(sM, hmmExpectations, target, eventSequence, anchors, p,
                                diagonalCalculation_Expectations, 1, 1);
    sequence_destruct(target);
    stList_destruct(anchors);
}

double scoreByPosteriorProbabilityIgnoringGaps(stList *pairs) {
    double s = 0.0;
    for (int64_t i = 0; i < stList_length(pairs); i++) s += (double)stIntTuple_get((stIntTuple *)stList_get(pairs, i), 0);
    return 100.0 * s / ((double)stList_length(pairs) * PAIR_ALIGNMENT_PROB_1);
}

int sortByXPlusYCoordinate2(const void *i, const void *j) {
    const int64_t k = stIntTuple_get((stIntTuple *)i, 1) + stIntTuple_get((stIntTuple *)i, 2);
    const int64_t l = stIntTuple_get((stIntTuple *)j, 1) + stIntTuple_get((stIntTuple *)j, 2);
    return k > l ? 1 : (k < l ? -1 : 0);
}

stList *path_findDegeneratePositions(char *kmer, int64_t kmerLength) {
    stList *pos = stList_construct3(0, free);
    if (!kmer) return pos;
    for (int64_t i = 0; i < kmerLength; i++) {
        if (kmer[i] == 'X') {
            int64_t *v = (int64_t *)xalloc(sizeof(int64_t));
            *v = i;
            stList_append(pos, v);
        }
    }
    return pos;
}

/* ------------------------------------------------------------------------ */
/* TSV writers (signalMachine.c:49-260)                                      */
/* ------------------------------------------------------------------------ */
static bool same_orientation(Strand strand, bool forward) {
    return (strand == 0 && forward) || (strand == 1 && !forward);
}

static int64_t reference_coordinate(int64_t x, int64_t offset, int64_t lenKmers, int64_t len, Strand strand,
                                    bool forward) {
    return same_orientation(strand, forward) ? x + offset : lenKmers - (x + (len - offset));
}

static int64_t query_position(int64_t pos, int64_t k, Strand strand, bool forward) {
    return same_orientation(strand, forward) ? pos : (k - 1) - pos;
}

static char complement_base(char c) {
    switch (c) {
        case 'A': return 'T';
        case 'T': return 'A';
        case 'C': return 'G';
        case 'G': return 'C';
        case 'a': return 't';
        case 't': return 'a';
        case 'c': return 'g';
        case 'g': return 'c';
        default: return c;
    }
}

/* k-mer at x of the target, and the reference-strand k-mer it stands for */
static void kmers_at(const char *target, int64_t x, int64_t k, Strand strand, bool forward, char *k_i, char *ref) {
    memcpy(k_i, target + x, (size_t)k);
    k_i[k] = '\0';
    if (same_orientation(strand, forward)) {
        memcpy(ref, k_i, (size_t)k + 1);
    } else {
        for (int64_t i = 0; i < k; i++) ref[i] = complement_base(k_i[k - 1 - i]);
        ref[k] = '\0';
    }
}

static FILE *append_to(const char *path) {
    FILE *f = fopen(path, "a");
    if (!f) fail("outputAlignment", "cannot open %s", path, 0, 0);
    return f;
}

static void check_pair(stIntTuple *t) {
    if (stIntTuple_length(t) != 4)
        fail("outputAlignment", "%saligned pair tuples should have length 4, this one has length %lld", "",
             stIntTuple_length(t), 0);
}

static double descale_mean(double x, double mu, double scale, double shift, double var) {
    /* emissions_signal_descaleEventMean_JordanStyle (impl/stateMachine.c:285-289) */
    return (x + var * mu - scale * mu - shift) / var;
}

void writePosteriorProbsFull(char *path, char *readLabel, StateMachine *sM, NanoporeReadAdjustmentParameters npp,
                             double *events, char *target, bool forward, char *contig, int64_t eventSequenceOffset,
                             int64_t referenceSequenceOffset, stList *pairs, Strand strand) {
    const char *label = strand == 0 ? "t" : "c";
    const int64_t k = sM->kmerLength;
    const int64_t len = (int64_t)strlen(target), lenKmers = len - k;
    char *k_i = (char *)xalloc((size_t)k + 1), *ref = (char *)xalloc((size_t)k + 1);
    FILE *f = append_to(path);
    for (int64_t i = 0; i < stList_length(pairs); i++) {
        stIntTuple *t = (stIntTuple *)stList_get(pairs, i);
        check_pair(t);
        const int64_t x = stIntTuple_get(t, 1);
        const int64_t x_adj = reference_coordinate(x, referenceSequenceOffset, lenKmers, len, strand, forward);
        const int64_t y = stIntTuple_get(t, 2) + eventSequenceOffset;
        const double p = (double)stIntTuple_get(t, 0) / PAIR_ALIGNMENT_PROB_1;
        char *path_kmer = (char *)(intptr_t)stIntTuple_get(t, 3);
        const double *e = events + y * NB_EVENT_PARAMS;
        kmers_at(target, x, k, strand, forward, k_i, ref);
        const int64_t id = kmer_id(path_kmer, sM->alphabet, sM->alphabetSize, k);
        const double mu = sM->EMISSION_MATCH_MATRIX[id * MODEL_PARAMS];
        const double noise = sM->EMISSION_MATCH_MATRIX[id * MODEL_PARAMS + 2];
        fprintf(f, "%s\t%" PRId64 "\t%s\t%s\t%s\t%" PRId64 "\t%f\t%f\t%f\t%s\t%f\t%f\t%f\t%f\t%f\t%s\n", contig, x_adj,
                ref, readLabel, label, y, e[0], e[1], e[2], k_i, mu * npp.scale + npp.shift, noise * npp.scale_sd, p,
                descale_mean(e[0], mu, npp.scale, npp.shift, npp.var), mu, path_kmer);
    }
    fclose(f);
    free(k_i);
    free(ref);
}

void writePosteriorProbsVC(char *path, char *readLabel, StateMachine *sM, char *target, bool forward,
                           int64_t eventSequenceOffset, int64_t referenceSequenceOffset, stList *pairs, Strand strand) {
    const char *label = strand == 0 ? "t" : "c";
    const char *dir = forward ? "forward" : "backward";
    const int64_t k = sM->kmerLength;
    const int64_t len = (int64_t)strlen(target), lenKmers = len - k;
    char *k_i = (char *)xalloc((size_t)k + 1), *ref = (char *)xalloc((size_t)k + 1);
    FILE *f = append_to(path);
    for (int64_t i = 0; i < stList_length(pairs); i++) {
        stIntTuple *t = (stIntTuple *)stList_get(pairs, i);
        check_pair(t);
        const int64_t x = stIntTuple_get(t, 1);
        kmers_at(target, x, k, strand, forward, k_i, ref);
        stList *qs = path_findDegeneratePositions(ref, k);
        if (stList_length(qs) > 0) {
            const int64_t x_adj = reference_coordinate(x, referenceSequenceOffset, lenKmers, len, strand, forward);
            const int64_t y = stIntTuple_get(t, 2) + eventSequenceOffset;
            const double p = (double)stIntTuple_get(t, 0) / PAIR_ALIGNMENT_PROB_1;
            const char *path_kmer = (const char *)(intptr_t)stIntTuple_get(t, 3);
            for (int64_t q = 0; q < stList_length(qs); q++) {
                const int64_t pos = *(int64_t *)stList_get(qs, q);
                const char base = path_kmer[query_position(pos, k, strand, forward)];
                fprintf(f, "%" PRId64 "\t%" PRId64 "\t%c\t%f\t%s\t%s\t%s\n", y, x_adj + pos, base, p, label, dir,
                        readLabel);
            }
        }
        stList_destruct(qs);
    }
    fclose(f);
    free(k_i);
    free(ref);
}

void writeAssignments(char *path, StateMachine *sM, double *events, int64_t eventSequenceOffset,
                      NanoporeReadAdjustmentParameters npp, stList *pairs, Strand strand) {
    const char *label = strand == 0 ? "t" : "c";
    FILE *f = append_to(path);
    for (int64_t i = 0; i < stList_length(pairs); i++) {
        stIntTuple *t = (stIntTuple *)stList_get(pairs, i);
        check_pair(t);
        const int64_t y = stIntTuple_get(t, 2) + eventSequenceOffset;
        const double p = (double)stIntTuple_get(t, 0) / PAIR_ALIGNMENT_PROB_1;
        char *path_kmer = (char *)(intptr_t)stIntTuple_get(t, 3);
        const double x = events[y * NB_EVENT_PARAMS];
        const int64_t id = kmer_id(path_kmer, sM->alphabet, sM->alphabetSize, sM->kmerLength);
        const double mu = sM->EMISSION_MATCH_MATRIX[id * MODEL_PARAMS];
        fprintf(f, "%s\t%s\t%lf\t%lf\n", path_kmer, label, descale_mean(x, mu, npp.scale, npp.shift, npp.var), p);
    }
    fclose(f);
}

void outputAlignment(OutputFormat fmt, char *path, char *readLabel, StateMachine *sM,
                     NanoporeReadAdjustmentParameters npp, double *events, char *target, bool forward, char *contig,
                     int64_t eventSequenceOffset, int64_t referenceSequenceOffset, stList *pairs, Strand strand) {
    switch (fmt) {
        case full:
            writePosteriorProbsFull(path, readLabel, sM, npp, events, target, forward, contig, eventSequenceOffset,
                                    referenceSequenceOffset, pairs, strand);
            break;
        case variantCaller:
            writePosteriorProbsVC(path, readLabel, sM, target, forward, eventSequenceOffset, referenceSequenceOffset,
                                  pairs, strand);
            break;
        case assignments:
            writeAssignments(path, sM, events, eventSequenceOffset, npp, pairs, strand);
            break;
        default:
            fprintf(stderr, "signalAlign - No valid output format provided\n");
    }
}
