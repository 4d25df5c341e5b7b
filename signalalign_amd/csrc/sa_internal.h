/*
 * sa_internal.h -- host-side state shared by the library's translation units:
 * device and pinned buffers, contexts, resident model tables, error reporting
 * and split points (internal to libsignalalign_amd).
 */
#ifndef SA_INTERNAL_H
#define SA_INTERNAL_H

#include "signalalign_amd.h"

#include "sa_kernels.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

namespace sa_rt {


/* The stream a thread's batch buffers are ordered on while it holds a context's lock:
   growing a buffer then frees and allocates on that stream (stream-ordered pool)
   instead of hipFree, which waits for every stream on the device -- other
   contexts' work included. */
inline thread_local hipStream_t t_alloc_stream = nullptr;

struct AllocStream {
    explicit AllocStream(hipStream_t s) { t_alloc_stream = s; }
    ~AllocStream() { t_alloc_stream = nullptr; }
};

struct DBuf {
    void *p = nullptr;
    size_t cap = 0;
    hipStream_t owner = nullptr; /* stream of a stream-ordered allocation */
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap && p) return hipSuccess;
        release();
        size_t want = std::max<size_t>(bytes + bytes / 4, 256); /* headroom: refilled batches grow rarely */
        hipStream_t s = t_alloc_stream;
        hipError_t e = s ? hipMallocAsync(&p, want, s) : hipMalloc(&p, want);
        if (e != hipSuccess && s) { /* no stream-ordered pool: plain allocation */
            (void)hipGetLastError();
            p = nullptr;
            s = nullptr;
            e = hipMalloc(&p, want);
        }
        if (e == hipSuccess) { cap = want; owner = s; }
        else p = nullptr;
        return e;
    }
    void release() {
        if (p) {
            if (owner) (void)hipFreeAsync(p, t_alloc_stream ? t_alloc_stream : owner);
            else (void)hipFree(p);
        }
        p = nullptr;
        cap = 0;
        owner = nullptr;
    }
    template <typename U> U *as() const { return reinterpret_cast<U *>(p); }
};

/* pinned host staging (DMA at full PCIe rate; grows, never shrinks) */
struct HBuf {
    void *p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap && p) return hipSuccess;
        if (p) { (void)hipHostFree(p); p = nullptr; cap = 0; }
        size_t want = std::max<size_t>(bytes + bytes / 4, 1 << 20);
        hipError_t e = hipHostMalloc(&p, want, hipHostMallocNonCoherent); /* host-cached */
        if (e == hipSuccess) cap = want;
        return e;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
    template <typename U> U *as() const { return reinterpret_cast<U *>(p); }
};

/* Pinned buffers that large pair lists are copied into.  A result keeps the buffer
   its pairs arrived in and hands it back when it is freed, so a batch run again and
   again neither copies its pairs a second time on the host nor pins new memory.
   Results hold the pool by shared_ptr: one may outlive its context. */
struct PinnedPool {
    /* free buffers kept: at least kKeep, up to kKeepMax while they total at most
       kKeepBytes (a serving lane's waves finish while earlier waves' results are still
       held, and pinning or unpinning host memory waits for the whole device) */
    static constexpr size_t kKeep = 2, kKeepMax = 8;
    static constexpr size_t kKeepBytes = (size_t)2 << 30;
    std::mutex mu;
    std::vector<HBuf> idle;
    HBuf take() {
        std::lock_guard<std::mutex> lk(mu);
        if (idle.empty()) return HBuf();
        /* the largest: a repeated run needs the size it needed before */
        auto it = std::max_element(idle.begin(), idle.end(),
                                   [](const HBuf &a, const HBuf &b) { return a.cap < b.cap; });
        HBuf h = *it;
        idle.erase(it);
        return h;
    }
    void give(HBuf h) {
        if (!h.p) return;
        std::lock_guard<std::mutex> lk(mu);
        idle.push_back(h);
        for (;;) {
            size_t total = 0;
            for (const HBuf &b : idle) total += b.cap;
            if (idle.size() <= kKeep || (idle.size() <= kKeepMax && total <= kKeepBytes)) break;
            auto it = std::min_element(idle.begin(), idle.end(),
                                       [](const HBuf &a, const HBuf &b) { return a.cap < b.cap; });
            it->release();
            idle.erase(it);
        }
    }
    ~PinnedPool() {
        for (HBuf &h : idle) h.release();
    }
};


template <typename T>
struct DevTables {
    DBuf mmu, misd, mca, mnmu, mnb, ymu, yisd, yca, ynmu, ynb, gx, hmu, hrow, gridx, ys;
    void release() {
        DBuf *all[] = {&mmu, &misd, &mca, &mnmu, &mnb, &ymu, &yisd, &yca, &ynmu, &ynb, &gx, &hmu, &hrow, &gridx, &ys};
        for (DBuf *b : all) b->release();
    }
};

} // namespace sa_rt

using sa_rt::DBuf;
using sa_rt::DevTables;
using sa_rt::HBuf;

struct sa_context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t aux[3] = {};   /* sub-alignment groups of different window sizes run side by side */
    int n_streams = 4;         /* 1: everything on `stream` (aux[] alias it) */
    HBuf stage;                /* pinned staging for inputs and pairs (calls are serialised by mu) */
    sa_batch *batch = nullptr; /* sa_align_batch's batch, refilled call after call */
    std::shared_ptr<sa_rt::PinnedPool> pair_pool = std::make_shared<sa_rt::PinnedPool>();
    std::mutex mu;
    char err[512] = {};        /* text of the last error (fixed storage: setting it allocates nothing) */
    hipEvent_t ev[24] = {};    /* run steps; 16..23: each stream lane's emission plane */
    hipEvent_t gev[160] = {};  /* start/end of each forward and backward launch */
    int n_cu = 0;
    /* device address of the no-allocation entry point's scratch array on this
       context's device (resolved once per device at context creation) */
    unsigned char *scratch = nullptr;
};

struct sa_model {
    sa_context *ctx = nullptr;
    int device = 0;
    int32_t kind = 0;
    int32_t k = 0;
    int32_t A = 0;
    int32_t n_kmers = 0;
    std::string alphabet;
    int8_t lut[256];
    double trans[9];
    double endv[3], rendv[3];
    int32_t grid_length = 0;
    double inv_dx = 0.0;
    double grid0 = 0.0, grid1 = 0.0, gridn = 0.0; /* first, second and last grid knots */
    int32_t ek = 1;          /* register-kernel emission variant: 0 standard Gaussian, 1 any Gaussian, 2 HDP */
    int32_t n_rows = 0;
    double gx_const = -2.3025850929940455;
    DBuf d_lut;
    DevTables<double> t64;
    DevTables<float> t32;
};

namespace sa_rt {

/* An entry point's first step: the context's device current, and no error left over
   from an earlier call that a launch check of this one would pick up */
inline void enter(const sa_context *ctx) {
    (void)hipSetDevice(ctx->device);
    (void)hipGetLastError();
}

inline int fail(sa_context *ctx, int code, const char *msg) {
    if (ctx) std::snprintf(ctx->err, sizeof(ctx->err), "%s", msg);
    return code;
}
inline int fail(sa_context *ctx, int code, const std::string &msg) { return fail(ctx, code, msg.c_str()); }

inline int fail_hip(sa_context *ctx, hipError_t e, const char *what) {
    if (ctx) std::snprintf(ctx->err, sizeof(ctx->err), "%s: %s", what, hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? SA_ERR_OUT_OF_MEMORY : SA_ERR_HIP;
}

#define SA_HIP(ctx, call)                                                                                     \
    do {                                                                                                      \
        hipError_t _e = (call);                                                                               \
        if (_e != hipSuccess) return sa_rt::fail_hip((ctx), _e, #call);                                       \
    } while (0)

template <typename T>
sa::Model<T> device_model(const sa_model *m, const DevTables<T> &t) {
    sa::Model<T> r;
    std::memset(&r, 0, sizeof(r));
    r.kind = m->kind;
    r.kmer_length = m->k;
    r.alphabet_size = m->A;
    r.n_kmers = m->n_kmers;
    r.mmu = (const T *)t.mmu.p; r.misd = (const T *)t.misd.p; r.mca = (const T *)t.mca.p; r.mnmu = (const T *)t.mnmu.p;
    r.mnb = (const T *)t.mnb.p;
    r.ymu = (const T *)t.ymu.p; r.yisd = (const T *)t.yisd.p; r.yca = (const T *)t.yca.p; r.ynmu = (const T *)t.ynmu.p;
    r.ynb = (const T *)t.ynb.p;
    r.gx = (const T *)t.gx.p;
    r.hmu = (const T *)t.hmu.p;
    r.hrow = (const int32_t *)t.hrow.p;
    r.gridx = (const T *)t.gridx.p;
    r.ys = (const T *)t.ys.p;
    r.grid_length = m->grid_length;
    r.use_swx = m->trans[7] > -HUGE_VAL ? 1 : 0;
    r.std_y = m->kind != SA_EMIT_HDP && m->ek == 0 ? 1 : 0;
    r.inv_dx = (T)m->inv_dx;
    r.grid_x0 = (T)m->grid0;
    r.grid_xn = (T)m->gridn;
    r.grid_dx = (T)m->grid1 - (T)m->grid0; /* as the kernels computed it from the knots */
    r.gx_const = (T)m->gx_const;
    for (int i = 0; i < 9; i++) r.trans[i] = (T)m->trans[i];
    for (int i = 0; i < 3; i++) { r.endv[i] = (T)m->endv[i]; r.rendv[i] = (T)m->rendv[i]; }
    return r;
}

template <typename T> const DevTables<T> &tables(const sa_model *m);
template <> inline const DevTables<double> &tables<double>(const sa_model *m) { return m->t64; }
template <> inline const DevTables<float> &tables<float>(const sa_model *m) { return m->t32; }

inline int64_t correct_length(int64_t ref_length, int k) {
    /* sequence_correctSeqLength (impl/pairwiseAligner.c:456-471) for a k-mer sequence */
    if (ref_length <= 0) return 0;
    int64_t l = ref_length - (k - 1);
    return l > 0 ? l : 0;
}

/* getSplitPoints (impl/pairwiseAligner.c:1706-1757): calls emit(x1, y1, x2, y2) for
 * each block in order.  Returns false on anchors the reference asserts against. */
template <typename Emit>
inline bool split_points_each(const int64_t *anc, int64_t n, int64_t lX, int64_t lY, int64_t split, bool rl,
                              bool rr, Emit emit) {
    int64_t x1 = 0, y1 = 0, x2 = 0, y2 = 0;
    auto P = [&](int64_t x2_, int64_t y2_, int64_t x3, int64_t y3, bool skip) {
        int64_t lX2 = x3 - x2_, lY2 = y3 - y2_;
        if (lX2 * lY2 > split) {
            int64_t ml = (int64_t)std::sqrt((double)split);
            int64_t hX = lX2 / 2 > ml ? ml : lX2 / 2;
            int64_t hY = lY2 / 2 > ml ? ml : lY2 / 2;
            if (!skip) emit(x1, y1, x2_ + hX, y2_ + hY);
            x1 = x3 - hX;
            y1 = y3 - hY;
            return true;
        }
        return false;
    };
    for (int64_t i = 0; i < n; i++) {
        int64_t x3 = anc[2 * i], y3 = anc[2 * i + 1];
        P(x2, y2, x3, y3, rl && i == 0);
        if (x3 < x2 || y3 < y2 || x3 >= lX || y3 >= lY) return false;
        x2 = x3 + 1;
        y2 = y3 + 1;
    }
    if (!P(x2, y2, lX, lY, rl && n == 0) || !rr) emit(x1, y1, lX, lY);
    return true;
}

inline bool split_points(const int64_t *anc, int64_t n, int64_t lX, int64_t lY, int64_t split, bool rl, bool rr,
                         std::vector<int64_t> &out) {
    out.clear();
    return split_points_each(anc, n, lX, lY, split, rl, rr, [&](int64_t a, int64_t b, int64_t c, int64_t d) {
        out.push_back(a); out.push_back(b); out.push_back(c); out.push_back(d);
    });
}

/* The device's timeline anchor: an event recorded once per device, at its first
   context (on that context's stream); nullptr before. */
hipEvent_t timeline_anchor(int device);
void timeline_init(int device, hipStream_t st);
/* ms from the device's anchor to `ev` (a completed event on that device), 0 if none */
double timeline_ms(int device, hipEvent_t ev);

/* Resolves the no-allocation entry point's scratch array on the current device
   (sa_noalloc.cpp): its device address, or nullptr with the HIP error in *err. */
unsigned char *scratch_setup(int device, hipError_t *err);

} // namespace sa_rt

#endif
