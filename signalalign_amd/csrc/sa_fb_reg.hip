/*
 * sa_fb_reg.hip -- register-resident forward and backward sweeps (one wavefront
 * per sub-alignment / per traceback window).
 *
 * Layout: lane i owns the CPT consecutive reference positions
 * x = X0 + CPT*i + c (c < CPT) of a window [X0, X0 + 64*CPT) that slides along the
 * band.  On anti-diagonal d a lane's cells are (x, d - x); the three states of the
 * two previous anti-diagonals stay in registers.  The neighbours of
 * stateMachine3_cellCalculate (impl/stateMachine.c:1194-1257) are then:
 *   forward  lower  (x-1, d-1): own cell c-1, or lane i-1's last cell (DPP wave_shr:1)
 *            upper  (x,   d-1): own cell c
 *            middle (x-1, d-2): own cell c-1 of d-2, or what arrived as lower one step ago
 *   backward lower  (x+1, e+1): own cell c+1, or lane i+1's first cell (DPP wave_shl:1)
 *            upper  (x,   e+1): own cell c
 *            middle (x+1, e+2): own cell c+1 of e+2, or what arrived as lower one step ago
 * The k-mer constants of a lane's positions live in registers and are reloaded
 * only when the window slides; event constants move one lane per anti-diagonal
 * with the same DPP shift, the lane at the window edge loading the next event
 * one step ahead.  Per-diagonal band bounds are read 64 diagonals at a time (one
 * per lane) and broadcast with readlane.
 *
 * Arithmetic per cell is identical to the LDS kernels of sa_kernels.hip: the
 * same terms, in the same order, through the same lookup log-add.
 */
#include "sa_device.h"

#include <type_traits>

namespace sa {

/* ------------------------------------------------------------------------ */
/* per-position and per-event constants                                       */
/* ------------------------------------------------------------------------ */

/* EK 0: Gaussian tables as the .model loader builds them (extra-event table = match
 * table with level_sd x 1.75, one skip emission for every k-mer): the extra-event
 * emission follows from the match constants.  EK 1: any Gaussian tables.  EK 2: HDP,
 * densities gathered by the sweep.  EK 3: HDP, each cell's density emission read
 * from the emission plane hdp_plane_kernel filled (no per-position or per-event
 * constants in registers). */
template <typename T, int EK> struct KCon;
template <typename T> struct KCon<T, 0> {
    T mmu, misd, mca, mnmu, mnb;
};
template <typename T> struct KCon<T, 1> {
    T gx, mmu, misd, mca, mnmu, mnb, ymu, yisd, yca, ynmu, ynb;
};
template <typename T> struct KCon<T, 2> {
    T hmu;
    int32_t row;
};

template <typename T> struct KCon<T, 3> {};

/* Element i of one sub-alignment's plane, addressed as the plane's base (uniform, in
   scalar registers) plus a 32-bit byte offset in one vector register rather than a
   64-bit vector sum (the backward sweep's loads; the forward sweeps, short of scalar
   registers, keep 64-bit addresses); the runtime gives the register sweeps only
   sub-alignments whose planes stay under 4 GB. */
template <typename T> __device__ __forceinline__ T &pl(T *base, int64_t i) {
    return *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + (uint32_t)i * (uint32_t)sizeof(T));
}
template <typename T> __device__ __forceinline__ const T &pl(const T *base, int64_t i) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + (uint32_t)i * (uint32_t)sizeof(T));
}

/* the three states of one diagonal's cells of a lane (forward sweeps) */
template <typename T, int CPT> struct FStates { T M[CPT], X[CPT], Y[CPT]; };

template <typename T, int EK> struct ECon;
template <typename T> struct ECon<T, 0> { T m, n, l, i; };
template <typename T> struct ECon<T, 1> { T m, n, l, i; };
template <typename T> struct ECon<T, 2> { T m; };
template <typename T> struct ECon<T, 3> {};

template <typename T>
__device__ __forceinline__ void load_kcon(KCon<T, 3> &, const Model<T> &, const int32_t *, int64_t, bool) {}
template <typename T> __device__ __forceinline__ void load_econ(ECon<T, 3> &, const Args<T> &, int64_t, bool) {}

/* k-mer (x-1) of cell x; an absent k-mer (x = 0 or past the end) gives -inf emissions */
template <typename T>
__device__ __forceinline__ void load_kcon(KCon<T, 0> &k, const Model<T> &m, const int32_t *kid, int64_t pos,
                                          bool ok) {
    if (ok) {
        const int32_t id = kid[pos];
        k.mmu = m.mmu[id]; k.misd = m.misd[id]; k.mca = m.mca[id]; k.mnmu = m.mnmu[id]; k.mnb = m.mnb[id];
    } else {
        k.mmu = 0; k.misd = 0; k.mca = ninf<T>(); k.mnmu = 1; k.mnb = 0;
    }
}
template <typename T>
__device__ __forceinline__ void load_kcon(KCon<T, 1> &k, const Model<T> &m, const int32_t *kid, int64_t pos,
                                          bool ok) {
    if (ok) {
        const int32_t id = kid[pos];
        k.gx = m.gx[id];
        k.mmu = m.mmu[id]; k.misd = m.misd[id]; k.mca = m.mca[id]; k.mnmu = m.mnmu[id]; k.mnb = m.mnb[id];
        k.ymu = m.ymu[id]; k.yisd = m.yisd[id]; k.yca = m.yca[id]; k.ynmu = m.ynmu[id]; k.ynb = m.ynb[id];
    } else {
        const T NI = ninf<T>();
        k.gx = NI;
        k.mmu = 0; k.misd = 0; k.mca = NI; k.mnmu = 1; k.mnb = 0;
        k.ymu = 0; k.yisd = 0; k.yca = NI; k.ynmu = 1; k.ynb = 0;
    }
}
template <typename T>
__device__ __forceinline__ void load_kcon(KCon<T, 2> &k, const Model<T> &m, const int32_t *kid, int64_t pos,
                                          bool ok) {
    if (ok) {
        const int32_t id = kid[pos];
        k.hmu = m.hmu[id];
        k.row = m.hrow[id];
    } else {
        k.hmu = 0;
        k.row = -1;
    }
}

/* event j; an absent event gives finite constants (its cells have no in-band
 * neighbour that reads it) */
template <typename T, int EK>
__device__ __forceinline__ void load_econ(ECon<T, EK> &e, const Args<T> &a, int64_t j, bool ok) {
    if (ok) { e.m = a.evm[j]; e.n = a.evn[j]; e.l = a.evl[j]; e.i = a.evi[j]; }
    else { e.m = 0; e.n = 1; e.l = 0; e.i = 1; }
}
template <typename T>
__device__ __forceinline__ void load_econ(ECon<T, 2> &e, const Args<T> &a, int64_t j, bool ok) {
    e.m = ok ? a.evm[j] : (T)0;
}

template <typename T, int EK>
__device__ __forceinline__ ECon<T, EK> shr1(const ECon<T, EK> &e, const ECon<T, EK> &f) {
    return ECon<T, EK>{shr1(e.m, f.m), shr1(e.n, f.n), shr1(e.l, f.l), shr1(e.i, f.i)};
}
template <typename T, int EK>
__device__ __forceinline__ ECon<T, EK> shl1(const ECon<T, EK> &e, const ECon<T, EK> &f) {
    return ECon<T, EK>{shl1(e.m, f.m), shl1(e.n, f.n), shl1(e.l, f.l), shl1(e.i, f.i)};
}
template <typename T> __device__ __forceinline__ ECon<T, 2> shr1(const ECon<T, 2> &e, const ECon<T, 2> &f) {
    return ECon<T, 2>{shr1(e.m, f.m)};
}
template <typename T> __device__ __forceinline__ ECon<T, 2> shl1(const ECon<T, 2> &e, const ECon<T, 2> &f) {
    return ECon<T, 2>{shl1(e.m, f.m)};
}
template <typename T> __device__ __forceinline__ ECon<T, 3> shr1(const ECon<T, 3> &, const ECon<T, 3> &) {
    return ECon<T, 3>{};
}
template <typename T> __device__ __forceinline__ ECon<T, 3> shl1(const ECon<T, 3> &, const ECon<T, 3> &) {
    return ECon<T, 3>{};
}

/* the match/gapX constants of the next position, for the backward sweep */
template <typename T> __device__ __forceinline__ KCon<T, 0> shl1_match(const KCon<T, 0> &k) {
    return KCon<T, 0>{shl1(k.mmu, (T)0), shl1(k.misd, (T)0), shl1(k.mca, ninf<T>()), shl1(k.mnmu, (T)1),
                      shl1(k.mnb, (T)0)};
}
template <typename T> __device__ __forceinline__ KCon<T, 1> shl1_match(const KCon<T, 1> &k) {
    KCon<T, 1> r = k;
    const T NI = ninf<T>();
    r.gx = shl1(k.gx, NI);
    r.mmu = shl1(k.mmu, (T)0); r.misd = shl1(k.misd, (T)0); r.mca = shl1(k.mca, NI);
    r.mnmu = shl1(k.mnmu, (T)1); r.mnb = shl1(k.mnb, (T)0);
    return r;
}
template <typename T> __device__ __forceinline__ KCon<T, 2> shl1_match(const KCon<T, 2> &k) {
    return KCon<T, 2>{shl1(k.hmu, (T)0), shl1(k.row, -1)};
}
template <typename T> __device__ __forceinline__ KCon<T, 3> shl1_match(const KCon<T, 3> &) { return KCon<T, 3>{}; }

/* emissions from registers.  Model-file tables (EK 0): both emissions of one k-mer and
   event at once, sharing the standardised mean and the noise term.  The backward sweep
   takes a cell's match emission from its right neighbour's evaluation on the diagonal
   before. */
template <typename T>
__device__ __forceinline__ void em_both(const Model<T> &, const KCon<T, 0> &k, const ECon<T, 0> &e, T s, T,
                                        T &eM, T &eY) {
    gauss_both_std(k.mmu, k.misd, k.mca, k.mnmu, k.mnb, e.m, e.n, e.l, e.i, s, eM, eY);
}
template <typename T>
__device__ __forceinline__ T em_gapx(const Model<T> &m, const KCon<T, 0> &k) {
    return k.mca > ninf<T>() ? m.gx_const : ninf<T>();
}
template <typename T>
__device__ __forceinline__ T em_match(const Model<T> &, const KCon<T, 1> &k, const ECon<T, 1> &e, T s, T) {
    return gauss_emit(k.mmu, k.misd, k.mca, k.mnmu, k.mnb, e.m, e.n, e.l, e.i, s);
}
template <typename T>
__device__ __forceinline__ T em_gapy(const Model<T> &, const KCon<T, 1> &k, const ECon<T, 1> &e, T s, T) {
    return gauss_emit(k.ymu, k.yisd, k.yca, k.ynmu, k.ynb, e.m, e.n, e.l, e.i, s);
}
template <typename T>
__device__ __forceinline__ void em_both(const Model<T> &m, const KCon<T, 1> &k, const ECon<T, 1> &e, T s, T iv,
                                        T &eM, T &eY) {
    eM = em_match<T>(m, k, e, s, iv);
    eY = em_gapy<T>(m, k, e, s, iv);
}
template <typename T> __device__ __forceinline__ T em_gapx(const Model<T> &, const KCon<T, 1> &k) { return k.gx; }
template <typename T>
__device__ __forceinline__ T em_match(const Model<T> &m, const KCon<T, 2> &k, const ECon<T, 2> &e, T s, T iv) {
    return hdp_emit_row(m, k.row, k.hmu, e.m, s, iv);
}
template <typename T>
__device__ __forceinline__ T em_gapy(const Model<T> &m, const KCon<T, 2> &k, const ECon<T, 2> &e, T s, T iv) {
    return hdp_emit_row(m, k.row, k.hmu, e.m, s, iv);
}
template <typename T>
__device__ __forceinline__ void em_both(const Model<T> &m, const KCon<T, 2> &k, const ECon<T, 2> &e, T s, T iv,
                                        T &eM, T &eY) {
    eM = hdp_emit_row(m, k.row, k.hmu, e.m, s, iv);
    eY = eM;
}
template <typename T>
__device__ __forceinline__ T em_gapx(const Model<T> &, const KCon<T, 2> &) { return (T)-2.3025850929940455; }
template <typename T>
__device__ __forceinline__ T em_gapx(const Model<T> &, const KCon<T, 3> &) { return (T)-2.3025850929940455; }

/* Uniform-address loads through the constant address space become scalar loads:
 * they are counted by lgkmcnt, so waiting for one does not wait for the vector
 * stores of the forward planes (vmcnt counts vector loads and stores together). */
template <typename T> __device__ __forceinline__ T ld_scalar(const T *p, int i) {
    typedef __attribute__((address_space(4))) const T CT;
    return ((const CT *)p)[__builtin_amdgcn_readfirstlane(i)];
}
/* byte i of p through the scalar cache: the aligned dword holding it (flag arrays
 * are allocated with 4 bytes of padding) */
__device__ __forceinline__ int ld_scalar_byte(const int8_t *p, int i) {
    const uintptr_t ad = (uintptr_t)p + (uintptr_t)__builtin_amdgcn_readfirstlane(i);
    const int32_t w = ld_scalar((const int32_t *)(ad & ~(uintptr_t)3), 0);
    return (w >> (int)((ad & 3) * 8)) & 0xff;
}
/* the aligned dword holding flag byte i, extracted later (get) so that the load's wait
   falls where the flag is used */
struct FlagWord {
    int32_t w, sh;
    __device__ __forceinline__ int get() const { return (w >> sh) & 0xff; }
};
__device__ __forceinline__ FlagWord ld_flag_word(const int8_t *p, int i) {
    const uintptr_t ad = (uintptr_t)p + (uintptr_t)__builtin_amdgcn_readfirstlane(i);
    return FlagWord{ld_scalar((const int32_t *)(ad & ~(uintptr_t)3), 0), (int)((ad & 3) * 8)};
}
template <typename T>
__device__ __forceinline__ void load_econ_uniform(ECon<T, 2> &e, const Args<T> &a, int64_t base, int j, bool ok) {
    e.m = ok ? ld_scalar(a.evm + base, j) : (T)0;
}
template <typename T>
__device__ __forceinline__ void load_econ_uniform(ECon<T, 3> &, const Args<T> &, int64_t, int, bool) {}
template <typename T, int EK>
__device__ __forceinline__ void load_econ_uniform(ECon<T, EK> &e, const Args<T> &a, int64_t base, int j, bool ok) {
    if (ok) {
        e.m = ld_scalar(a.evm + base, j); e.n = ld_scalar(a.evn + base, j);
        e.l = ld_scalar(a.evl + base, j); e.i = ld_scalar(a.evi + base, j);
    } else {
        e.m = 0; e.n = 1; e.l = 0; e.i = 1;
    }
}

/* Make the window-slide reloads complete inside the (rare) slide branch: an empty
 * asm that reads the registers.  Otherwise the first use on the common path
 * carries a vmcnt wait that also waits for the previous diagonal's stores. */
__device__ __forceinline__ void sink(float v) { asm volatile("" ::"v"(v)); }
__device__ __forceinline__ void sink(double v) { asm volatile("" ::"v"(v)); }
__device__ __forceinline__ void sink(int v) { asm volatile("" ::"v"(v)); }
template <typename T> __device__ __forceinline__ void sink(const KCon<T, 0> &k) {
    sink(k.mmu); sink(k.misd); sink(k.mca); sink(k.mnmu); sink(k.mnb);
}
template <typename T> __device__ __forceinline__ void sink(const KCon<T, 1> &k) {
    sink(k.gx); sink(k.mmu); sink(k.misd); sink(k.mca); sink(k.mnmu); sink(k.mnb);
    sink(k.ymu); sink(k.yisd); sink(k.yca); sink(k.ynmu); sink(k.ynb);
}
template <typename T> __device__ __forceinline__ void sink(const KCon<T, 2> &k) { sink(k.hmu); sink(k.row); }
template <typename T> __device__ __forceinline__ void sink(const KCon<T, 3> &) {}
template <typename T> __device__ __forceinline__ void sink(const ECon<T, 2> &e) { sink(e.m); }
template <typename T> __device__ __forceinline__ void sink(const ECon<T, 3> &) {}
template <typename T, int EK> __device__ __forceinline__ void sink(const ECon<T, EK> &e) {
    sink(e.m); sink(e.n); sink(e.l); sink(e.i);
}

template <typename T> __device__ __forceinline__ T shfl_idx(T v, int src, T fill) {
    T r = __shfl(v, src & 63, 64);
    return (src >= 0 && src < 64) ? r : fill;
}

__device__ __forceinline__ int rdl(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
/* lane `lane`'s value (uniform lane index) */
__device__ __forceinline__ double readlane(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

/* ------------------------------------------------------------------------ */
/* forward                                                                   */
/* ------------------------------------------------------------------------ */

template <typename T, int CPT, int EK, int NW>
__global__ void __launch_bounds__(64 * NW) forward_reg_kernel(Args<T> a, const int32_t *__restrict__ order) {
    static_assert(NW > 1, "one-wavefront sweeps are forward_sw_kernel");
    constexpr bool NORM = sizeof(T) == 4;
    constexpr int WL = 64 * NW;   /* lanes of the workgroup */
    constexpr int W = WL * CPT;   /* positions of the window */
    /* the last cell of each wavefront crosses to the next through LDS, double-buffered
       by diagonal parity (one barrier per diagonal); the wavefront maxima of a
       renormalised diagonal cross at the same barrier */
    __shared__ T bx[2][NW][3];
    __shared__ T red[NW];
    __shared__ T shbuf[6 * W];
    __shared__ LutEntry<T> lut[LutSize<T>::n];
    /* the several-wavefront sweeps carry the longest dependency chains of a launch
       sequence: their instructions go first on shared SIMDs (the launch's priority) */
    set_priority(a.prio);
    lut_fill(lut, (int)threadIdx.x);
    __syncthreads();
    const int ti = order[blockIdx.x];
    const TaskDev t = a.tasks[ti];
    const Model<T> &m = a.m;
    const int gl = threadIdx.x;
    const int lane = gl & 63;
    const int wave = __builtin_amdgcn_readfirstlane(gl >> 6); /* uniform: scalar branches on it */
    const int woff = 64 * CPT * wave; /* the wavefront's first position in the window */
    const int lX = t.lX, lY = t.lY, N = lX + lY;
    const int32_t *xlo = a.xlo + t.diag_off;
    const int32_t *xhi = a.xhi + t.diag_off;
    const int8_t *dfl = a.dflag + t.diag_off;
    int32_t *doff = a.doff + t.diag_off;
    int32_t *doffxy = a.doffxy + t.diag_off;
    double *foff = a.foff + t.diag_off;
    const int32_t *kid = a.kid + t.pos_off;
    const int64_t evc = t.evc_off;
    T *fM = a.fM + t.cell_off;
    T *fX = a.fX + t.xy_off;
    T *fY = a.fY + t.xy_off;
    const T s = (T)(m.kind == 1 ? 1.0 : t.scale / t.var);
    const T iv = (T)(1.0 / t.var);
    const T tMC = m.trans[0], tMFX = m.trans[1], tMFY = m.trans[2], tGOX = m.trans[3], tGOY = m.trans[4];
    const T tGEX = m.trans[5], tGEY = m.trans[6], tSWX = m.trans[7];
    const T NI = ninf<T>();
    const bool useSWX = m.use_swx != 0;

    /* band bounds and keep-flag of each diagonal come through the scalar cache, one
       diagonal ahead; the per-diagonal plane offsets are written 64 diagonals at a
       time (lane j of wave 0 holds diagonal dc + j) */
    int dc = 0;
    int md_doff = 0, md_dxy = -1;
    double md_foff = 0.0;

    int X0 = 0;
    KCon<T, EK> kc[CPT];
    ECon<T, EK> ev[CPT];
    T pM[CPT], pX[CPT], pY[CPT], qM[CPT], qX[CPT], qY[CPT];
    T LpM = NI, LpX = NI, LpY = NI;

    /* diagonal 0: start or ragged-start values (impl/stateMachine.c:1022-1031) */
    const int lo0 = ld_scalar(xlo, 0), hi0 = ld_scalar(xhi, 0), f0 = ld_scalar_byte(dfl, 0);
    int nlo = 0, nhi = -1, nfl = 0;
    if (N >= 1) { nlo = ld_scalar(xlo, 1); nhi = ld_scalar(xhi, 1); nfl = ld_scalar_byte(dfl, 1); }
#pragma unroll
    for (int c = 0; c < CPT; c++) {
        const int x = CPT * gl + c;
        const bool v = x >= lo0 && x <= hi0;
        pM[c] = v ? (t.ragged_left ? NI : (T)0) : NI;
        pX[c] = v ? (t.ragged_left ? (T)0 : NI) : NI;
        pY[c] = pX[c];
        qM[c] = NI; qX[c] = NI; qY[c] = NI;
        if (v) {
            fM[x - lo0] = pM[c];
            if (f0) { fX[x - lo0] = pX[c]; fY[x - lo0] = pY[c]; }
        }
        load_kcon(kc[c], m, kid, x - 1, x >= 1 && x <= lX);
        const int j = 1 - x - 1;
        load_econ(ev[c], a, evc + j, j >= 0 && j < lY);
    }
    if (gl == 0) { md_doff = 0; md_dxy = f0 ? 0 : -1; md_foff = 0.0; }
    if (lane == 63) { bx[0][wave][0] = pM[CPT - 1]; bx[0][wave][1] = pX[CPT - 1]; bx[0][wave][2] = pY[CPT - 1]; }
    __syncthreads();
    int64_t run = hi0 - lo0 + 1;
    int64_t runxy = f0 ? run : 0;
    double off1 = 0.0, off2 = 0.0;
    int lo1 = lo0, lo2 = lo0;
    ECon<T, EK> pre;
    /* the event the wavefront's first cell reads on the next diagonal (each wavefront
       loads its own through the scalar cache; cell x on d reads event d - x - 1) */
    load_econ_uniform(pre, a, evc, 1 - X0 - woff, 1 - X0 - woff >= 0 && 1 - X0 - woff < lY);
    bool fresh = true;
    T mxp = 0; /* the renormalisation shift of d-1 (its edge cells crossed unshifted) */
    constexpr bool PF = EK == 2;
    HdpFetch<T> gp[PF ? CPT : 1];
    bool have_gp = false;
    /* emission plane: the lane's cells of the next diagonal, loaded a diagonal ahead */
    constexpr bool PL = EK == 3;
    const T *em = PL ? a.em + t.cell_off : nullptr;
    T pe[PL ? CPT : 1];
    bool have_pe = false;
    bool bad = false;

    for (int d = 1; d <= N; d++) {
        int j = d - dc;
        if (j == 64) {
            if (wave == 0 && dc + lane <= N) {
                doff[dc + lane] = md_doff; doffxy[dc + lane] = md_dxy; foff[dc + lane] = md_foff;
            }
            dc += 64;
            j = 0;
        }
        const int lo = nlo, hi = nhi, fl = nfl;
        if (d < N) { nlo = ld_scalar(xlo, d + 1); nhi = ld_scalar(xhi, d + 1); nfl = ld_scalar_byte(dfl, d + 1); }
        /* band bounds are non-decreasing; anything else goes to the LDS kernels.  The
           sweep runs on to its end (its results are discarded) so that the loop has a
           single exit: early returns cost the loop its register assignment. */
        bad |= lo < X0;
        /* fill values for the first lane of each wavefront: the last cell of the
           previous wavefront on d-1 (and on d-2 for the middle state) */
        T fLM = NI, fLX = NI, fLY = NI;
        if (wave > 0) {
            const int pb = (d - 1) & 1;
            fLM = bx[pb][wave - 1][0] - mxp; fLX = bx[pb][wave - 1][1] - mxp; fLY = bx[pb][wave - 1][2] - mxp;
        }
        if (hi > X0 + W - 1) {
            /* slide the window as far as the two previous diagonals allow */
            const int minlo = lo1 < lo2 ? lo1 : lo2;
            int smax = (minlo - X0) / CPT;
            const int sreq = (hi - (X0 + W - 1) + CPT - 1) / CPT;
            bad |= smax < sreq;
            smax = smax < 0 ? 0 : smax;
            {
#pragma unroll
                for (int c = 0; c < CPT; c++) {
                    const int i = CPT * gl + c;
                    shbuf[0 * W + i] = pM[c]; shbuf[1 * W + i] = pX[c]; shbuf[2 * W + i] = pY[c];
                    shbuf[3 * W + i] = qM[c]; shbuf[4 * W + i] = qX[c]; shbuf[5 * W + i] = qY[c];
                }
                __syncthreads();
                const bool in = gl + smax < WL;
#pragma unroll
                for (int c = 0; c < CPT; c++) {
                    const int i = CPT * (gl + smax) + c;
                    pM[c] = in ? shbuf[0 * W + i] : NI; pX[c] = in ? shbuf[1 * W + i] : NI;
                    pY[c] = in ? shbuf[2 * W + i] : NI; qM[c] = in ? shbuf[3 * W + i] : NI;
                    qX[c] = in ? shbuf[4 * W + i] : NI; qY[c] = in ? shbuf[5 * W + i] : NI;
                }
                /* the previous position's last cell: source lane gl-1+smax */
                const int sl = gl - 1 + smax;
                const bool ok = sl >= 0 && sl < WL;
                const int i = CPT * sl + CPT - 1;
                fLM = ok ? shbuf[0 * W + i] : NI; fLX = ok ? shbuf[1 * W + i] : NI; fLY = ok ? shbuf[2 * W + i] : NI;
                LpM = ok ? shbuf[3 * W + i] : NI; LpX = ok ? shbuf[4 * W + i] : NI; LpY = ok ? shbuf[5 * W + i] : NI;
                __syncthreads();
            }
            X0 += smax * CPT;
#pragma unroll
            for (int c = 0; c < CPT; c++) {
                const int x = X0 + CPT * gl + c;
                load_kcon(kc[c], m, kid, x - 1, x >= 1 && x <= lX);
                const int jj = d - x - 1;
                load_econ(ev[c], a, evc + jj, jj >= 0 && jj < lY);
            }
#pragma unroll
            for (int c = 0; c < CPT; c++) { sink(kc[c]); sink(ev[c]); }
            fresh = true;
            have_gp = false;
            have_pe = false;
        }
        if (!fresh) {
            /* cell x on d reads event d-x-1: its left neighbour's event of d-1 */
            const ECon<T, EK> inc = shr1(ev[CPT - 1], pre);
#pragma unroll
            for (int c = CPT - 1; c >= 1; c--) ev[c] = ev[c - 1];
            ev[0] = inc;
        }
        fresh = false;
        {
            const int jn = d - X0 - woff; /* event of the wavefront's first cell on d+1 */
            load_econ_uniform(pre, a, evc, jn, jn >= 0 && jn < lY);
        }
        const T LM = shr1(pM[CPT - 1], fLM), LX = shr1(pX[CPT - 1], fLX), LY = shr1(pY[CPT - 1], fLY);
        const T dl2 = NORM ? (T)(off2 - off1) : (T)0;
        /* the rows are renormalised every 8th diagonal, so d-2 and d-1 carry different
           offsets only on the diagonal after one */
        const bool shift2 = NORM && off2 != off1;
        const bool renorm = NORM && (d & 7) == 0;
        if (shift2) {
            /* d-2 carries its own offset: bring its states to d-1's */
#pragma unroll
            for (int c = 0; c < CPT; c++) { qM[c] += dl2; qX[c] += dl2; qY[c] += dl2; }
            LpM += dl2; LpX += dl2; LpY += dl2;
        }
        T nM[CPT], nX[CPT], nY[CPT];
        T mx = NI;
#pragma unroll
        for (int c = 0; c < CPT; c++) {
            const int x = X0 + CPT * gl + c;
            const bool valid = x >= lo && x <= hi;
            const T lM = c ? pM[c - 1] : LM, lXv = c ? pX[c - 1] : LX, lYv = c ? pY[c - 1] : LY;
            const T mM = c ? qM[c - 1] : LpM, mXv = c ? qX[c - 1] : LpX, mYv = c ? qY[c - 1] : LpY;
            const T eX = em_gapx<T>(m, kc[c]);
            T eM, eY;
            if constexpr (PL) {
                eM = have_pe ? pe[PL ? c : 0] : valid ? em[run + (x - lo)] : (T)0;
                eY = eM;
            } else if (PF && have_gp) {
                eM = hdp_finish(m, gp[PF ? c : 0], iv);
                eY = eM;
            } else {
                em_both<T>(m, kc[c], ev[c], s, iv, eM, eY);
            }
            /* lower -> gapX, middle -> match, upper -> gapY */
            T X = lM + (eX + tGOX);
            X = logadd_lut(X, lXv + (eX + tGEX), lut);
            if (useSWX) X = logadd_lut(X, lYv + (eX + tSWX), lut);
            T M = mM + (eM + tMC);
            M = logadd_lut(M, mXv + (eM + tMFX), lut);
            M = logadd_lut(M, mYv + (eM + tMFY), lut);
            T Y = pM[c] + (eY + tGOY);
            Y = logadd_lut(Y, pY[c] + (eY + tGEY), lut);
            nM[c] = valid ? M : NI;
            nX[c] = valid ? X : NI;
            nY[c] = valid ? Y : NI;
            if (renorm) {
                T m3 = nM[c] > nX[c] ? nM[c] : nX[c];
                m3 = m3 > nY[c] ? m3 : nY[c];
                mx = m3 > mx ? m3 : mx;
            }
        }
        if constexpr (PF) {
            /* HDP: the density gathers of diagonal d+1 are issued here, before the
               barrier and the stores, and used a diagonal later (two diagonals ahead
               gained nothing: the wait before their use also covers every later
               store and gather).  The events of d+1 are this diagonal's shifted one
               lane; each wavefront's first lane takes `pre` */
            const ECon<T, EK> inc = shr1(ev[CPT - 1], pre);
#pragma unroll
            for (int c = 0; c < CPT; c++) {
                const T evm = c ? ev[c - 1].m : inc.m;
                gp[c] = hdp_fetch(m, kc[c].row, kc[c].hmu, evm, s);
            }
            have_gp = true;
        }
        if constexpr (PL) {
            /* the emissions of d+1, read before the barrier and the stores */
            const int64_t nrun = run + (hi - lo + 1);
#pragma unroll
            for (int c = 0; c < CPT; c++) {
                const int x = X0 + CPT * gl + c;
                pe[c] = d < N && x >= nlo && x <= nhi ? em[nrun + (x - nlo)] : (T)0;
            }
            have_pe = true;
        }
        /* renormalisation: the wavefront maxima cross at the diagonal's barrier with
           the edge cells (unshifted; the next diagonal's reader applies the shift) */
        if (renorm) {
            mx = wave_max(mx);
            if (lane == 0) red[wave] = mx;
        }
        if (lane == 63) {
            const int pb = d & 1;
            bx[pb][wave][0] = nM[CPT - 1]; bx[pb][wave][1] = nX[CPT - 1]; bx[pb][wave][2] = nY[CPT - 1];
        }
        __syncthreads();
        double offd = off1;
        mxp = 0;
        if (renorm) {
            mx = red[0];
#pragma unroll
            for (int w = 1; w < NW; w++) mx = red[w] > mx ? red[w] : mx;
            if (mx > NI) {
#pragma unroll
                for (int c = 0; c < CPT; c++) { nM[c] -= mx; nX[c] -= mx; nY[c] -= mx; }
                offd = off1 + (double)mx;
                mxp = mx;
            }
        }
#pragma unroll
        for (int c = 0; c < CPT; c++) {
            const int x = X0 + CPT * gl + c;
            if (x >= lo && x <= hi) {
                fM[run + (x - lo)] = nM[c];
                if (fl) { fX[runxy + (x - lo)] = nX[c]; fY[runxy + (x - lo)] = nY[c]; }
            }
        }
        if (gl == j) { md_doff = (int32_t)run; md_dxy = fl ? (int32_t)runxy : -1; md_foff = offd; }
        run += hi - lo + 1;
        if (fl) runxy += hi - lo + 1;
        LpM = LM; LpX = LX; LpY = LY;
#pragma unroll
        for (int c = 0; c < CPT; c++) {
            qM[c] = pM[c]; qX[c] = pX[c]; qY[c] = pY[c];
            pM[c] = nM[c]; pX[c] = nX[c]; pY[c] = nY[c];
        }
        off2 = off1;
        off1 = offd;
        lo2 = lo1;
        lo1 = lo;
    }
    if (wave == 0 && dc + lane <= N) { doff[dc + lane] = md_doff; doffxy[dc + lane] = md_dxy; foff[dc + lane] = md_foff; }
    if (bad && gl == 0) atomicOr(a.err, ERR_WINDOWS);
}

/* ------------------------------------------------------------------------ */
/* forward, several wavefronts synchronised every kHaloK diagonals           */
/* ------------------------------------------------------------------------ */

/* Wave w's lane i holds position X0 + kHaloS*w + i; for w > 0 its first kHaloH lanes
 * repeat the last positions of wave w-1 (a halo), the other lanes are its own.  Between
 * two synchronisations a wavefront runs on its own: its first lane takes -inf for the
 * positions left of it.  A cell reads x-1 on d-1 and d-2 and x on d-1, so the wrong
 * values spread by at most one position per diagonal and stay inside the halo for
 * kHaloK = kHaloH diagonals.  Every kHaloK-th diagonal (where the FP32 rows are also
 * renormalised) each wavefront writes the states of d and d-1 at the positions it owns
 * to LDS, the window slides if the next kHaloK diagonals need it, and every lane reads
 * its position's states back: the halo lanes are exact again.  One barrier per
 * kHaloK diagonals instead of one per diagonal; the arithmetic of every stored cell is
 * the other sweeps' own, so the planes are bit-identical to theirs.  Each diagonal step
 * is one cell's dependency chain; its scalar loads (band bounds and flag, the first
 * lane's event) are issued two diagonals ahead, so that the step's LDS waits (the
 * log-add table; they share a counter with scalar loads) find them done.  The table
 * form of the log-add measured faster here than selecting the cubic by compares
 * (0.70 against 0.76 us per diagonal on a 50 kb read alone).  The host picks
 * this sweep when the band fits the window at every synchronisation:
 * xhi(min(d + kHaloK, N)) - min(xlo(d-1), xlo(d)) + 1 <= W at d = 0, kHaloK, 2 kHaloK, ...
 * (prep_kernel's second span per sub-alignment). */
template <typename T, int EK, int NW>
__global__ void __launch_bounds__(64 * NW) forward_halo_kernel(Args<T> a, const int32_t *__restrict__ order) {
    static_assert(NW > 1, "one-wavefront sweeps are forward_sw_kernel");
    constexpr bool NORM = sizeof(T) == 4;
    constexpr int W = 64 + kHaloS * (NW - 1); /* positions of the window */
    /* states of d and d-1 by position, double-buffered by synchronisation parity (a
       wavefront writes the next buffer only after the barrier every wavefront reaches
       once it has read this one) */
    __shared__ T shb[2][6 * W];
    __shared__ T red[2][NW];
    __shared__ LutEntry<T> lut[LutSize<T>::n];
    set_priority(a.prio);
    lut_fill(lut, (int)threadIdx.x);
    __syncthreads();
    const int ti = order[blockIdx.x];
    const TaskDev t = a.tasks[ti];
    const Model<T> &m = a.m;
    const int gl = threadIdx.x;
    const int lane = gl & 63;
    const int wave = __builtin_amdgcn_readfirstlane(gl >> 6); /* uniform: scalar branches on it */
    const int woff = kHaloS * wave;                            /* the wavefront's first position */
    const bool own = wave == 0 || lane >= kHaloH;              /* position owned (stored, exchanged) */
    const int lX = t.lX, lY = t.lY, N = lX + lY;
    const int32_t *xlo = a.xlo + t.diag_off;
    const int32_t *xhi = a.xhi + t.diag_off;
    const int8_t *dfl = a.dflag + t.diag_off;
    int32_t *doff = a.doff + t.diag_off;
    int32_t *doffxy = a.doffxy + t.diag_off;
    double *foff = a.foff + t.diag_off;
    const int32_t *kid = a.kid + t.pos_off;
    const int64_t evc = t.evc_off;
    T *fM = a.fM + t.cell_off;
    T *fX = a.fX + t.xy_off;
    T *fY = a.fY + t.xy_off;
    const T s = (T)(m.kind == 1 ? 1.0 : t.scale / t.var);
    const T iv = (T)(1.0 / t.var);
    const T tMC = m.trans[0], tMFX = m.trans[1], tMFY = m.trans[2], tGOX = m.trans[3], tGOY = m.trans[4];
    const T tGEX = m.trans[5], tGEY = m.trans[6], tSWX = m.trans[7];
    const T NI = ninf<T>();
    const bool useSWX = m.use_swx != 0;

    /* plane offsets of each diagonal, written 64 diagonals at a time by wave 0 (lane j
       holds diagonal dc + j) */
    int dc = 0;
    int md_doff = 0, md_dxy = -1;
    double md_foff = 0.0;

    int X0 = 0;
    KCon<T, EK> kc;
    ECon<T, EK> ev;
    T pM, pX, pY, qM, qX, qY;
    T LpM = NI, LpX = NI, LpY = NI;

    /* diagonal 0: start or ragged-start values (impl/stateMachine.c:1022-1031); every
       lane evaluates its own position, so the halo starts exact */
    const int lo0 = ld_scalar(xlo, 0), hi0 = ld_scalar(xhi, 0), f0 = ld_scalar_byte(dfl, 0);
    /* bounds and flag of d+1 and d+2: each diagonal's scalar loads are issued two
       diagonals before their use */
    int nlo = 0, nhi = -1, nlo2 = 0, nhi2 = -1;
    FlagWord nfl{0, 0}, nfl2{0, 0};
    if (N >= 1) { nlo = ld_scalar(xlo, 1); nhi = ld_scalar(xhi, 1); nfl = ld_flag_word(dfl, 1); }
    if (N >= 2) { nlo2 = ld_scalar(xlo, 2); nhi2 = ld_scalar(xhi, 2); nfl2 = ld_flag_word(dfl, 2); }
    {
        const int x = woff + lane;
        const bool v = x >= lo0 && x <= hi0;
        pM = v ? (t.ragged_left ? NI : (T)0) : NI;
        pX = v ? (t.ragged_left ? (T)0 : NI) : NI;
        pY = pX;
        qM = NI; qX = NI; qY = NI;
        if (v && own) {
            fM[x - lo0] = pM;
            if (f0) { fX[x - lo0] = pX; fY[x - lo0] = pY; }
        }
        load_kcon(kc, m, kid, x - 1, x >= 1 && x <= lX);
        const int j = 1 - x - 1;
        load_econ(ev, a, evc + j, j >= 0 && j < lY);
        /* complete the loads here: a load still pending at the loop's entry makes every
           iteration wait on the vector memory counter, which also counts the stores */
        sink(kc);
        sink(ev);
    }
    if (gl == 0) { md_doff = 0; md_dxy = f0 ? 0 : -1; md_foff = 0.0; }
    int64_t run = hi0 - lo0 + 1;
    int64_t runxy = f0 ? run : 0;
    double off1 = 0.0, off2 = 0.0;
    int lo1 = lo0;
    /* pre: the event the wavefront's first lane reads on d+1 (cell x on d reads event
       d - x - 1); pre2, the one of d+2, loaded a diagonal earlier through the scalar
       cache.  Diagonal 1's events are loaded above, so pre2 starts as diagonal 2's. */
    ECon<T, EK> pre, pre2;
    load_econ_uniform(pre2, a, evc, 1 - X0 - woff, 1 - X0 - woff >= 0 && 1 - X0 - woff < lY);
    pre = pre2;
    bool fresh = true;
    constexpr bool PF = EK == 2;
    HdpFetch<T> gp;
    bool have_gp = false;
    constexpr bool PL = EK == 3;
    const T *em = PL ? a.em + t.cell_off : nullptr;
    T pe = (T)0;
    bool have_pe = false;
    bool bad = false;
    int sp = 0; /* synchronisation parity */

    for (int d = 1; d <= N; d++) {
        int j = d - dc;
        if (j == 64) {
            if (wave == 0 && dc + lane <= N) {
                doff[dc + lane] = md_doff; doffxy[dc + lane] = md_dxy; foff[dc + lane] = md_foff;
            }
            dc += 64;
            j = 0;
        }
        const int lo = nlo, hi = nhi;
        const FlagWord flw = nfl;
        nlo = nlo2; nhi = nhi2; nfl = nfl2;
        if (d + 2 <= N) { nlo2 = ld_scalar(xlo, d + 2); nhi2 = ld_scalar(xhi, d + 2); nfl2 = ld_flag_word(dfl, d + 2); }
        const bool sync = (d % kHaloK) == 0;
        /* the band's right end kHaloK diagonals on (the next synchronisation's window) */
        const int hnext = sync ? ld_scalar(xhi, d + kHaloK < N ? d + kHaloK : N) : 0;
        bad |= lo < X0; /* bounds are non-decreasing; the sweep runs on to a single exit */
        if (!fresh) {
            /* cell x on d reads event d-x-1: its left neighbour's event of d-1 */
            ev = shr1(ev, pre);
        }
        fresh = false;
        pre = pre2; /* the first lane's event on d+1 */
        {
            const int jn = d + 1 - X0 - woff; /* on d+2 */
            load_econ_uniform(pre2, a, evc, jn, jn >= 0 && jn < lY);
        }
        const T LM = shr1(pM, NI), LX = shr1(pX, NI), LY = shr1(pY, NI);
        if (NORM && off2 != off1) {
            /* d-2 carries its own offset (the diagonal after a renormalised one): bring
               its states to d-1's */
            const T dl2 = (T)(off2 - off1);
            qM += dl2; qX += dl2; qY += dl2;
            LpM += dl2; LpX += dl2; LpY += dl2;
        }
        T nM, nX, nY;
        {
            const int x = X0 + woff + lane;
            const bool valid = x >= lo && x <= hi;
            const T eX = em_gapx<T>(m, kc);
            T eM, eY;
            if constexpr (PL) {
                eM = have_pe ? pe : valid ? em[run + (x - lo)] : (T)0;
                eY = eM;
            } else if (PF && have_gp) {
                eM = hdp_finish(m, gp, iv);
                eY = eM;
            } else {
                em_both<T>(m, kc, ev, s, iv, eM, eY);
            }
            /* lower -> gapX, middle -> match, upper -> gapY */
            T X = LM + (eX + tGOX);
            X = logadd_lut(X, LX + (eX + tGEX), lut);
            if (useSWX) X = logadd_lut(X, LY + (eX + tSWX), lut);
            T M = LpM + (eM + tMC);
            M = logadd_lut(M, LpX + (eM + tMFX), lut);
            M = logadd_lut(M, LpY + (eM + tMFY), lut);
            T Y = pM + (eY + tGOY);
            Y = logadd_lut(Y, pY + (eY + tGEY), lut);
            nM = valid ? M : NI;
            nX = valid ? X : NI;
            nY = valid ? Y : NI;
        }
        if constexpr (PF) {
            /* HDP: the density gathers of diagonal d+1, issued a diagonal ahead */
            const ECon<T, EK> inc = shr1(ev, pre);
            gp = hdp_fetch(m, kc.row, kc.hmu, inc.m, s);
            have_gp = true;
        }
        if constexpr (PL) {
            /* the emissions of d+1, read a diagonal ahead */
            const int64_t nrun = run + (hi - lo + 1);
            const int x = X0 + woff + lane;
            pe = d < N && x >= nlo && x <= nhi ? em[nrun + (x - nlo)] : (T)0;
            have_pe = true;
        }
        double offd = off1;
        if (sync) {
            /* renormalisation (FP32) over the positions owned; the states of d and d-1
               at every owned position cross through LDS */
            if constexpr (NORM) {
                T m3 = nM > nX ? nM : nX;
                m3 = m3 > nY ? m3 : nY;
                const T mx = wave_max(own ? m3 : NI);
                if (lane == 0) red[sp][wave] = mx;
            }
            T *sb = shb[sp];
            if (own) {
                const int i = woff + lane;
                sb[0 * W + i] = nM; sb[1 * W + i] = nX; sb[2 * W + i] = nY;
                sb[3 * W + i] = pM; sb[4 * W + i] = pX; sb[5 * W + i] = pY;
            }
            __syncthreads();
            /* slide as far as d and d-1 allow when the next kHaloK diagonals need it */
            int X0n = X0;
            if (hnext > X0 + W - 1) {
                X0n = lo1 < lo ? lo1 : lo;
                X0n = X0n > X0 ? X0n : X0;
                bad |= hnext > X0n + W - 1;
            }
            {
                const int i = X0n - X0 + woff + lane;
                const bool in = i < W;
                const int ii = in ? i : 0;
                nM = in ? sb[0 * W + ii] : NI; nX = in ? sb[1 * W + ii] : NI; nY = in ? sb[2 * W + ii] : NI;
                pM = in ? sb[3 * W + ii] : NI; pX = in ? sb[4 * W + ii] : NI; pY = in ? sb[5 * W + ii] : NI;
            }
            if constexpr (NORM) {
                T mx = red[sp][0];
#pragma unroll
                for (int w = 1; w < NW; w++) mx = red[sp][w] > mx ? red[sp][w] : mx;
                if (mx > NI) {
                    nM -= mx; nX -= mx; nY -= mx;
                    offd = off1 + (double)mx;
                }
            }
            sp ^= 1;
            if (X0n != X0) {
                X0 = X0n;
                /* constants of the new positions for diagonal d+1 (cell x reads k-mer
                   x-1 and event d-x) */
                const int x = X0 + woff + lane;
                load_kcon(kc, m, kid, x - 1, x >= 1 && x <= lX);
                const int jj = d - x;
                load_econ(ev, a, evc + jj, jj >= 0 && jj < lY);
                sink(kc);
                sink(ev);
                /* the first lane's event on d+2 from the new window */
                const int jn = d + 1 - X0 - woff;
                load_econ_uniform(pre2, a, evc, jn, jn >= 0 && jn < lY);
                fresh = true;
                have_gp = false;
                have_pe = false;
            }
        }
        const int fl = flw.get(); /* extracted here, a diagonal after its load was issued */
        {
            const int x = X0 + woff + lane;
            if (own && x >= lo && x <= hi) {
                fM[run + (x - lo)] = nM;
                if (fl) { fX[runxy + (x - lo)] = nX; fY[runxy + (x - lo)] = nY; }
            }
        }
        if (gl == j) { md_doff = (int32_t)run; md_dxy = fl ? (int32_t)runxy : -1; md_foff = offd; }
        run += hi - lo + 1;
        if (fl) runxy += hi - lo + 1;
        qM = pM; qX = pX; qY = pY;
        pM = nM; pX = nX; pY = nY;
        if (sync) {
            /* the middle neighbours of d+1 from the exchanged states of d-1 */
            LpM = shr1(qM, NI); LpX = shr1(qX, NI); LpY = shr1(qY, NI);
        } else {
            LpM = LM; LpX = LX; LpY = LY;
        }
        off2 = off1;
        off1 = offd;
        lo1 = lo;
    }
    if (wave == 0 && dc + lane <= N) { doff[dc + lane] = md_doff; doffxy[dc + lane] = md_dxy; foff[dc + lane] = md_foff; }
    if (bad && gl == 0) atomicOr(a.err, ERR_WINDOWS);
}

/* ------------------------------------------------------------------------ */
/* forward, one wavefront, window width following the band                  */
/* ------------------------------------------------------------------------ */

/* Register-window class of a block of 64 diagonals (the narrowest of 1, 2, 3 cells per
 * lane that holds the band there), limited to the widest window of the launch; a
 * launch wider than three cells per lane keeps its width. */
__device__ __forceinline__ int window_class(int stored, int cmax) {
    if (cmax > 3) return cmax;
    return stored < 1 || stored > cmax ? cmax : stored;
}

/* What a one-wavefront forward sweep carries from one window width to the next. */
struct FwdRun {
    int d;        /* next diagonal to compute */
    int X0;       /* window of the previous segment, its width */
    int wprev;
    int dc;       /* first diagonal of the current block of 64 (plane offsets, classes) */
    int md_doff, md_dxy;
    double md_foff;
    int64_t run, runxy;
    double off1;  /* offset of d-1 (the two diagonals a cell reads share it) */
    int lo1, lo2;
    int ncls;     /* stored class byte of the next block */
    bool bad;
};

/* Diagonals st.d.. of one sub-alignment with CPT cells per lane, until the band's class
 * changes at a block boundary (returns the new class; the states of the last two
 * diagonals are left in xs, indexed by position - X0) or the sweep ends (returns 0).
 * xs holds 6 planes of stride xsw. */
template <typename T, int CPT, int EK>
__device__ __forceinline__ int fwd_segment(const Args<T> &a, const TaskDev &t, FwdRun &st, T *xs, int xsw,
                                           int cmax, const LutEntry<T> *lut) {
    constexpr bool NORM = sizeof(T) == 4;
    constexpr int W = 64 * CPT;
    const Model<T> &m = a.m;
    const int lane = threadIdx.x;
    const int lX = t.lX, lY = t.lY, N = lX + lY;
    const int32_t *xlo = a.xlo + t.diag_off;
    const int32_t *xhi = a.xhi + t.diag_off;
    const int8_t *dfl = a.dflag + t.diag_off;
    const int8_t *dcl = a.dcls + t.diag_off;
    int32_t *doff = a.doff + t.diag_off;
    int32_t *doffxy = a.doffxy + t.diag_off;
    double *foff = a.foff + t.diag_off;
    const int32_t *kid = a.kid + t.pos_off;
    const int64_t evc = t.evc_off;
    T *fM = a.fM + t.cell_off;
    T *fX = a.fX + t.xy_off;
    T *fY = a.fY + t.xy_off;
    const T s = (T)(m.kind == 1 ? 1.0 : t.scale / t.var);
    const T iv = (T)(1.0 / t.var);
    const T tMC = m.trans[0], tMFX = m.trans[1], tMFY = m.trans[2], tGOX = m.trans[3], tGOY = m.trans[4];
    const T tGEX = m.trans[5], tGEY = m.trans[6], tSWX = m.trans[7];
    const T NI = ninf<T>();
    const bool useSWX = m.use_swx != 0;

    KCon<T, EK> kc[CPT];
    ECon<T, EK> ev[CPT];
    T pM[CPT], pX[CPT], pY[CPT], qM[CPT], qX[CPT], qY[CPT];
    T LpM, LpX, LpY;
    using FSt = FStates<T, CPT>;
    int d = st.d;
    int X0;
    if (d == 0) {
        /* diagonal 0: start or ragged-start values (impl/stateMachine.c:1022-1031) */
        X0 = 0;
        const int lo0 = ld_scalar(xlo, 0), hi0 = ld_scalar(xhi, 0), f0 = ld_scalar_byte(dfl, 0);
#pragma unroll
        for (int c = 0; c < CPT; c++) {
            const int x = CPT * lane + c;
            const bool v = x >= lo0 && x <= hi0;
            pM[c] = v ? (t.ragged_left ? NI : (T)0) : NI;
            pX[c] = v ? (t.ragged_left ? (T)0 : NI) : NI;
            pY[c] = pX[c];
            qM[c] = NI; qX[c] = NI; qY[c] = NI;
            if (v) {
                fM[x - lo0] = pM[c];
                if (f0) { fX[x - lo0] = pX[c]; fY[x - lo0] = pY[c]; }
            }
            load_kcon(kc[c], m, kid, x - 1, x >= 1 && x <= lX);
            const int j = 1 - x - 1;
            load_econ(ev[c], a, evc + j, j >= 0 && j < lY);
        }
        if (lane == 0) { st.md_doff = 0; st.md_dxy = f0 ? 0 : -1; st.md_foff = 0.0; }
        st.run = hi0 - lo0 + 1;
        st.runxy = f0 ? st.run : 0;
        st.off1 = 0.0;
        st.lo1 = lo0; st.lo2 = lo0;
        LpM = NI; LpX = NI; LpY = NI;
        d = 1;
    } else {
        /* a new width: the window starts at the leftmost cell the diagonal and the two
           before it hold; the states of d-1 and d-2 come from xs */
        const int lo = ld_scalar(xlo, d);
        X0 = lo < st.lo1 ? lo : st.lo1;
        X0 = X0 < st.lo2 ? X0 : st.lo2;
#pragma unroll
        for (int c = 0; c < CPT; c++) {
            const int x = X0 + CPT * lane + c;
            const int o = x - st.X0;
            const bool in = o >= 0 && o < st.wprev;
            const int oi = in ? o : 0;
            pM[c] = in ? xs[0 * xsw + oi] : NI; pX[c] = in ? xs[1 * xsw + oi] : NI;
            pY[c] = in ? xs[2 * xsw + oi] : NI; qM[c] = in ? xs[3 * xsw + oi] : NI;
            qX[c] = in ? xs[4 * xsw + oi] : NI; qY[c] = in ? xs[5 * xsw + oi] : NI;
            load_kcon(kc[c], m, kid, x - 1, x >= 1 && x <= lX);
            const int jj = d - x - 1;
            load_econ(ev[c], a, evc + jj, jj >= 0 && jj < lY);
        }
        {
            const int o = X0 + CPT * lane - 1 - st.X0; /* middle of the lane's first cell */
            const bool in = o >= 0 && o < st.wprev;
            const int oi = in ? o : 0;
            LpM = in ? xs[3 * xsw + oi] : NI; LpX = in ? xs[4 * xsw + oi] : NI; LpY = in ? xs[5 * xsw + oi] : NI;
        }
#pragma unroll
        for (int c = 0; c < CPT; c++) { sink(kc[c]); sink(ev[c]); }
        __syncthreads(); /* xs is read before the next segment writes it */
    }
    int nlo = 0, nhi = -1, nfl = 0;
    if (d <= N) { nlo = ld_scalar(xlo, d); nhi = ld_scalar(xhi, d); nfl = ld_scalar_byte(dfl, d); }
    ECon<T, EK> pre;
    load_econ_uniform(pre, a, evc, d - X0, d - X0 >= 0 && d - X0 < lY);
    bool fresh = true;
    bool bad = st.bad;
    constexpr bool PF = EK == 2;
    HdpFetch<T> gp[PF ? CPT : 1];
    bool have_gp = false;
    /* emission plane: the lane's cells of the next diagonal, loaded a diagonal ahead */
    constexpr bool PL = EK == 3;
    const T *em = PL ? a.em + t.cell_off : nullptr;
    T pe[PL ? CPT : 1];
    bool have_pe = false;
    int64_t run = st.run, runxy = st.runxy;
    double off1 = st.off1;
    int lo1 = st.lo1, lo2 = st.lo2;
    int dc = st.dc;
    int md_doff = st.md_doff, md_dxy = st.md_dxy;
    double md_foff = st.md_foff;
    int ncls = st.ncls;
    int next = 0;
    const int d0 = d; /* the segment's first diagonal */

    /* One diagonal: P holds d-1's states, Q d-2's, Nw receives d's (at three cells per
       lane with rotated events a third set, so that the roles repeat every third diagonal
       as the events do; otherwise Q itself, cells from the right, so that each of d-2's
       states is read before its register is reused).  False: the segment ends before
       d.  The loop below rotates the roles, so no state is copied between diagonals.
       Events: at two or three cells per lane cell c's event is in ev[(c - K) mod CPT]
       on the loop's K-th diagonal (K = ph), so the shift by one cell per diagonal is a
       renaming too; elsewhere the registers are shifted. */
    /* (FP64 at three cells per lane: the third set costs more than the copies) */
    constexpr bool ROT = (CPT == 2 || (CPT == 3 && sizeof(T) == 4)) && EK != 3;
    auto step = [&](FSt &P, FSt &Q, FSt &Nw, auto ph) -> bool {
        constexpr int K = decltype(ph)::value;
        auto evr = [&](int c) -> ECon<T, EK> & { return ev[ROT ? ((c - K) % CPT + CPT) % CPT : c]; };
        int j = d - dc;
        const int lo = nlo, hi = nhi, fl = nfl;
        if (j == 64) {
            if (dc + lane <= N) { doff[dc + lane] = md_doff; doffxy[dc + lane] = md_dxy; foff[dc + lane] = md_foff; }
            dc += 64;
            j = 0;
            const int want = window_class(ncls & 7, cmax);
            ncls = d + 64 <= N ? ld_scalar_byte(dcl, d + 64) : 0;
            if (want != CPT) {
                next = want;
                return false;
            }
        }
        if (d < N) { nlo = ld_scalar(xlo, d + 1); nhi = ld_scalar(xhi, d + 1); nfl = ld_scalar_byte(dfl, d + 1); }
        /* band bounds are non-decreasing; anything else goes to the LDS kernels.  The
           sweep runs on to its end (its results are discarded) so that the loop has a
           single exit per segment: early returns cost the loop its register assignment. */
        bad |= lo < X0;
        if (hi > X0 + W - 1) {
            /* the band leaves the window: the segment ends here and starts again at d
               with the window starting at the leftmost cell of d, d-1 and d-2 (the way
               a width change re-seats it).  Sliding inside the loop would give the
               k-mer and event registers a second definition, which costs the loop
               register copies and memory waits on every diagonal. */
            if (d != d0) {
                next = CPT;
                return false;
            }
            bad = true; /* the window holds fewer positions than the band: the LDS sweeps redo it */
        }
        if (!fresh) {
            /* cell x on d reads event d-x-1: its left neighbour's event of d-1 (the
               register of the lane's last cell on d-1 becomes the first cell's) */
            if constexpr (ROT) {
                evr(0) = shr1(evr(0), pre);
            } else {
                const ECon<T, EK> inc = shr1(ev[CPT - 1], pre);
#pragma unroll
                for (int c = CPT - 1; c >= 1; c--) ev[c] = ev[c - 1];
                ev[0] = inc;
            }
        }
        fresh = false;
        {
            const int jn = d - X0; /* event of the first cell of the window on d+1 */
            load_econ_uniform(pre, a, evc, jn, jn >= 0 && jn < lY);
        }
        T LM = shr1(P.M[CPT - 1], NI), LX = shr1(P.X[CPT - 1], NI), LY = shr1(P.Y[CPT - 1], NI);
        const bool renorm = NORM && (d & 7) == 0;
        T mx = NI;
#pragma unroll
        for (int c = CPT - 1; c >= 0; c--) {
            const int x = X0 + CPT * lane + c;
            const bool valid = x >= lo && x <= hi;
            const T lM = c ? P.M[c - 1] : LM, lXv = c ? P.X[c - 1] : LX, lYv = c ? P.Y[c - 1] : LY;
            const T mM = c ? Q.M[c - 1] : LpM, mXv = c ? Q.X[c - 1] : LpX, mYv = c ? Q.Y[c - 1] : LpY;
            const T eX = em_gapx<T>(m, kc[c]);
            T eM, eY;
            if constexpr (PL) {
                eM = have_pe ? pe[PL ? c : 0] : valid ? em[run + (x - lo)] : (T)0;
                eY = eM;
            } else if (PF && have_gp) {
                eM = hdp_finish(m, gp[PF ? c : 0], iv);
                eY = eM;
            } else {
                em_both<T>(m, kc[c], evr(c), s, iv, eM, eY);
            }
            /* lower -> gapX, middle -> match, upper -> gapY */
            T X = lM + (eX + tGOX);
            X = logadd_lut(X, lXv + (eX + tGEX), lut);
            if (useSWX) X = logadd_lut(X, lYv + (eX + tSWX), lut);
            T M = mM + (eM + tMC);
            M = logadd_lut(M, mXv + (eM + tMFX), lut);
            M = logadd_lut(M, mYv + (eM + tMFY), lut);
            T Y = P.M[c] + (eY + tGOY);
            Y = logadd_lut(Y, P.Y[c] + (eY + tGEY), lut);
            Nw.M[c] = valid ? M : NI;
            Nw.X[c] = valid ? X : NI;
            Nw.Y[c] = valid ? Y : NI;
            if (renorm) {
                T m3 = Nw.M[c] > Nw.X[c] ? Nw.M[c] : Nw.X[c];
                m3 = m3 > Nw.Y[c] ? m3 : Nw.Y[c];
                mx = m3 > mx ? m3 : mx;
            }
        }
        double offd = off1;
        if (renorm) {
            /* every 8th diagonal: the largest state becomes 0, its log carried in the
               diagonal's offset; d-1's states (d-2's on the next diagonal) move with them,
               so that the two diagonals a cell reads always share one offset */
            mx = wave_max(mx);
            if (mx > NI) {
#pragma unroll
                for (int c = 0; c < CPT; c++) {
                    Nw.M[c] -= mx; Nw.X[c] -= mx; Nw.Y[c] -= mx;
                    P.M[c] -= mx; P.X[c] -= mx; P.Y[c] -= mx;
                }
                LM -= mx; LX -= mx; LY -= mx;
                offd = off1 + (double)mx;
            }
        }
        if constexpr (PF) {
            /* HDP: the density gathers of diagonal d+1 are issued at the end of diagonal
               d, before its stores, and used a diagonal later.  The events of d+1 are
               this diagonal's shifted one lane; the first lane's comes from `pre` */
            const ECon<T, EK> inc = shr1(evr(CPT - 1), pre);
#pragma unroll
            for (int c = 0; c < CPT; c++) {
                const T evm = c ? evr(c - 1).m : inc.m;
                gp[c] = hdp_fetch(m, kc[c].row, kc[c].hmu, evm, s);
            }
            have_gp = true;
        }
        if constexpr (PL) {
            /* the emissions of d+1, read before this diagonal's stores */
            const int64_t nrun = run + (hi - lo + 1);
#pragma unroll
            for (int c = 0; c < CPT; c++) {
                const int x = X0 + CPT * lane + c;
                pe[c] = d < N && x >= nlo && x <= nhi ? em[nrun + (x - nlo)] : (T)0;
            }
            have_pe = true;
        }
#pragma unroll
        for (int c = 0; c < CPT; c++) {
            const int x = X0 + CPT * lane + c;
            if (x >= lo && x <= hi) {
                fM[run + (x - lo)] = Nw.M[c];
                if (fl) { fX[runxy + (x - lo)] = Nw.X[c]; fY[runxy + (x - lo)] = Nw.Y[c]; }
            }
        }
        if (lane == j) { md_doff = (int32_t)run; md_dxy = fl ? (int32_t)runxy : -1; md_foff = offd; }
        run += hi - lo + 1;
        if (fl) runxy += hi - lo + 1;
        LpM = LM; LpX = LX; LpY = LY;
        off1 = offd;
        lo2 = lo1;
        lo1 = lo;
        d++;
        return true;
    };
    FSt A, B, C; /* the states of d-1 and d-2, and (three cells per lane, rotated events) of d */
#pragma unroll
    for (int c = 0; c < CPT; c++) {
        A.M[c] = pM[c]; A.X[c] = pX[c]; A.Y[c] = pY[c];
        B.M[c] = qM[c]; B.X[c] = qX[c]; B.Y[c] = qY[c];
    }
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    int role = 0; /* d-1's and d-2's states: 0 in A and B, 1 in B and A (C and A with the
                     third set), 2 in B and C */
    if constexpr (CPT == 3 && ROT) {
        while (d <= N) {
            role = 0;
            if (!step(A, B, C, I0{})) break;
            role = 1;
            if (d > N || !step(C, A, B, I1{})) break;
            role = 2;
            if (d > N || !step(B, C, A, I2{})) break;
            role = 0;
        }
    } else {
        while (d <= N) {
            role = 0;
            if (!step(A, B, B, I0{})) break;
            role = 1;
            if (d > N || !step(B, A, A, I1{})) break;
            role = 0;
        }
    }
    if (next) {
        const FSt &P = role == 0 ? A : CPT == 3 && ROT ? (role == 1 ? C : B) : B;
        const FSt &Q = role == 0 ? B : CPT == 3 && ROT ? (role == 1 ? A : C) : A;
#pragma unroll
        for (int c = 0; c < CPT; c++) {
            const int i = CPT * lane + c;
            xs[0 * xsw + i] = P.M[c]; xs[1 * xsw + i] = P.X[c]; xs[2 * xsw + i] = P.Y[c];
            xs[3 * xsw + i] = Q.M[c]; xs[4 * xsw + i] = Q.X[c]; xs[5 * xsw + i] = Q.Y[c];
        }
        __syncthreads();
    }
    st.d = d;
    st.X0 = X0;
    st.wprev = W;
    st.dc = dc;
    st.md_doff = md_doff; st.md_dxy = md_dxy; st.md_foff = md_foff;
    st.run = run; st.runxy = runxy;
    st.off1 = off1;
    st.lo1 = lo1; st.lo2 = lo2;
    st.ncls = ncls;
    st.bad = bad;
    return next;
}

/* One wavefront per sub-alignment; the window runs at 1, 2 or 3 cells per lane as the
 * band's width allows (block classes from prep_kernel), never wider than CMAX. */
/* The one-wavefront forward sweep of one sub-alignment (the kernel below and the
 * persistent sweeps' forward item). */
template <typename T, int CMAX, int EK>
__device__ __forceinline__ void fwd_task(const Args<T> &a, const TaskDev t, T *xs, const LutEntry<T> *lut) {
    const int lane = threadIdx.x;
    const int N = t.lX + t.lY;
    const int8_t *dcl = a.dcls + t.diag_off;
    FwdRun st;
    st.d = 0; st.X0 = 0; st.wprev = 0; st.dc = 0;
    st.md_doff = 0; st.md_dxy = -1; st.md_foff = 0.0;
    st.run = 0; st.runxy = 0; st.off1 = 0.0; st.lo1 = 0; st.lo2 = 0;
    st.ncls = 64 <= N ? ld_scalar_byte(dcl, 64) : 0;
    st.bad = false;
    int cls = window_class(ld_scalar_byte(dcl, 0) & 7, CMAX);
    while (cls) {
        if constexpr (CMAX == 1 || CMAX > 3) {
            cls = fwd_segment<T, CMAX, EK>(a, t, st, xs, 64 * CMAX, CMAX, lut);
        } else {
            if (cls == 1) cls = fwd_segment<T, 1, EK>(a, t, st, xs, 64 * CMAX, CMAX, lut);
            else if (CMAX == 3 && cls == 2) cls = fwd_segment<T, 2, EK>(a, t, st, xs, 64 * CMAX, CMAX, lut);
            else cls = fwd_segment<T, CMAX, EK>(a, t, st, xs, 64 * CMAX, CMAX, lut);
        }
    }
    int32_t *doff = a.doff + t.diag_off;
    int32_t *doffxy = a.doffxy + t.diag_off;
    double *foff = a.foff + t.diag_off;
    if (st.dc + lane <= N) { doff[st.dc + lane] = st.md_doff; doffxy[st.dc + lane] = st.md_dxy; foff[st.dc + lane] = st.md_foff; }
    if (st.bad && lane == 0) atomicOr(a.err, ERR_WINDOWS);
}

template <typename T, int CMAX, int EK>
__global__ void __launch_bounds__(64) forward_sw_kernel(Args<T> a, const int32_t *__restrict__ order) {
    extern __shared__ __align__(16) unsigned char dsm[];
    T *xs = reinterpret_cast<T *>(dsm); /* [6][64 * CMAX]: states across a width change */
    __shared__ LutEntry<T> lut[LutSize<T>::n];
    set_priority(a.prio);
    lut_fill(lut, (int)threadIdx.x);
    __syncthreads();
    fwd_task<T, CMAX, EK>(a, a.tasks[order[blockIdx.x]], xs, lut);
}

/* ------------------------------------------------------------------------ */
/* backward + total refresh + posterior pairs                                 */
/* ------------------------------------------------------------------------ */

/* Sequential log-add chain of dpDiagonal_dotProduct (impl/pairwiseAligner.c:1014-1027)
 * over the terms held CPT per lane in x order.  Terms at least 7.5 below the running
 * maximum are exact no-ops of logAdd and are skipped; terms more than 24 below the
 * largest term are skipped too (each weighs under 4e-11 of the total; together they move
 * the log total by under 1e-8).  The remaining few terms are visited in order with a
 * ballot walk. */
template <int CPT>
__device__ __forceinline__ double chain_reg(const double (&t)[CPT], const LutEntry<double> *lutd) {
    double mx = -__builtin_huge_val();
#pragma unroll
    for (int c = 0; c < CPT; c++) mx = t[c] > mx ? t[c] : mx;
    mx = wave_max(mx);
    const double cut = mx - 24.0;
    unsigned long long act[CPT];
    unsigned long long any = 0;
#pragma unroll
    for (int c = 0; c < CPT; c++) {
        act[c] = __ballot(t[c] > cut);
        any |= act[c];
    }
    /* the walk runs on wave-uniform values: each term is read from its lane into
       scalar registers, the log-add reads its cubic from the LDS table */
    double acc = -__builtin_huge_val(), pm = -__builtin_huge_val();
    while (any) {
        const int L = __builtin_ctzll(any);
        any &= any - 1;
#pragma unroll
        for (int c = 0; c < CPT; c++) {
            if ((act[c] >> L) & 1ull) {
                const double v = readlane(t[c], L);
                if (v > pm - 7.5) {
                    acc = logadd_lut(acc, v, lutd);
                    pm = v > pm ? v : pm;
                }
            }
        }
    }
    return acc;
}

/* What a backward sweep carries from one window width to the next. */
struct BwdRun {
    int e;          /* next diagonal to compute */
    int X0, wprev;  /* window of the previous segment, its width */
    int ncls;       /* stored class byte of the next block down */
    double ob1;     /* offset of B[e+1] (B[e+2] is moved to it at each renormalisation) */
    int hi1, hi2;
    int cursor, posterior_count;
    double total;
    bool bad;
};

/* Diagonals st.e down to T+1 of one traceback window with CPT cells per lane, until the
 * band's class changes at a block boundary (returns the new class; the states of e+1
 * and e+2 are left in xs, indexed by position - X0) or the window ends (returns 0).
 * xs holds 4 planes of stride xsw. */
template <typename T, int CPT, int EK>
__device__ __forceinline__ int bwd_segment(const Args<T> &a, const int wi, const WinDev &Wd, const TaskDev &t,
                                           BwdRun &st, T *xs, int xsw, int cmax, const LutEntry<T> *lut,
                                           const LutEntry<double> *lutd) {
    constexpr bool NORM = sizeof(T) == 4;
    constexpr int W = 64 * CPT;
    const Model<T> &m = a.m;
    const int lane = threadIdx.x;
    const int lX = t.lX, lY = t.lY;
    const int32_t *xlo = a.xlo + t.diag_off;
    const int32_t *xhi = a.xhi + t.diag_off;
    const int32_t *doff = a.doff + t.diag_off;
    const int32_t *doffxy = a.doffxy + t.diag_off;
    const double *foff = a.foff + t.diag_off;
    const int8_t *dcl = a.dcls + t.diag_off;
    const int32_t *kid = a.kid + t.pos_off;
    const int64_t evc = t.evc_off;
    const T *fM = a.fM + t.cell_off;
    const T *fX = a.fX + t.xy_off;
    const T *fY = a.fY + t.xy_off;
    const T s = (T)(m.kind == 1 ? 1.0 : t.scale / t.var);
    const T iv = (T)(1.0 / t.var);
    const T tMC = m.trans[0], tMFX = m.trans[1], tMFY = m.trans[2], tGOX = m.trans[3], tGOY = m.trans[4];
    const T tGEX = m.trans[5], tGEY = m.trans[6], tSWX = m.trans[7];
    const T NI = ninf<T>();
    const bool useSWX = m.use_swx != 0;
    const T thr = (T)a.threshold;
    /* candidates for the threshold: exponents at least log(threshold) - 1e-4 */
    const T lthr = a.threshold > 0.0 ? (T)(log(a.threshold) - 1e-4) : -__builtin_huge_valf();
    const int D = Wd.D, TF = Wd.TF, TB = Wd.T;
    const bool ragged_end = Wd.last && t.ragged_right;
    const T e0 = ragged_end ? m.rendv[0] : m.endv[0];
    const T e1 = ragged_end ? m.rendv[1] : m.endv[1];
    const T e2 = ragged_end ? m.rendv[2] : m.endv[2];
    int32_t *out = a.out + 3 * Wd.out_off;

    const int es = st.e;
    /* band bounds and plane offsets of each diagonal come through the scalar cache,
       one diagonal ahead */
    int nlo = ld_scalar(xlo, es), nhi = ld_scalar(xhi, es), ndo = ld_scalar(doff, es);
    double nfo = ld_scalar(foff, es);
    int X0;
    KCon<T, EK> kc[CPT];
    ECon<T, EK> ev[CPT];
    T pM[CPT], pX[CPT], pY[CPT], qM[CPT];
    T RpM; /* lane i+1's first cell of e+2, match state */
    using FSt = FStates<T, CPT>;
    if (es == D) {
        X0 = nhi - W + 1;
#pragma unroll
        for (int c = 0; c < CPT; c++) {
            pM[c] = NI; pX[c] = NI; pY[c] = NI; qM[c] = NI;
        }
        RpM = NI;
    } else {
        /* a new width: the window ends at the rightmost cell of e, e+1 and e+2; their
           states come from xs */
        int rh = nhi > st.hi1 ? nhi : st.hi1;
        rh = rh > st.hi2 ? rh : st.hi2;
        X0 = rh - W + 1;
#pragma unroll
        for (int c = 0; c < CPT; c++) {
            const int o = X0 + CPT * lane + c - st.X0;
            const bool in = o >= 0 && o < st.wprev;
            const int oi = in ? o : 0;
            pM[c] = in ? xs[0 * xsw + oi] : NI; pX[c] = in ? xs[1 * xsw + oi] : NI;
            pY[c] = in ? xs[2 * xsw + oi] : NI; qM[c] = in ? xs[3 * xsw + oi] : NI;
        }
        const int o = X0 + CPT * lane + CPT - st.X0;
        const bool in = o >= 0 && o < st.wprev;
        RpM = in ? xs[3 * xsw + (in ? o : 0)] : NI;
        __syncthreads(); /* xs is read before the next segment writes it */
    }
#pragma unroll
    for (int c = 0; c < CPT; c++) {
        const int x = X0 + CPT * lane + c;
        load_kcon(kc[c], m, kid, x - 1, x >= 1 && x <= lX);
        const int j = es - x;
        load_econ(ev[c], a, evc + j, j >= 0 && j < lY);
    }
    if (es != D) {
#pragma unroll
        for (int c = 0; c < CPT; c++) { sink(kc[c]); sink(ev[c]); }
    }
    /* the extra-event skip emission of k-mer x+1 for the lane's last cell (a segment
       keeps its k-mers).  FP64 Gaussian sweeps, which evaluate every match emission,
       shift that k-mer in on every diagonal instead: one register more costs them
       spills */
    constexpr bool XN = EK >= 2 || sizeof(T) == 4;
    const T eXn = XN ? em_gapx<T>(m, shl1_match(kc[0])) : (T)0;
    ECon<T, EK> pre;
    {
        const int jn = (es - 1) - (X0 + W - 1); /* event of lane 63's last cell on es-1 */
        load_econ_uniform(pre, a, evc, jn, jn >= 0 && jn < lY);
    }
    double ob1 = st.ob1; /* offset of B[e+1] */
    int hi1 = st.hi1, hi2 = st.hi2;    /* xhi of e+1, e+2 (-1: outside the window of diagonals) */
    int cursor = st.cursor, posterior_count = st.posterior_count;
    double total = st.total;
    int ncls = st.ncls;
    bool fresh = true;
    bool bad = st.bad;
    int next = 0;
    /* The match emission of cell x on e (k-mer x+1, event e-x) is the emission of the
       k-mer and event its right neighbour evaluated on e+1: each cell evaluates its own
       k-mer against its event once (both emissions, em_both) and keeps the match one
       for its left neighbour's next diagonal.  HDP: the match and extra-event emissions
       are the same density, its gathers issued a diagonal ahead */
    constexpr bool PF = EK == 2;
    HdpFetch<T> gp[PF ? CPT : 1];
    /* emission plane (EK 3): the same reuse, the extra-event emissions of e-1's cells
       (E(e, x)) loaded a diagonal ahead */
    constexpr bool PL = EK == 3;
    const T *em = PL ? a.em + t.cell_off : nullptr;
    T pe[PL ? CPT : 1];
    bool have_pe = false;
    /* FP64 Gaussian sweeps evaluate every match emission: the registers the reuse takes
       cost them more spills than the evaluation */
    constexpr bool RE = EK >= 2 || sizeof(T) == 4;
    /* the match emission of each cell's own k-mer and event: from e+1 until the cell
       reads its right neighbour's, then its own of e (cells run from the left, so each
       value is read before its register is rewritten) */
    T eYp[CPT];
    bool have_gp = false, have_eY = false;
    /* E(dd, xx) straight from the plane (window tops and slides) */
    auto plane_at = [&](int dd, int xx) -> T {
        if (dd > lX + lY) return (T)0;
        const int l = ld_scalar(xlo, dd), h = ld_scalar(xhi, dd), o = ld_scalar(doff, dd);
        return xx >= l && xx <= h ? pl(em, o + (xx - l)) : (T)0;
    };
    /* the forward match states of the next diagonal's cells for its posteriors, loaded
       a diagonal ahead (they would otherwise stall every diagonal on HBM latency) */
    T fpre[CPT];
    bool have_fpre = false;
    /* the forward states the next total refresh reads (gapX/gapY of its diagonal r,
       all three states of r-1), loaded a diagonal ahead as well */
    T rfx[CPT], rfy[CPT], rgm[CPT], rgx[CPT], rgy[CPT];
    bool have_rpre = false;
    auto load_refresh = [&](int r) {
        const int lor = ld_scalar(xlo, r), hir = ld_scalar(xhi, r), dxy = ld_scalar(doffxy, r);
        const bool h1r = r + 1 <= D; /* (r = D: no later diagonal in the window) */
        const int lo1 = h1r ? ld_scalar(xlo, r + 1) : 1, hi1r = h1r ? ld_scalar(xhi, r + 1) : 0;
        const int lom = ld_scalar(xlo, r - 1), him = ld_scalar(xhi, r - 1);
        const int dm = ld_scalar(doff, r - 1), dxym = ld_scalar(doffxy, r - 1);
#pragma unroll
        for (int c = 0; c < CPT; c++) {
            const int x = X0 + CPT * lane + c;
            const bool in = x >= lor && x <= hir;
            rfx[c] = in ? pl(fX, dxy + (x - lor)) : NI;
            rfy[c] = in ? pl(fY, dxy + (x - lor)) : NI;
            /* cell (x, r+1-x) of r+1; its middle (x-1, r-x) on r-1 */
            const bool hm = x >= lo1 && x <= hi1r && x - 1 >= lom && x - 1 <= him && x >= 1 && r + 1 - x >= 1;
            const int fi = x - 1 - lom;
            rgm[c] = hm ? pl(fM, dm + fi) : NI;
            rgx[c] = hm ? pl(fX, dxym + fi) : NI;
            rgy[c] = hm ? pl(fY, dxym + fi) : NI;
        }
    };

    int e = es;
    /* One diagonal: P holds e+1's states, Q e+2's (its match states only); Q receives
       e's (cells from the left, so that each of e+2's match states is read before its
       register is reused).  False: the segment ends before e.  The loop below
       alternates the two roles, so no state is copied between diagonals. */
    auto step = [&](FSt &P, FSt &Q) -> bool {
        if ((e & 63) == 63 && e != es) {
            /* top of block e/64: its class decides the width */
            const int want = window_class((ncls >> 3) & 7, cmax);
            ncls = e - 64 > TB ? ld_scalar_byte(dcl, e - 127) : 0;
            if (want != CPT) {
                next = want;
                return false;
            }
        }
        const int lo = nlo, hi = nhi, dof = ndo;
        const double fo = nfo;
        if (e - 1 > TB) {
            nlo = ld_scalar(xlo, e - 1); nhi = ld_scalar(xhi, e - 1); ndo = ld_scalar(doff, e - 1);
            nfo = ld_scalar(foff, e - 1);
        }
        T LowM = NI, LowX = NI; /* lane i+1's first cell of e+1 */
        double obe = 0.0;
        if (e == D) {
#pragma unroll
            for (int c = 0; c < CPT; c++) {
                const int x = X0 + CPT * lane + c;
                const bool valid = x >= lo && x <= hi;
                Q.M[c] = valid ? e0 : NI; Q.X[c] = valid ? e1 : NI; Q.Y[c] = valid ? e2 : NI;
            }
            fresh = false;
        } else {
            if (lo < X0) {
                /* the band leaves the window: the segment ends here and starts again at
                   e with the window ending at the rightmost cell of e, e+1 and e+2 (the
                   way a width change re-seats it).  Sliding inside the loop would give
                   the k-mer and event registers a second definition, which costs the
                   loop register copies and memory waits on every diagonal. */
                if (e != es) {
                    next = CPT;
                    return false;
                }
                bad = true; /* the window holds fewer positions than the band: the LDS sweeps redo it */
            }
            bad |= hi > X0 + W - 1;
            if (!fresh) {
                /* cell x on e reads event e-x: its right neighbour's event of e+1 */
                const ECon<T, EK> in = shl1(ev[0], pre);
#pragma unroll
                for (int c = 0; c < CPT - 1; c++) ev[c] = ev[c + 1];
                ev[CPT - 1] = in;
            }
            fresh = false;
            LowM = shl1(P.M[0], NI);
            LowX = shl1(P.X[0], NI);
            const T eYn = have_eY ? shl1(eYp[0], NI) : NI;
            const KCon<T, EK> knd = XN ? kc[0] : shl1_match(kc[0]); /* (FP64 Gaussian only) */
            const bool renorm = NORM && ((D - e) & 7) == 0;
            T mx = NI;
#pragma unroll
            for (int c = 0; c < CPT; c++) {
                const int x = X0 + CPT * lane + c;
                const bool valid = x >= lo && x <= hi;
                /* k-mer x+1 (the lane's last cell: its right neighbour's first, shifted
                   in only where an emission is evaluated from it) */
                auto kx = [&]() { return c < CPT - 1 ? kc[c < CPT - 1 ? c + 1 : 0] : XN ? shl1_match(kc[0]) : knd; };
                const T ClowX = c < CPT - 1 ? P.X[c + 1] : LowX;
                const T CupY = P.Y[c];
                const T CmidM = c < CPT - 1 ? Q.M[c + 1] : RpM;
                const T eX = c < CPT - 1 ? em_gapx<T>(m, kc[c < CPT - 1 ? c + 1 : 0]) : XN ? eXn : em_gapx<T>(m, knd);
                T eM, eY;
                if constexpr (PL) {
                    eY = have_pe ? pe[PL ? c : 0] : plane_at(e + 1, x);
                    eM = have_eY ? (c < CPT - 1 ? eYp[c < CPT - 1 ? c + 1 : 0] : eYn) : plane_at(e + 2, x + 1);
                    eYp[c] = eY;
                } else if constexpr (PF) {
                    eY = have_gp ? hdp_finish(m, gp[c], iv) : em_gapy<T>(m, kc[c], ev[c], s, iv);
                    eM = have_eY ? (c < CPT - 1 ? eYp[c < CPT - 1 ? c + 1 : 0] : eYn) : em_match<T>(m, kx(), ev[c], s, iv);
                    eYp[c] = eY;
                } else {
                    T mself; /* (FP64: the refresh below evaluates it again) */
                    em_both<T>(m, kc[c], ev[c], s, iv, mself, eY);
                    if (RE && have_eY) {
                        eM = c < CPT - 1 ? eYp[c < CPT - 1 ? c + 1 : 0] : eYn;
                    } else {
                        T unused;
                        em_both<T>(m, kx(), ev[c], s, iv, eM, unused);
                    }
                    eYp[c] = mself;
                }
                /* contributions arrive middle (e+2), then upper, then lower (e+1) */
                T M = CmidM + (eM + tMC);
                M = logadd_lut(M, CupY + (eY + tGOY), lut);
                M = logadd_lut(M, ClowX + (eX + tGOX), lut);
                T X = CmidM + (eM + tMFX);
                X = logadd_lut(X, ClowX + (eX + tGEX), lut);
                T Y = CmidM + (eM + tMFY);
                Y = logadd_lut(Y, CupY + (eY + tGEY), lut);
                if (useSWX) Y = logadd_lut(Y, ClowX + (eX + tSWX), lut);
                Q.M[c] = valid ? M : NI;
                Q.X[c] = valid ? X : NI;
                Q.Y[c] = valid ? Y : NI;
                if (renorm) {
                    T m3 = Q.M[c] > Q.X[c] ? Q.M[c] : Q.X[c];
                    m3 = m3 > Q.Y[c] ? m3 : Q.Y[c];
                    mx = m3 > mx ? m3 : mx;
                }
            }
            have_eY = true;
            obe = ob1;
            if (renorm) {
                /* every 8th diagonal: the largest state becomes 0, its log carried in the
                   diagonal's offset; e+1's match states (e+2's on the next diagonal) move
                   with them, so that the two diagonals a cell reads always share one offset */
                mx = wave_max(mx);
                if (mx > NI) {
#pragma unroll
                    for (int c = 0; c < CPT; c++) {
                        Q.M[c] -= mx; Q.X[c] -= mx; Q.Y[c] -= mx;
                        P.M[c] -= mx;
                    }
                    LowM -= mx;
                    obe = ob1 + (double)mx;
                }
            }
        }
        {
            const int jn = (e - 1) - (X0 + W - 1); /* event of lane 63's last cell on e-1 */
            load_econ_uniform(pre, a, evc, jn, jn >= 0 && jn < lY);
        }
        if constexpr (PF) {
            /* the events of e-1 are this diagonal's shifted one lane left */
            const ECon<T, EK> in = shl1(ev[0], pre);
#pragma unroll
            for (int c = 0; c < CPT; c++) {
                const T evm = c < CPT - 1 ? ev[c < CPT - 1 ? c + 1 : 0].m : in.m;
                gp[c] = hdp_fetch(m, kc[c].row, kc[c].hmu, evm, s);
            }
            have_gp = true;
        }
        if constexpr (PL) {
            /* E(e, x): the extra-event emissions of e-1's cells */
#pragma unroll
            for (int c = 0; c < CPT; c++) {
                const int x = X0 + CPT * lane + c;
                pe[c] = x >= lo && x <= hi ? pl(em, dof + (x - lo)) : (T)0;
            }
            have_pe = true;
        }

        if (e <= TF) {
            const double offe = NORM ? fo + obe : 0.0;
            if (posterior_count % 10 == 0) {
                /* diagonalCalculationTotalProbability (impl/pairwiseAligner.c:1182-1200) */
                if (!have_rpre) load_refresh(e);
                have_rpre = false;
                double tA[CPT], tB[CPT];
#pragma unroll
                for (int c = 0; c < CPT; c++) {
                    const int x = X0 + CPT * lane + c;
                    tA[c] = -__builtin_huge_val();
                    tB[c] = -__builtin_huge_val();
                    if (x >= lo && x <= hi) {
                        T cell = (have_fpre ? fpre[c] : pl(fM, dof + (x - lo))) + Q.M[c];
                        cell = logadd_lut(cell, rfx[c] + Q.X[c], lut);
                        cell = logadd_lut(cell, rfy[c] + Q.Y[c], lut);
                        tA[c] = (double)cell + offe;
                    }
                }
                const bool h1 = e + 1 <= D;
                if (h1) {
                    /* match flux: B[e+1] against match transitions out of F[e-1] (a cell
                       whose middle lies outside F[e-1] contributes -inf, as in the reference) */
                    const int lo1e = ld_scalar(xlo, e + 1), hi1e = ld_scalar(xhi, e + 1);
                    const double offm = NORM ? foff[e - 1] + obe : 0.0; /* B[e+1] is at e's offset */
#pragma unroll
                    for (int c = 0; c < CPT; c++) {
                        const int x = X0 + CPT * lane + c; /* cell (x, e+1-x); middle (x-1, e-x) on e-1 */
                        if (x >= lo1e && x <= hi1e) {
                            /* the match emission of k-mer x, event e-x */
                            T eM;
                            if constexpr (RE) {
                                eM = eYp[c];
                            } else {
                                T unused;
                                em_both<T>(m, kc[c], ev[c], s, iv, eM, unused);
                            }
                            T Mv = rgm[c] + (eM + tMC);
                            Mv = logadd_lut(Mv, rgx[c] + (eM + tMFX), lut);
                            Mv = logadd_lut(Mv, rgy[c] + (eM + tMFY), lut);
                            tB[c] = (double)(Mv + P.M[c]) + offm;
                        }
                    }
                }
                double tot = chain_reg<CPT>(tA, lutd);
                if (h1) tot = logadd(tot, chain_reg<CPT>(tB, lutd));
                total = tot;
                if (e == TF && lane == 0) a.win_total[wi] = total;
            }
            posterior_count++;
            /* diagonalCalculationPosteriorMatchProbs (impl/pairwiseAligner.c:1202-1253): a
               cell is kept when exp(F + B - total) >= threshold.  The exponent is first
               compared with log(threshold) less a margin far above the rounding of exp;
               only diagonals with a cell past it take the exp and the exact comparison. */
            const T cst = NORM ? (T)(offe - total) : (T)0;
            T sv[CPT];
            bool cand[CPT];
            bool anyc = false;
            if (have_fpre && lo > 0 && hi < e) {
                /* no cell of e lies in row 0 or column 0, and the prefetched forward
                   states are NaN outside the band: the exponent is NaN there and fails
                   the test, so no per-cell bounds check is needed */
#pragma unroll
                for (int c = 0; c < CPT; c++) {
                    sv[c] = NORM ? (fpre[c] + Q.M[c]) + cst : (fpre[c] + Q.M[c]) - (T)total;
                    cand[c] = sv[c] >= lthr;
                    anyc |= cand[c];
                }
            } else {
#pragma unroll
                for (int c = 0; c < CPT; c++) {
                    const int x = X0 + CPT * lane + c;
                    const int y = e - x;
                    cand[c] = false;
                    sv[c] = NI;
                    if (x >= lo && x <= hi && x > 0 && y > 0) {
                        const T fm = have_fpre ? fpre[c] : pl(fM, dof + (x - lo));
                        sv[c] = NORM ? (fm + Q.M[c]) + cst : (fm + Q.M[c]) - (T)total;
                        cand[c] = sv[c] >= lthr;
                    }
                    anyc |= cand[c];
                }
            }
            if (__ballot(anyc)) {
                bool keep[CPT];
                T pv[CPT];
                unsigned long long bal[CPT];
#pragma unroll
                for (int c = 0; c < CPT; c++) {
                    pv[c] = fexp(sv[c]);
                    keep[c] = cand[c] && pv[c] >= thr;
                    bal[c] = __ballot(keep[c]);
                }
                const unsigned long long lt = (1ull << lane) - 1ull;
                int base = cursor;
                int ntot = 0;
#pragma unroll
                for (int c = 0; c < CPT; c++) {
                    base += __popcll(bal[c] & lt);
                    ntot += __popcll(bal[c]);
                }
                int own = 0;
#pragma unroll
                for (int c = 0; c < CPT; c++) {
                    if (keep[c]) {
                        const int idx = base + own;
                        if (idx < Wd.out_cap) {
                            const int x = X0 + CPT * lane + c;
                            const double p = pv[c] > (T)1 ? 1.0 : (double)pv[c];
                            out[3 * idx + 0] = (int32_t)floor(p * (double)kPairProb1);
                            out[3 * idx + 1] = x - 1;
                            out[3 * idx + 2] = e - x - 1;
                        }
                        own++;
                    }
                }
                cursor += ntot;
            }
        }
        if (e - 1 <= TF && e - 1 > TB) {
#pragma unroll
            for (int c = 0; c < CPT; c++) {
                const int x = X0 + CPT * lane + c;
                const bool v = x >= nlo && x <= nhi;
                fpre[c] = v ? pl(fM, ndo + (x - nlo)) : __builtin_nanf(""); /* NaN: see the posterior test */
            }
            have_fpre = true;
            /* e-1 refreshes the total (FP64: its 5·CPT states are loaded there instead; held
               across the diagonal they cost spills) */
            if (sizeof(T) == 4 && posterior_count % 10 == 0) {
                load_refresh(e - 1);
                have_rpre = true;
            }
        }
        RpM = LowM;
        ob1 = obe;
        hi2 = hi1;
        hi1 = hi;
        e--;
        return true;
    };
    FSt A, B; /* the states of e+1 and e+2 */
#pragma unroll
    for (int c = 0; c < CPT; c++) {
        A.M[c] = pM[c]; A.X[c] = pX[c]; A.Y[c] = pY[c];
        B.M[c] = qM[c]; B.X[c] = NI; B.Y[c] = NI;
    }
    bool swapped = false; /* e+1's states are in B */
    if constexpr (sizeof(T) == 4) {
        while (e > TB) {
            if (!step(A, B)) break;
            swapped = true;
            if (e <= TB || !step(B, A)) break;
            swapped = false;
        }
    } else {
        /* FP64: two full sets of states cost the sweep more spills than the copies */
        while (e > TB) {
            if (!step(A, B)) break;
#pragma unroll
            for (int c = 0; c < CPT; c++) {
                const T q = A.M[c];
                A.M[c] = B.M[c]; A.X[c] = B.X[c]; A.Y[c] = B.Y[c];
                B.M[c] = q;
            }
        }
    }
    if (next) {
        const FSt &P = swapped ? B : A, &Q = swapped ? A : B;
#pragma unroll
        for (int c = 0; c < CPT; c++) {
            const int i = CPT * lane + c;
            xs[0 * xsw + i] = P.M[c]; xs[1 * xsw + i] = P.X[c]; xs[2 * xsw + i] = P.Y[c]; xs[3 * xsw + i] = Q.M[c];
        }
        __syncthreads();
    }
    st.e = e;
    st.X0 = X0;
    st.wprev = W;
    st.ncls = ncls;
    st.ob1 = ob1;
    st.hi1 = hi1; st.hi2 = hi2;
    st.cursor = cursor; st.posterior_count = posterior_count;
    st.total = total;
    st.bad = bad;
    return next;
}

/* One wavefront per traceback window; the window runs at 1, 2 or 3 cells per lane as
 * the band's width allows (block classes from prep_kernel), never wider than CMAX. */
/* Waves per SIMD the backward sweep is compiled for: the FP64 sweeps at three cells
   per lane would otherwise hold one wave per SIMD; at two they spill a few registers
   to scratch and still run 10% faster.  The FP32 ones at three cells are held to 168
   registers, three waves per SIMD.  Otherwise the compiler's choice. */
template <typename T, int CMAX> struct BwdWaves {
    static constexpr int n = CMAX != 3 ? 1 : sizeof(T) == 8 ? 2 : 3;
};

/* One traceback window (the kernel below and the persistent sweeps' window item). */
template <typename T, int CMAX, int EK>
__device__ __forceinline__ void bwd_window(const Args<T> &a, const int wi, T *xs, const LutEntry<T> *lut,
                                           const LutEntry<double> *lutd) {
    const WinDev Wd = a.wins[wi];
    const TaskDev t = a.tasks[Wd.task];
    const int lane = threadIdx.x;
    const int8_t *dcl = a.dcls + t.diag_off;
    BwdRun st;
    st.e = Wd.D;
    st.X0 = 0; st.wprev = 0;
    const int b0 = Wd.D >> 6;
    st.ncls = 64 * b0 - 1 > Wd.T ? ld_scalar_byte(dcl, 64 * (b0 - 1)) : 0;
    st.ob1 = 0.0;
    st.hi1 = -1; st.hi2 = -1;
    st.cursor = 0; st.posterior_count = 0;
    st.total = -__builtin_huge_val();
    st.bad = false;
    int cls = window_class((ld_scalar_byte(dcl, 64 * b0) >> 3) & 7, CMAX);
    while (cls) {
        if constexpr (CMAX == 1 || CMAX > 3) {
            cls = bwd_segment<T, CMAX, EK>(a, wi, Wd, t, st, xs, 64 * CMAX, CMAX, lut, lutd);
        } else {
            if (cls == 1) cls = bwd_segment<T, 1, EK>(a, wi, Wd, t, st, xs, 64 * CMAX, CMAX, lut, lutd);
            else if (CMAX == 3 && cls == 2) cls = bwd_segment<T, 2, EK>(a, wi, Wd, t, st, xs, 64 * CMAX, CMAX, lut, lutd);
            else cls = bwd_segment<T, CMAX, EK>(a, wi, Wd, t, st, xs, 64 * CMAX, CMAX, lut, lutd);
        }
    }
    if (lane == 0) a.win_count[wi] = st.cursor;
    if (st.bad && lane == 0) atomicOr(a.err, ERR_WINDOWS);
}

template <typename T, int CMAX, int EK>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(BwdWaves<T, CMAX>::n)))
backward_reg_kernel(Args<T> a, const int32_t *__restrict__ order) {
    extern __shared__ __align__(16) unsigned char dsm[];
    T *xs = reinterpret_cast<T *>(dsm); /* [4][64 * CMAX]: states across a width change */
    __shared__ LutEntry<T> lut[LutSize<T>::n];
    __shared__ LutEntry<double> lutd2[sizeof(T) == 8 ? 1 : 16]; /* FP64 table for the totals */
    lut_fill(lut, (int)threadIdx.x);
    if constexpr (sizeof(T) != 8) lut_fill(lutd2, (int)threadIdx.x);
    __syncthreads();
    const LutEntry<double> *lutd = sizeof(T) == 8 ? reinterpret_cast<const LutEntry<double> *>(lut) : lutd2;
    bwd_window<T, CMAX, EK>(a, order[blockIdx.x], xs, lut, lutd);
}

/* ------------------------------------------------------------------------ */
/* persistent sweeps: forward items and their traceback windows, one launch  */
/* ------------------------------------------------------------------------ */

/* Worker wavefronts claim forward items in the launch's order (longest first).  A
 * worker whose sweep is done makes the sub-alignment's planes and per-diagonal offsets
 * visible device-wide (agent-scope release) and publishes it in the next slot of
 * `done`.  Once every forward item is claimed, workers take the traceback windows of
 * the published sub-alignments in publication order (an agent-scope acquire before
 * each), each window once.  A worker waits only for a sub-alignment whose forward
 * sweep another running worker has claimed, and forward sweeps never wait, so every
 * wait ends; a worker that makes no progress for wait_ticks of the real-time clock
 * (seconds; a step takes milliseconds) raises ERR_QUEUE and leaves, so the launch ends
 * whatever the queue holds.  The sub-alignments' per-diagonal arrays
 * start on 128-byte boundaries (sa_batch_create), so no scalar-cache line a window's
 * sweep reads holds another sub-alignment's offsets. */
/* registers for three workers per SIMD in FP32, as the separate sweeps get (more spill to scratch),
   two in FP64 (as the backward sweep at three positions per lane) */
template <typename T> struct PersistWaves { static constexpr int n = sizeof(T) == 8 ? 2 : 3; };

template <typename T, int EK>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(PersistWaves<T>::n)))
fb_persistent_kernel(Args<T> a, PersistQ q) {
    constexpr int CMAX = 3;
    extern __shared__ __align__(16) unsigned char dsm[];
    T *xs = reinterpret_cast<T *>(dsm); /* [6][64 * CMAX] forward, [4][64 * CMAX] backward */
    __shared__ LutEntry<T> lut[LutSize<T>::n];
    __shared__ LutEntry<double> lutd2[sizeof(T) == 8 ? 1 : 16]; /* FP64 table for the totals */
    lut_fill(lut, (int)threadIdx.x);
    if constexpr (sizeof(T) != 8) lut_fill(lutd2, (int)threadIdx.x);
    __syncthreads();
    const LutEntry<double> *lutd = sizeof(T) == 8 ? reinterpret_cast<const LutEntry<double> *>(lut) : lutd2;
    const int lane = threadIdx.x;
    /* forward items: their chains are the launch's longest, so they issue first on a
       shared SIMD */
    __builtin_amdgcn_s_setprio(1);
    while (true) {
        int f = 0;
        if (lane == 0) f = atomicAdd(&q.ctr[0], 1);
        f = __builtin_amdgcn_readfirstlane(f);
        if (f >= q.n_fwd) break;
        const int ti = q.order[f];
        fwd_task<T, CMAX, EK>(a, a.tasks[ti], xs, lut);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        if (lane == 0) {
            const int slot = atomicAdd(&q.ctr[1], 1);
            atomicExch(&q.done[slot], ti + 1); /* a read-modify-write, as every queue access */
        }
    }
    __builtin_amdgcn_s_setprio(0);
    /* windows of the published sub-alignments.  Every access to the queue, reads
       included, is an atomic read-modify-write (an OR with a zero the compiler cannot
       see, so that no read becomes a plain load): all of them meet at the one point
       where read-modify-writes are ordered.  The cursor moves by the compare-and-swap's
       own result.  A worker that makes no progress for q.wait_ticks of the 100 MHz
       real-time clock raises ERR_QUEUE and leaves, however long its naps are. */
    const int zero = q.zero;
    uint64_t last = __builtin_amdgcn_s_memrealtime();
    int k = 0;
    if (lane == 0) k = atomicOr(&q.ctr[2], zero);
    k = __builtin_amdgcn_readfirstlane(k);
    int naps = 1;
    while (k < q.n_fwd) {
        int t1 = 0;
        if (lane == 0) t1 = atomicOr(&q.done[k], zero);
        t1 = __builtin_amdgcn_readfirstlane(t1);
        if (t1 != 0) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            const int ti = t1 - 1;
            int w = 0;
            if (lane == 0) w = atomicAdd(&q.wclaim[ti], 1);
            w = __builtin_amdgcn_readfirstlane(w);
            if (w < a.task_nwin[ti]) {
                bwd_window<T, CMAX, EK>(a, q.first_win[ti] + w, xs, lut, lutd);
                last = __builtin_amdgcn_s_memrealtime();
                naps = 1;
                continue; /* the same publication slot may hold more windows */
            }
            /* this sub-alignment's windows are all taken: move the cursor past it (or
               follow whoever did) */
            int old = 0;
            if (lane == 0) old = atomicCAS(&q.ctr[2], k, k + 1);
            old = __builtin_amdgcn_readfirstlane(old);
            k = old == k ? k + 1 : old;
            last = __builtin_amdgcn_s_memrealtime();
        } else {
            /* the k-th forward sweep is still running: back off so that the waiting
               workers do not crowd one address */
            for (int z = 0; z < naps; z++) __builtin_amdgcn_s_sleep(127);
            naps = naps < 8 ? naps + 1 : 8;
            if (lane == 0) k = atomicOr(&q.ctr[2], zero);
            k = __builtin_amdgcn_readfirstlane(k);
            if (__builtin_amdgcn_s_memrealtime() - last > (uint64_t)q.wait_ticks) {
                if (lane == 0) atomicOr(a.err, ERR_QUEUE);
                break;
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* HDP emission plane                                                        */
/* ------------------------------------------------------------------------ */

/* The density emission of every band cell of the listed sub-alignments, in the
 * layout of the forward match plane (cell x of diagonal d at cell_off + doff[d] +
 * x - xlo[d]).  Cell x on d: k-mer x-1 against event d-x-1, the value the sweeps'
 * own gathers give (hdp_fetch/hdp_finish: emissions_signal_getHdpKmerDensity,
 * impl/stateMachine.c:468-495; the match and extra-event emissions are the same
 * density); -inf where x has no k-mer, an absent event reads mean 0, as in the
 * sweeps.  The register sweeps then read a cell's emission with one coalesced load
 * a diagonal ahead instead of gathering from the density grid on their dependency
 * chain, and the backward pass does not gather again.
 *
 * Gathered cell by cell, the grid (3125 rows of 2000 knots: 50 MB in FP32) does not
 * stay in L2 and every cell costs a cache line from beyond it.  So the plane is
 * built in two passes.  The column pass takes the band's columns grouped by density
 * row (the host sorts them): a workgroup stages one row (16 KB) and the knots in LDS
 * and computes whole columns x (diagonals col_a[x] .. col_b[x], consecutive events)
 * into the column plane.  The transpose pass then moves each chunk of kPlaneDiags
 * diagonals from the column plane to the diagonal layout through an LDS tile. */
template <typename T>
__global__ void __launch_bounds__(256) hdp_columns_kernel(Args<T> a, const int32_t *__restrict__ ent,
                                                          const int32_t *__restrict__ groups) {
    constexpr int NG = 2048, NQ = kColumnGroup;
    __shared__ T s_ys[2 * NG], s_gx[NG];
    /* per column of the group: where its cells go, its first diagonal, its cells'
       first flat index, and what its emissions need */
    __shared__ int64_t s_dst[NQ];
    __shared__ int32_t s_ca[NQ], s_x[NQ], s_f[NQ + 1], s_lY[NQ];
    __shared__ int64_t s_evc[NQ];
    __shared__ T s_hmu[NQ], s_s[NQ], s_iv[NQ];
    const Model<T> &m = a.m;
    const int row = groups[3 * blockIdx.x], e0 = groups[3 * blockIdx.x + 1], ne = groups[3 * blockIdx.x + 2];
    const int tid = threadIdx.x;
    const int G = m.grid_length;
    const bool lg = G <= NG;
    if (row >= 0 && lg) {
        const T *yr = m.ys + (int64_t)row * G * 2;
        for (int i = tid; i < 2 * G; i += 256) s_ys[i] = yr[i];
        for (int i = tid; i < G; i += 256) s_gx[i] = m.gridx[i];
    }
    int h = 0;
    if (tid < NQ) {
        if (tid < ne) {
            const int ti = ent[2 * (e0 + tid)], x = ent[2 * (e0 + tid) + 1];
            const TaskDev t = a.tasks[ti];
            const int64_t cs = t.pos_off + ti + x;
            const int ca = a.col_a[cs], cb = a.col_b[cs], co = a.col_off[cs];
            /* (a band whose bounds are not non-decreasing fails the sweeps, which then
               run again on the LDS kernels; its columns are only kept inside its cells) */
            h = (int)max<int64_t>(0, min<int64_t>((int64_t)cb - ca + 1, a.task_cells[ti] - co));
            s_dst[tid] = t.cell_off + co;
            s_ca[tid] = ca;
            s_x[tid] = x;
            s_lY[tid] = t.lY;
            s_evc[tid] = t.evc_off;
            s_s[tid] = (T)(t.scale / t.var);
            s_iv[tid] = (T)(1.0 / t.var);
            s_hmu[tid] = row >= 0 ? m.hmu[a.kid[t.pos_off + x - 1]] : (T)0;
        }
        /* exclusive prefix of the column heights (the group fits one wavefront) */
        int v = h;
        for (int o = 1; o < 64; o <<= 1) {
            const int u = __shfl_up(v, o, 64);
            if (tid >= o) v += u;
        }
        s_f[tid + 1] = v;
        if (tid == 0) s_f[0] = 0;
    }
    __syncthreads();
    /* one column per wavefront at a time, two columns in flight; the knots come from
       LDS (ds_read) when they fit, else from HBM */
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    auto run = [&](const T *yrow, const T *gx) {
        for (int q0 = wave; q0 < ne; q0 += 8) {
#pragma unroll
            for (int off = 0; off < 128; off += 64) {
                T e[2];
                int d[2];
#pragma unroll
                for (int u = 0; u < 2; u++) {
                    const int q = q0 + 4 * u;
                    const int h = q < ne ? s_f[q + 1] - s_f[q] : 0;
                    const int i = off + lane;
                    d[u] = q < ne ? s_ca[q] + i : 0;
                    const int j = q < ne ? d[u] - s_x[q] - 1 : -1;
                    e[u] = i < h && row >= 0 && j >= 0 && j < s_lY[q] ? a.evm[s_evc[q] + j] : (T)0;
                }
#pragma unroll
                for (int u = 0; u < 2; u++) {
                    const int q = q0 + 4 * u;
                    const int h = q < ne ? s_f[q + 1] - s_f[q] : 0;
                    if (off + lane >= h) continue;
                    const T v = row < 0 ? ninf<T>()
                                        : hdp_finish(m, hdp_fetch_r(m, row, yrow, s_hmu[q], e[u], s_s[q], gx), s_iv[q]);
                    a.emc[s_dst[q] + (d[u] - s_ca[q])] = v;
                }
            }
            /* columns taller than 128 cells: the rest one by one */
#pragma unroll
            for (int u = 0; u < 2; u++) {
                const int q = q0 + 4 * u;
                if (q >= ne) continue;
                const int h = s_f[q + 1] - s_f[q];
                for (int i = 128 + lane; i < h; i += 64) {
                    const int dd = s_ca[q] + i, j = dd - s_x[q] - 1;
                    const T e = row >= 0 && j >= 0 && j < s_lY[q] ? a.evm[s_evc[q] + j] : (T)0;
                    const T v = row < 0 ? ninf<T>()
                                        : hdp_finish(m, hdp_fetch_r(m, row, yrow, s_hmu[q], e, s_s[q], gx), s_iv[q]);
                    a.emc[s_dst[q] + i] = v;
                }
            }
        }
    };
    if (lg) run(s_ys, s_gx);
    else run(m.ys + (int64_t)(row < 0 ? 0 : row) * G * 2, m.gridx);
}

/* Chunks of kPlaneDiags diagonals, grid-stride (a chunk is a few thousand cells:
   one workgroup per chunk would spend its time being launched) */
template <typename T>
__global__ void __launch_bounds__(256) hdp_plane_kernel(Args<T> a, const int32_t *__restrict__ ctask,
                                                        const int32_t *__restrict__ chunk0, int c0, int c1) {
    constexpr int NP = 192, TS = kPlaneDiags + 1; /* tile: NP columns x kPlaneDiags diagonals (+1: banks) */
    __shared__ int32_t s_off[kPlaneDiags + 1], s_lo[kPlaneDiags], s_hi[kPlaneDiags];
    __shared__ int32_t s_ca[NP], s_cb[NP], s_co[NP];
    __shared__ T tile[NP * TS];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    for (int b = c0 + blockIdx.x; b < c1; b += gridDim.x) {
        const int ti = ctask[b]; /* the chunk's sub-alignment, its first chunk */
        const TaskDev t = a.tasks[ti];
        const int N = t.lX + t.lY;
        const int d0 = (b - chunk0[ti]) * kPlaneDiags;
        const int nd = min(kPlaneDiags, N + 1 - d0);
        const int32_t *xlo = a.xlo + t.diag_off;
        const int32_t *xhi = a.xhi + t.diag_off;
        const int32_t *doff = a.doff + t.diag_off;
        const int32_t *ca = a.col_a + t.pos_off + ti, *cb = a.col_b + t.pos_off + ti;
        const int32_t *co = a.col_off + t.pos_off + ti;
        T *em = a.em + t.cell_off;
        const T *emc = a.emc + t.cell_off;
        const int64_t cells = a.task_cells[ti];
        __syncthreads(); /* the previous chunk is done with the shared arrays */
        if (tid < nd) {
            const int d = d0 + tid;
            s_lo[tid] = xlo[d];
            s_hi[tid] = xhi[d];
            s_off[tid] = doff[d];
        }
        __syncthreads();
        int pmin = s_lo[0], pmax = s_hi[0]; /* the chunk's columns */
        for (int k = 1; k < nd; k++) {
            pmin = min(pmin, s_lo[k]);
            pmax = max(pmax, s_hi[k]);
        }
        const int np = pmax - pmin + 1;
        const bool lt = np <= NP;
        if (lt) {
            for (int i = tid; i < np; i += 256) {
                s_ca[i] = ca[pmin + i];
                s_cb[i] = cb[pmin + i];
                s_co[i] = co[pmin + i];
            }
            __syncthreads();
            /* each column's segment of the chunk's diagonals: consecutive values */
            for (int i = tid; i < np * kPlaneDiags; i += 256) {
                const int c = i / kPlaneDiags, k = i % kPlaneDiags, d = d0 + k;
                const int64_t ic = (int64_t)s_co[c] + (d - s_ca[c]);
                if (k < nd && d >= s_ca[c] && d <= s_cb[c] && ic < cells) tile[c * TS + k] = emc[ic];
            }
            __syncthreads();
        }
        /* out along the diagonals, one wavefront per diagonal */
        for (int k = wave; k < nd; k += 4) {
            const int l = s_lo[k], hh = s_hi[k], d = d0 + k;
            T *row = em + s_off[k] - l;
            for (int x = l + lane; x <= hh; x += 64) {
                T v;
                if (lt) v = tile[(x - pmin) * TS + k];
                else {
                    const int64_t ic = (int64_t)co[x] + (d - ca[x]);
                    v = ic >= 0 && ic < cells ? emc[ic] : (T)0;
                }
                row[x] = v;
            }
        }
    }
}

template <typename T>
hipError_t launch_hdp_plane(const Args<T> &a, const int32_t *entries, const int32_t *groups, int n_groups,
                            const int32_t *ctask, const int32_t *chunk0, int c0, int c1, int n_cu, hipStream_t st) {
    n_cu = n_cu > 0 ? n_cu : 256;
    if (c1 <= c0) return hipSuccess;
    if (n_groups > 0) hipLaunchKernelGGL(hdp_columns_kernel<T>, dim3(n_groups), dim3(256), 0, st, a, entries, groups);
    const int grid = c1 - c0 < 8 * n_cu ? c1 - c0 : 8 * n_cu;
    hipLaunchKernelGGL(hdp_plane_kernel<T>, dim3(grid), dim3(256), 0, st, a, ctask, chunk0, c0, c1);
    return hipGetLastError();
}
template hipError_t launch_hdp_plane<float>(const Args<float> &, const int32_t *, const int32_t *, int,
                                            const int32_t *, const int32_t *, int, int, int, hipStream_t);
template hipError_t launch_hdp_plane<double>(const Args<double> &, const int32_t *, const int32_t *, int,
                                             const int32_t *, const int32_t *, int, int, int, hipStream_t);

/* ------------------------------------------------------------------------ */
/* launchers                                                                 */
/* ------------------------------------------------------------------------ */

template <typename T, int CPT, int EK>
static hipError_t fwd_reg_inst(const Args<T> &a, const int32_t *order, int n, int nw, size_t shmem, hipStream_t st) {
    if constexpr (CPT == 1) {
        if (nw == 2) {
            hipLaunchKernelGGL((forward_reg_kernel<T, CPT, EK, 2>), dim3(n), dim3(128), shmem, st, a, order);
            return hipGetLastError();
        }
        if (nw == 3) {
            hipLaunchKernelGGL((forward_reg_kernel<T, CPT, EK, 3>), dim3(n), dim3(192), shmem, st, a, order);
            return hipGetLastError();
        }
    }
    if constexpr (CPT <= 2) {
        if (nw == 4) {
            hipLaunchKernelGGL((forward_reg_kernel<T, CPT, EK, 4>), dim3(n), dim3(256), shmem, st, a, order);
            return hipGetLastError();
        }
    }
    const size_t xs = sizeof(T) * 6 * 64 * CPT; /* states across a width change */
    hipLaunchKernelGGL((forward_sw_kernel<T, CPT, EK>), dim3(n), dim3(64), shmem > xs ? shmem : xs, st, a, order);
    return hipGetLastError();
}

template <typename T, int CPT>
static hipError_t fwd_reg_ek(const Args<T> &a, const int32_t *order, int n, int nw, int ek, size_t shmem,
                             hipStream_t st) {
    switch (ek) {
    case 0: return fwd_reg_inst<T, CPT, 0>(a, order, n, nw, shmem, st);
    case 1: return fwd_reg_inst<T, CPT, 1>(a, order, n, nw, shmem, st);
    case 2: return fwd_reg_inst<T, CPT, 2>(a, order, n, nw, shmem, st);
    default: return fwd_reg_inst<T, CPT, 3>(a, order, n, nw, shmem, st);
    }
}

template <typename T, int CPT>
static hipError_t bwd_reg_ek(const Args<T> &a, const int32_t *order, int n, int ek, size_t shmem,
                             hipStream_t st) {
    const size_t xs = sizeof(T) * 4 * 64 * CPT; /* states across a width change */
    shmem = shmem > xs ? shmem : xs;
    switch (ek) {
    case 0: hipLaunchKernelGGL((backward_reg_kernel<T, CPT, 0>), dim3(n), dim3(64), shmem, st, a, order); break;
    case 1: hipLaunchKernelGGL((backward_reg_kernel<T, CPT, 1>), dim3(n), dim3(64), shmem, st, a, order); break;
    case 2: hipLaunchKernelGGL((backward_reg_kernel<T, CPT, 2>), dim3(n), dim3(64), shmem, st, a, order); break;
    default: hipLaunchKernelGGL((backward_reg_kernel<T, CPT, 3>), dim3(n), dim3(64), shmem, st, a, order); break;
    }
    return hipGetLastError();
}

template <typename T>
hipError_t launch_forward_reg(const Args<T> &a, const int32_t *order, int n, int cpt, int nw, int ek,
                              size_t shmem, hipStream_t st) {
    if (n == 0) return hipSuccess;
    if (nw < 1 || nw > 4) return hipErrorInvalidValue;
    if ((nw == 2 || nw == 3) && cpt != 1) return hipErrorInvalidValue;
    if (nw == 4 && cpt > 2) return hipErrorInvalidValue;
    switch (cpt) {
    case 1: return fwd_reg_ek<T, 1>(a, order, n, nw, ek, shmem, st);
    case 2: return fwd_reg_ek<T, 2>(a, order, n, nw, ek, shmem, st);
    case 3: return fwd_reg_ek<T, 3>(a, order, n, 1, ek, shmem, st);
    case 4: return fwd_reg_ek<T, 4>(a, order, n, 1, ek, shmem, st);
    case 6: return fwd_reg_ek<T, 6>(a, order, n, 1, ek, shmem, st);
    case 8: return fwd_reg_ek<T, 8>(a, order, n, 1, ek, shmem, st);
    default: return hipErrorInvalidValue;
    }
}

template <typename T, int EK>
static hipError_t fwd_halo_inst(const Args<T> &a, const int32_t *order, int n, int nw, hipStream_t st) {
    switch (nw) {
    case 2: hipLaunchKernelGGL((forward_halo_kernel<T, EK, 2>), dim3(n), dim3(128), 0, st, a, order); break;
    case 3: hipLaunchKernelGGL((forward_halo_kernel<T, EK, 3>), dim3(n), dim3(192), 0, st, a, order); break;
    default: hipLaunchKernelGGL((forward_halo_kernel<T, EK, 4>), dim3(n), dim3(256), 0, st, a, order); break;
    }
    return hipGetLastError();
}

template <typename T>
hipError_t launch_forward_halo(const Args<T> &a, const int32_t *order, int n, int nw, int ek, hipStream_t st) {
    if (n == 0) return hipSuccess;
    if (nw < 2 || nw > 4) return hipErrorInvalidValue;
    switch (ek) {
    case 0: return fwd_halo_inst<T, 0>(a, order, n, nw, st);
    case 1: return fwd_halo_inst<T, 1>(a, order, n, nw, st);
    case 2: return fwd_halo_inst<T, 2>(a, order, n, nw, st);
    default: return fwd_halo_inst<T, 3>(a, order, n, nw, st);
    }
}

template <typename T>
hipError_t launch_fb_persistent(const Args<T> &a, const PersistQ &q, int workers, int ek, hipStream_t st) {
    if (q.n_fwd == 0 || workers <= 0) return hipSuccess;
    const size_t shmem = sizeof(T) * 6 * 64 * 3;
    switch (ek) {
    case 0: hipLaunchKernelGGL((fb_persistent_kernel<T, 0>), dim3(workers), dim3(64), shmem, st, a, q); break;
    case 1: hipLaunchKernelGGL((fb_persistent_kernel<T, 1>), dim3(workers), dim3(64), shmem, st, a, q); break;
    case 2: hipLaunchKernelGGL((fb_persistent_kernel<T, 2>), dim3(workers), dim3(64), shmem, st, a, q); break;
    default: hipLaunchKernelGGL((fb_persistent_kernel<T, 3>), dim3(workers), dim3(64), shmem, st, a, q); break;
    }
    return hipGetLastError();
}

template <typename T>
hipError_t launch_backward_reg(const Args<T> &a, const int32_t *order, int n, int cpt, int ek, size_t shmem,
                               hipStream_t st) {
    if (n == 0) return hipSuccess;
    switch (cpt) {
    case 1: return bwd_reg_ek<T, 1>(a, order, n, ek, shmem, st);
    case 2: return bwd_reg_ek<T, 2>(a, order, n, ek, shmem, st);
    case 3: return bwd_reg_ek<T, 3>(a, order, n, ek, shmem, st);
    case 4: return bwd_reg_ek<T, 4>(a, order, n, ek, shmem, st);
    case 6: return bwd_reg_ek<T, 6>(a, order, n, ek, shmem, st);
    case 8: return bwd_reg_ek<T, 8>(a, order, n, ek, shmem, st);
    default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_forward_reg<float>(const Args<float> &, const int32_t *, int, int, int, int, size_t,
                                              hipStream_t);
template hipError_t launch_forward_reg<double>(const Args<double> &, const int32_t *, int, int, int, int, size_t,
                                               hipStream_t);
template hipError_t launch_fb_persistent<float>(const Args<float> &, const PersistQ &, int, int, hipStream_t);
template hipError_t launch_fb_persistent<double>(const Args<double> &, const PersistQ &, int, int, hipStream_t);
template hipError_t launch_forward_halo<float>(const Args<float> &, const int32_t *, int, int, int, hipStream_t);
template hipError_t launch_forward_halo<double>(const Args<double> &, const int32_t *, int, int, int, hipStream_t);
template hipError_t launch_backward_reg<float>(const Args<float> &, const int32_t *, int, int, int, size_t,
                                              hipStream_t);
template hipError_t launch_backward_reg<double>(const Args<double> &, const int32_t *, int, int, int, size_t,
                                               hipStream_t);

} // namespace sa
