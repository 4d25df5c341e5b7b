/*
 * sa_paths.hip -- the banded forward-backward over degenerate references ('X'
 * bases, multi-path cells; impl/pairwiseAligner.c:501-692), FP64.
 *
 * Column x of a sub-alignment (k-mer x-1; column 0 has no k-mer) holds one path
 * per combination of the degenerate base options at the 'X' positions of its
 * k-mer (hdCell_construct, :599-656), numbered with the first 'X' varying slowest
 * (path_listPotentialKmers, :569-585).  A cell keeps three states per path.  A
 * match or gapX transition joins path q of column x-1 and path p of column x when
 * their k-mers agree on the overlap (path_checkLegal, :540-552): the shared 'X'
 * sites take the same option.  GapY transitions stay on a path.
 *
 * The work unit is a path entry (cell, path): entries of a column are consecutive,
 * columns in order, so a diagonal's entries are one contiguous range and the
 * rings and forward planes are indexed by entry.  Recurrence order per state is the
 * reference's: stateMachine3_cellCalculate visits the current cell's paths and,
 * for each, the neighbour's paths in order (push form forward; the pull form
 * backward takes middle, then upper, then lower contributions, each over the
 * neighbour's paths in order).  Every diagonal's three states go to the forward
 * planes.
 */
#include "sa_device.h"

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace sa {

namespace {

__device__ __forceinline__ int ipow(int b, int e) {
    int r = 1;
    for (int i = 0; i < e; i++) r *= b;
    return r;
}

/* The legal predecessors of path p of column x (x >= 1) in column x-1, in
 * increasing order: path_checkLegal (impl/pairwiseAligner.c:540-552) accepts q
 * when the 'X' sites the two k-mers share take the same options.  Column x-1
 * holds sites [sa1, sb1), column x sites [sa0, sb0), sa1 <= sa0 <= sb1 <= sb0; a
 * path's digits are its sites' options, the first site most significant.  The
 * shared sites [sa0, sb1) are p's leading digits, so the legal q are
 * lead * B^(sb1 - sa0) + p / B^(sb0 - sb1) for every lead over the sites that leave,
 * [sa1, sa0).  Column 0 has one path and every transition out of it is legal. */
struct Range {
    int first, step, n;
};

__device__ __forceinline__ Range preds(const int32_t *SA, const int32_t *SB, int x, int p, int B) {
    if (x - 1 == 0) return Range{0, 1, 1};
    const int sa1 = SA[x - 1], sb1 = SB[x - 1], sa0 = SA[x], sb0 = SB[x];
    return Range{p / ipow(B, sb0 - sb1), ipow(B, sb1 - sa0), ipow(B, sa0 - sa1)};
}

/* The legal successors of path q of column x in column x+1, in increasing order:
 * column x+1 holds sites [sa2, sb2); the shared sites [sa2, sb0) are q's trailing
 * digits, so the legal p are (q mod B^(sb0 - sa2)) * B^(sb2 - sb0) + trail for every
 * trail over the sites that enter, [sb0, sb2).  From column 0 every path of column 1. */
__device__ __forceinline__ Range succs(const int32_t *SA, const int32_t *SB, const int32_t *pp, int x, int q,
                                       int B) {
    if (x == 0) return Range{0, 1, pp[2] - pp[1]};
    const int sa2 = SA[x + 1], sb2 = SB[x + 1], sb0 = SB[x];
    const int n = ipow(B, sb2 - sb0);
    return Range{(q % ipow(B, sb0 - sa2)) * n, 1, n};
}

struct PathView {
    const int32_t *pp, *SA, *SB, *ecol, *ekid;
    const int64_t *doffE;
    int64_t plane;
    int B;
};

__device__ __forceinline__ PathView view(const PathTab &P, const TaskDev &t, int ti) {
    PathView v;
    v.pp = P.pp + P.col_off[ti];
    v.SA = P.sa + P.col_off[ti];
    v.SB = P.sb + P.col_off[ti];
    v.ecol = P.ecol + P.ent_off[ti];
    v.ekid = P.ekid + P.ent_off[ti];
    v.doffE = P.doffE + t.diag_off;
    v.plane = P.plane_off[ti];
    v.B = P.nopt;
    return v;
}

} // namespace

/* The inputs of an entry's emissions that depend only on its k-mer and event
 * (Gaussian tables): loaded a diagonal before they are used. */
struct RawEmit {
    double gx, mmu, misd, mca, mnmu, mnb, ymu, yisd, yca, ynmu, ynb, evm, evn, evl, evi;
};

__device__ __forceinline__ PathEnt load_ent(const PathEnt *et, const PathDiag &D, int tid) {
    if (tid < D.e1 - D.e0) return et[D.e0 + tid];
    return PathEnt{0, 0, 0, 1, 0, 0};
}

__device__ __forceinline__ void load_raw(const Model<double> &m, const Args<double> &a, const PathEnt &E, bool act,
                                         int64_t d, int64_t evc, RawEmit &r) {
    const int kk = act ? E.kk : 0;
    const int64_t y = d - E.x;
    const int64_t j = evc + ((act && y >= 1) ? y - 1 : 0);
    r.gx = m.gx[kk];
    r.mmu = m.mmu[kk]; r.misd = m.misd[kk]; r.mca = m.mca[kk]; r.mnmu = m.mnmu[kk]; r.mnb = m.mnb[kk];
    r.ymu = m.ymu[kk]; r.yisd = m.yisd[kk]; r.yca = m.yca[kk]; r.ynmu = m.ynmu[kk]; r.ynb = m.ynb[kk];
    r.evm = a.evm[j]; r.evn = a.evn[j]; r.evl = a.evl[j]; r.evi = a.evi[j];
}


/* Diagonal d of the pipelined forward sweep.  Ec, Rc: this thread's entry on d and
 * its emission inputs; En: its entry on d+1, whose emission inputs go to Rn; En3
 * receives its entry on d+3 (the entry on d+2 is in flight in a fourth slot).  The
 * diagonal records slide by one. */
template <bool LDS>
__device__ __forceinline__ void fwd_step(const Args<double> &a, const LutEntry<double> *lut, const PathEnt *et,
                                         const PathDiag *dg, double *ring,
                                         int R, int64_t N, int64_t evc, bool gauss, bool ragged_left, double s,
                                         double inv_var, int64_t d, PathDiag &Dm2, PathDiag &Dm1, PathDiag &D0,
                                         PathDiag &D1, PathDiag &D2, PathDiag &D3, const PathEnt &Ec,
                                         const PathEnt &En, PathEnt &En3, const RawEmit &Rc, RawEmit &Rn) {
    const Model<double> &m = a.m;
    const int RM = R - 1, tid = threadIdx.x;
    const double NI = -__builtin_huge_val();
    const PathDiag EMPTY = {1, 0, 0, 0, 0};
    const bool act = tid < D0.e1 - D0.e0;
    /* next diagonals' inputs */
    if (gauss) load_raw(m, a, En, tid < D1.e1 - D1.e0, d + 1, evc, Rn);
    En3 = load_ent(et, D3, tid);
    const PathDiag D4 = d + 4 <= N ? dg[d + 4] : EMPTY;
    double *rc = ring + (d % 3) * 3 * R, *r1 = ring + ((d + 2) % 3) * 3 * R, *r2 = ring + ((d + 1) % 3) * 3 * R;
    if (act) {
        const double tMC = m.trans[0], tMFX = m.trans[1], tMFY = m.trans[2], tGOX = m.trans[3], tGOY = m.trans[4];
        const double tGEX = m.trans[5], tGEY = m.trans[6], tSWX = m.trans[7];
        const int x = Ec.x;
        const int64_t y = d - x;
        const int es = (D0.e0 + tid) & RM;
        double M = NI, X = NI, Y = NI;
        if (d == 0) {
            /* start or ragged-start values (impl/stateMachine.c:1022-1031) */
            M = ragged_left ? NI : 0.0;
            X = Y = ragged_left ? 0.0 : NI;
        } else {
            double eX, eM = 0.0, eY = 0.0;
            if (gauss) {
                eX = Rc.gx;
                if (y >= 1) {
                    if (m.std_y) {
                        gauss_both_std(Rc.mmu, Rc.misd, Rc.mca, Rc.mnmu, Rc.mnb, Rc.evm, Rc.evn, Rc.evl, Rc.evi, s, eM,
                                       eY);
                    } else {
                        eM = gauss_emit(Rc.mmu, Rc.misd, Rc.mca, Rc.mnmu, Rc.mnb, Rc.evm, Rc.evn, Rc.evl, Rc.evi, s);
                        eY = gauss_emit(Rc.ymu, Rc.yisd, Rc.yca, Rc.ynmu, Rc.ynb, Rc.evm, Rc.evn, Rc.evl, Rc.evi, s);
                    }
                }
            } else {
                eX = Emit<double>::gapx(m, Ec.kk);
                if (y >= 1) {
                    eM = Emit<double>::match(m, a, Ec.kk, evc + (y - 1), s, inv_var);
                    eY = eM;
                }
            }
            /* lower -> gapX from d-1 and middle -> match from d-2 read the same
               predecessor entries: one loop, the two chains independent (each keeps its
               own order of log-adds), a chain whose neighbour is outside the band left
               as it is */
            const bool hL = x - 1 >= Dm1.lo && x - 1 <= Dm1.hi, hM = x - 1 >= Dm2.lo && x - 1 <= Dm2.hi;
            const bool hU = x >= Dm1.lo && x <= Dm1.hi && x >= 1;
            const double aX0 = eX + tGOX, aX1 = eX + tGEX, aX2 = eX + tSWX;
            const double aM0 = eM + tMC, aM1 = eM + tMFX, aM2 = eM + tMFY;
            { /* upper -> gapY, the same path (a third independent chain) */
                double Yn = logadd_lut(Y, r1[es] + (eY + tGOY), lut);
                Yn = logadd_lut(Yn, r1[2 * R + es] + (eY + tGEY), lut);
                Y = hU ? Yn : Y;
            }
            if (hL || hM) {
                for (int i = 0; i < Ec.pn; i++) {
                    const int eq = (Ec.pfirst + i * Ec.pstep) & RM;
                    double Xn = logadd_lut(X, r1[eq] + aX0, lut);
                    double Mn = logadd_lut(M, r2[eq] + aM0, lut);
                    Xn = logadd_lut(Xn, r1[R + eq] + aX1, lut);
                    Mn = logadd_lut(Mn, r2[R + eq] + aM1, lut);
                    Xn = logadd_lut(Xn, r1[2 * R + eq] + aX2, lut);
                    Mn = logadd_lut(Mn, r2[2 * R + eq] + aM2, lut);
                    X = hL ? Xn : X;
                    M = hM ? Mn : M;
                }
            }
        }
        rc[es] = M;
        rc[R + es] = X;
        rc[2 * R + es] = Y;
        a.fM[D0.foff + tid] = M;
        a.fX[D0.foff + tid] = X;
        a.fY[D0.foff + tid] = Y;
    }
    if constexpr (LDS) {
        /* only the LDS ring is exchanged: the plane stores are read by the next
           kernel, and the loads issued above stay in flight across the barrier */
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    } else {
        __syncthreads();
    }
    Dm2 = Dm1;
    Dm1 = D0;
    D0 = D1;
    D1 = D2;
    D2 = D3;
    D3 = D4;
}

/* The forward sweep of a task whose diagonals all fit the workgroup (one entry per
 * thread): each step loads the emission inputs of the next diagonal and the entry
 * records of the third one on, so a step waits on LDS and the barrier rather than on
 * a chain of dependent loads.  Unrolled by four so that the four entry slots and the
 * two emission-input slots trade roles without register moves (a move would wait for
 * the loads in flight).  Same arithmetic, in the same order, as the strided loop. */
template <bool LDS>
__device__ __forceinline__ void forward_pipelined(const Args<double> &a, const LutEntry<double> *lut, const PathTab &P,
                                                  const TaskDev &t, int ti, double *ring, int R) {
    const PathEnt *et = P.ent + P.ent_off[ti];
    const PathDiag *dg = P.diag + t.diag_off;
    const int tid = threadIdx.x;
    const int64_t N = (int64_t)t.lX + t.lY;
    const int64_t evc = t.evc_off;
    const bool gauss = a.m.kind != 2;
    const bool ragged_left = t.ragged_left != 0;
    const double s = a.m.kind == 1 ? 1.0 : t.scale / t.var;
    const double inv_var = 1.0 / t.var;
    const PathDiag EMPTY = {1, 0, 0, 0, 0};
    PathDiag Dm2 = EMPTY, Dm1 = EMPTY, D0 = dg[0];
    PathDiag D1 = N >= 1 ? dg[1] : EMPTY, D2 = N >= 2 ? dg[2] : EMPTY, D3 = N >= 3 ? dg[3] : EMPTY;
    PathEnt EA = load_ent(et, D0, tid), EB = load_ent(et, D1, tid), EC = load_ent(et, D2, tid);
    PathEnt ED = {0, 0, 0, 1, 0, 0};
    RawEmit RA = {}, RB = {};
    if (gauss) load_raw(a.m, a, EA, tid < D0.e1 - D0.e0, 0, evc, RA);
#define SA_FWD_STEP(dd, E0, E1, E3, R0, R1)                                                                         \
    fwd_step<LDS>(a, lut, et, dg, ring, R, N, evc, gauss, ragged_left, s, inv_var, dd, Dm2, Dm1, D0, D1, D2, D3, E0, \
                  E1, E3, R0, R1)
    int64_t d = 0;
    for (; d + 3 <= N; d += 4) {
        SA_FWD_STEP(d, EA, EB, ED, RA, RB);
        SA_FWD_STEP(d + 1, EB, EC, EA, RB, RA);
        SA_FWD_STEP(d + 2, EC, ED, EB, RA, RB);
        SA_FWD_STEP(d + 3, ED, EA, EC, RB, RA);
    }
    for (; d <= N; d++) {
        SA_FWD_STEP(d, EA, EB, ED, RA, RB);
        EA = EB;
        EB = EC;
        EC = ED;
        RA = RB;
    }
#undef SA_FWD_STEP
}

__global__ void __launch_bounds__(256) paths_forward_kernel(Args<double> a, PathTab P) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int R = P.ring, RM = R - 1;
    unsigned char *mem = a.ring_g ? a.ring_g + (size_t)blockIdx.x * a.ring_g_stride : smem;
    double *ring = reinterpret_cast<double *>(mem); /* [3 diag][3 state][R] */
    const int ti = a.task_order ? a.task_order[blockIdx.x] : blockIdx.x;
    const TaskDev t = a.tasks[ti];
    const PathView v = view(P, t, ti);
    const Model<double> &m = a.m;
    const int tid = threadIdx.x, NT = blockDim.x;
    const int64_t N = (int64_t)t.lX + t.lY;
    const int32_t *xlo = a.xlo + t.diag_off, *xhi = a.xhi + t.diag_off;
    const int64_t evc = t.evc_off;
    const double s = m.kind == 1 ? 1.0 : t.scale / t.var;
    const double inv_var = 1.0 / t.var;
    const double tMC = m.trans[0], tMFX = m.trans[1], tMFY = m.trans[2], tGOX = m.trans[3], tGOY = m.trans[4];
    const double tGEX = m.trans[5], tGEY = m.trans[6], tSWX = m.trans[7];
    const double NI = -__builtin_huge_val();

    if (P.maxent[ti] <= NT) {
        /* the log-add's cubic coefficients from an LDS table (bit-identical to the
           compare-and-select form); an LDS ring is addressed as LDS (a pointer that
           may be either is a flat one) */
        __shared__ LutEntry<double> lut[16];
        lut_fill(lut, tid);
        __syncthreads();
        if (a.ring_g) forward_pipelined<false>(a, lut, P, t, ti, ring, R);
        else forward_pipelined<true>(a, lut, P, t, ti, reinterpret_cast<double *>(smem), R);
        return;
    }
    /* diagonals wider than the workgroup: entries strided over the threads */
    for (int64_t d = 0; d <= N; d++) {
        const int lo = xlo[d], hi = xhi[d];
        const int e0 = v.pp[lo], e1 = v.pp[hi + 1];
        const bool h2 = d >= 2;
        const int lo1 = d >= 1 ? xlo[d - 1] : 1, hi1 = d >= 1 ? xhi[d - 1] : 0;
        const int lo2 = h2 ? xlo[d - 2] : 1, hi2 = h2 ? xhi[d - 2] : 0;
        double *rc = ring + (d % 3) * 3 * R, *r1 = ring + ((d + 2) % 3) * 3 * R, *r2 = ring + ((d + 1) % 3) * 3 * R;
        const int64_t pb = v.plane + v.doffE[d];
        for (int en = e0 + tid; en < e1; en += NT) {
            const int x = v.ecol[en];
            const int p = en - v.pp[x];
            const int64_t y = d - x;
            double M = NI, X = NI, Y = NI;
            if (d == 0) {
                /* start or ragged-start values (impl/stateMachine.c:1022-1031) */
                M = t.ragged_left ? NI : 0.0;
                X = Y = t.ragged_left ? 0.0 : NI;
            } else {
                const int32_t kk = v.ekid[en];
                const Range pr = x >= 1 ? preds(v.SA, v.SB, x, p, v.B) : Range{0, 1, 0};
                if (x - 1 >= lo1 && x - 1 <= hi1) { /* lower -> gapX */
                    const double eX = Emit<double>::gapx(m, kk);
                    for (int i = 0; i < pr.n; i++) {
                        const int eq = (v.pp[x - 1] + pr.first + i * pr.step) & RM;
                        X = logadd(X, r1[eq] + (eX + tGOX));
                        X = logadd(X, r1[R + eq] + (eX + tGEX));
                        X = logadd(X, r1[2 * R + eq] + (eX + tSWX));
                    }
                }
                if (x - 1 >= lo2 && x - 1 <= hi2) { /* middle -> match (x >= 1, y >= 1) */
                    const double eM = Emit<double>::match(m, a, kk, evc + (y - 1), s, inv_var);
                    for (int i = 0; i < pr.n; i++) {
                        const int eq = (v.pp[x - 1] + pr.first + i * pr.step) & RM;
                        M = logadd(M, r2[eq] + (eM + tMC));
                        M = logadd(M, r2[R + eq] + (eM + tMFX));
                        M = logadd(M, r2[2 * R + eq] + (eM + tMFY));
                    }
                }
                if (x >= lo1 && x <= hi1 && x >= 1) { /* upper -> gapY, the same path */
                    const double eY = Emit<double>::gapy(m, a, kk, evc + (y - 1), s, inv_var);
                    const int eu = en & RM;
                    Y = logadd(Y, r1[eu] + (eY + tGOY));
                    Y = logadd(Y, r1[2 * R + eu] + (eY + tGEY));
                }
            }
            const int es = en & RM;
            rc[es] = M;
            rc[R + es] = X;
            rc[2 * R + es] = Y;
            a.fM[pb + (en - e0)] = M;
            a.fX[pb + (en - e0)] = X;
            a.fY[pb + (en - e0)] = Y;
        }
        __syncthreads();
    }
}

/* dpDiagonal_dotProduct's sequential log-add chain (impl/pairwiseAligner.c:1014-1027) */
__device__ __forceinline__ double chain_paths(const double *term, int n) {
    double acc = -__builtin_huge_val();
    for (int i = 0; i < n; i++) acc = logadd(acc, term[i]);
    return acc;
}

/* LDS: the ring and term rows are in LDS (addressed as LDS; a pointer that may be
 * either is a flat one), else in this workgroup's slice of HBM. */
template <bool LDS>
__global__ void __launch_bounds__(256) paths_backward_kernel(Args<double> a, PathTab P) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int R = P.ring, RM = R - 1;
    unsigned char *mem = LDS ? smem : a.ring_g + (size_t)blockIdx.x * a.ring_g_stride;
    double *termA = reinterpret_cast<double *>(mem); /* R (cells of a diagonal) */
    double *termB = termA + R;                       /* R */
    double *bcast = termB + R;                       /* 2 */
    double *ring = bcast + 2;                        /* [3 diag][3 state][R] */
    const int wi = a.win_order[blockIdx.x];
    const WinDev W = a.wins[wi];
    const TaskDev t = a.tasks[W.task];
    const PathView v = view(P, t, W.task);
    const Model<double> &m = a.m;
    const int tid = threadIdx.x, NT = blockDim.x;
    const int32_t *xlo = a.xlo + t.diag_off, *xhi = a.xhi + t.diag_off;
    const int64_t evc = t.evc_off;
    const double s = m.kind == 1 ? 1.0 : t.scale / t.var;
    const double inv_var = 1.0 / t.var;
    const double tMC = m.trans[0], tMFX = m.trans[1], tMFY = m.trans[2], tGOX = m.trans[3], tGOY = m.trans[4];
    const double tGEX = m.trans[5], tGEY = m.trans[6], tSWX = m.trans[7];
    const double NI = -__builtin_huge_val();
    const int D = W.D, TF = W.TF, TB = W.T;
    const bool ragged_end = W.last && t.ragged_right;
    const double e0v = ragged_end ? m.rendv[0] : m.endv[0];
    const double e1v = ragged_end ? m.rendv[1] : m.endv[1];
    const double e2v = ragged_end ? m.rendv[2] : m.endv[2];
    const double thr = a.threshold;
    int32_t *out = a.out + 4 * W.out_off;
    __shared__ int s_cursor;
    if (tid == 0) s_cursor = 0;
    /* the recurrence's log-add takes its cubic coefficients from an LDS table
       (bit-identical to the compare-and-select form) */
    __shared__ LutEntry<double> lut[16];
    lut_fill(lut, tid);
    __syncthreads();
    int posterior_count = 0;
    double total = NI;
    __shared__ int s_wave_count[4];

    for (int e = D; e > TB; e--) {
        const int lo = xlo[e], hi = xhi[e];
        const int en0 = v.pp[lo], en1 = v.pp[hi + 1];
        const bool h1 = e + 1 <= D, h2 = e + 2 <= D;
        const int lo1 = h1 ? xlo[e + 1] : 1, hi1 = h1 ? xhi[e + 1] : 0;
        const int lo2 = h2 ? xlo[e + 2] : 1, hi2 = h2 ? xhi[e + 2] : 0;
        double *rc = ring + (e % 3) * 3 * R, *r1 = ring + ((e + 1) % 3) * 3 * R, *r2 = ring + ((e + 2) % 3) * 3 * R;
        for (int en = en0 + tid; en < en1; en += NT) {
            const int x = v.ecol[en];
            const int q = en - v.pp[x];
            const int y = e - x;
            double M, X, Y;
            if (e == D) {
                M = e0v;
                X = e1v;
                Y = e2v;
            } else {
                M = X = Y = NI;
                const bool hmid = x + 1 >= lo2 && x + 1 <= hi2, hlow = x + 1 >= lo1 && x + 1 <= hi1;
                const Range sr = hmid || hlow ? succs(v.SA, v.SB, v.pp, x, q, v.B) : Range{0, 1, 0};
                if (hmid) { /* middle: cell (x+1, y+1) on e+2 */
                    for (int i = 0; i < sr.n; i++) {
                        const int ep = v.pp[x + 1] + sr.first + i;
                        const double eM = Emit<double>::match(m, a, v.ekid[ep], evc + y, s, inv_var);
                        const double BM = r2[ep & RM];
                        M = logadd_lut(M, BM + (eM + tMC), lut);
                        X = logadd_lut(X, BM + (eM + tMFX), lut);
                        Y = logadd_lut(Y, BM + (eM + tMFY), lut);
                    }
                }
                if (x >= lo1 && x <= hi1 && x >= 1) { /* upper: cell (x, y+1) on e+1, the same path */
                    const double eY = Emit<double>::gapy(m, a, v.ekid[en], evc + y, s, inv_var);
                    const double BY = r1[2 * R + (en & RM)];
                    M = logadd_lut(M, BY + (eY + tGOY), lut);
                    Y = logadd_lut(Y, BY + (eY + tGEY), lut);
                }
                if (hlow) { /* lower: cell (x+1, y) on e+1 */
                    for (int i = 0; i < sr.n; i++) {
                        const int ep = v.pp[x + 1] + sr.first + i;
                        const double eX = Emit<double>::gapx(m, v.ekid[ep]);
                        const double BX = r1[R + (ep & RM)];
                        M = logadd_lut(M, BX + (eX + tGOX), lut);
                        X = logadd_lut(X, BX + (eX + tGEX), lut);
                        Y = logadd_lut(Y, BX + (eX + tSWX), lut);
                    }
                }
            }
            const int es = en & RM;
            rc[es] = M;
            rc[R + es] = X;
            rc[2 * R + es] = Y;
        }
        __syncthreads();

        if (e <= TF) {
            const int64_t fb = v.plane + v.doffE[e];
            if (posterior_count % 10 == 0) {
                /* diagonalCalculationTotalProbability (impl/pairwiseAligner.c:1182-1200): per
                   cell the paths' dot products (hdCell_totalProbability, :677-692) */
                const int w = hi - lo + 1;
                for (int i = tid; i < w; i += NT) {
                    const int x = lo + i;
                    double hd = NI;
                    for (int r = v.pp[x]; r < v.pp[x + 1]; r++) {
                        const int64_t fi = fb + (r - en0);
                        double c = a.fM[fi] + rc[r & RM];
                        c = logadd(c, a.fX[fi] + rc[R + (r & RM)]);
                        c = logadd(c, a.fY[fi] + rc[2 * R + (r & RM)]);
                        hd = logadd(hd, c);
                    }
                    termA[i] = hd;
                }
                int w1 = 0;
                if (h1) {
                    /* the match flux into e+1 from e-1 (dotted with B[e+1]) */
                    const int lo1e = xlo[e + 1];
                    w1 = xhi[e + 1] - lo1e + 1;
                    const int lom = xlo[e - 1], him = xhi[e - 1];
                    const int64_t fbm = v.plane + v.doffE[e - 1];
                    const int enm0 = v.pp[lom];
                    for (int i = tid; i < w1; i += NT) {
                        const int x = lo1e + i;
                        const int y = e + 1 - x;
                        const bool hm = x - 1 >= lom && x - 1 <= him && x >= 1 && y >= 1;
                        double hd = NI;
                        for (int r = 0; r < v.pp[x + 1] - v.pp[x]; r++) {
                            const int er = v.pp[x] + r;
                            double Mv = NI;
                            if (hm) {
                                const double eM = Emit<double>::match(m, a, v.ekid[er], evc + (y - 1), s, inv_var);
                                const Range pr = preds(v.SA, v.SB, x, r, v.B);
                                for (int i = 0; i < pr.n; i++) {
                                    const int64_t fi = fbm + (v.pp[x - 1] + pr.first + i * pr.step - enm0);
                                    Mv = logadd(Mv, a.fM[fi] + (eM + tMC));
                                    Mv = logadd(Mv, a.fX[fi] + (eM + tMFX));
                                    Mv = logadd(Mv, a.fY[fi] + (eM + tMFY));
                                }
                            }
                            hd = logadd(hd, Mv + r1[er & RM]);
                        }
                        termB[i] = hd;
                    }
                }
                __syncthreads();
                if (tid == 0) {
                    double tot = chain_paths(termA, w);
                    if (h1) tot = logadd(tot, chain_paths(termB, w1));
                    bcast[0] = tot;
                }
                __syncthreads();
                total = bcast[0];
                if (e == TF && tid == 0) a.win_total[wi] = total;
            }
            posterior_count++;
            /* diagonalCalculationPosteriorMatchProbs (impl/pairwiseAligner.c:1202-1253):
               every path of every cell, entries in order */
            for (int b0 = en0; b0 < en1; b0 += NT) {
                const int en = b0 + tid;
                bool keep = false;
                double pv = 0.0;
                int x = 0, y = 0;
                if (en < en1) {
                    x = v.ecol[en];
                    y = e - x;
                    if (x > 0 && y > 0) {
                        pv = exp((a.fM[fb + (en - en0)] + rc[en & RM]) - total);
                        keep = pv >= thr;
                    }
                }
                /* compaction in entry order: per wavefront ballots, then the wavefronts in order */
                const unsigned long long bal = __ballot(keep);
                const int wv = tid >> 6, nwv = (NT + 63) >> 6;
                if ((tid & 63) == 0) s_wave_count[wv] = __popcll(bal);
                __syncthreads();
                int base = s_cursor;
                for (int k = 0; k < wv; k++) base += s_wave_count[k];
                if (keep) {
                    const int idx = base + __popcll(bal & ((1ull << (tid & 63)) - 1ull));
                    if (idx < W.out_cap) {
                        if (pv > 1.0) pv = 1.0;
                        out[4 * idx + 0] = (int32_t)floor(pv * (double)kPairProb1);
                        out[4 * idx + 1] = x - 1;
                        out[4 * idx + 2] = y - 1;
                        out[4 * idx + 3] = en - v.pp[x];
                    }
                }
                __syncthreads();
                if (tid == 0) {
                    int add = 0;
                    for (int k = 0; k < nwv; k++) add += s_wave_count[k];
                    s_cursor += add;
                }
                __syncthreads();
            }
        }
    }
    if (tid == 0) a.win_count[wi] = s_cursor;
}

hipError_t launch_paths_forward(const Args<double> &a, const PathTab &P, int n_tasks, hipStream_t st) {
    const size_t lds = a.ring_g ? 0 : paths_forward_bytes(P.ring);
    hipLaunchKernelGGL(paths_forward_kernel, dim3(n_tasks), dim3(256), lds, st, a, P);
    return hipGetLastError();
}

hipError_t launch_paths_backward(const Args<double> &a, const PathTab &P, int n_wins, hipStream_t st) {
    const size_t lds = a.ring_g ? 0 : paths_backward_bytes(P.ring);
    if (a.ring_g) hipLaunchKernelGGL(paths_backward_kernel<false>, dim3(n_wins), dim3(256), 0, st, a, P);
    else hipLaunchKernelGGL(paths_backward_kernel<true>, dim3(n_wins), dim3(256), lds, st, a, P);
    return hipGetLastError();
}

} // namespace sa
