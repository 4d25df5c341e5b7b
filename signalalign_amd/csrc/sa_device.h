/*
 * sa_device.h -- device arithmetic shared by the kernels: the reference's lookup
 * log-add, the emission densities, wave reductions and one-lane DPP shifts.
 */
#ifndef SA_DEVICE_H
#define SA_DEVICE_H

#include "sa_kernels.h"

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace sa {

/* ------------------------------------------------------------------------ */
/* numerics                                                                  */
/* ------------------------------------------------------------------------ */

template <typename T> __device__ __forceinline__ T ninf();
template <> __device__ __forceinline__ float ninf<float>() { return -__builtin_huge_valf(); }
template <> __device__ __forceinline__ double ninf<double>() { return -__builtin_huge_val(); }

__device__ __forceinline__ double poly3(double d, double c0, double c1, double c2, double c3) {
#pragma clang fp contract(off)
    return ((c0 * d + c1) * d + c2) * d + c3;
}
__device__ __forceinline__ float poly3(float d, float c0, float c1, float c2, float c3) {
    return __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(c0, d, c1), d, c2), d, c3);
}

/* logAdd (impl/pairwiseAligner.c:237-254): lo + lookup(hi - lo) while hi - lo < 7.5,
 * else hi.  A -inf operand makes the difference inf or NaN, which selects hi. */
template <typename T>
__device__ __forceinline__ T logadd(T a, T b) {
    T hi = a > b ? a : b;
    T lo = a > b ? b : a;
    T d = hi - lo;
    const bool p1 = d <= (T)1.00f, p2 = d <= (T)2.50f, p3 = d <= (T)4.50f;
    T c0 = p1 ? (T)-0.009350833524763f : p2 ? (T)-0.014532321752540f : p3 ? (T)-0.004605031767994f : (T)-0.000458661602210f;
    T c1 = p1 ? (T)0.130659527668286f : p2 ? (T)0.139942324101744f : p3 ? (T)0.063427417320019f : (T)0.009695946122598f;
    T c2 = p1 ? (T)0.498799810682272f : p2 ? (T)0.495635523139337f : p3 ? (T)0.695956496475118f : (T)0.930734667215156f;
    T c3 = p1 ? (T)0.693203116424741f : p2 ? (T)0.692140569840976f : p3 ? (T)0.514272634594009f : (T)0.168037164329057f;
    T r = poly3(d, c0, c1, c2, c3) + lo;
    return d < (T)7.5 ? r : hi;
}

/* The same log-add with the cubic's coefficients read from a 16-entry table in LDS
 * instead of selected by three compares.  Entry k = ceil(2 * min(d, 7.5)) holds the
 * piece containing (0.5(k-1), 0.5k]; the pieces change at 1.0, 2.5 and 4.5, all
 * multiples of 0.5, so each entry lies inside one piece and the reference's
 * inclusive bounds (d <= 1.0, d <= 2.5, d <= 4.5) select the same piece.  d is NaN
 * only when both operands are -inf (min(NaN, 7.5) = 7.5 picks some entry; the
 * result is hi either way).  Results are bit-identical to logadd().
 *
 * FP32 (the recurrence of the FP32 mode, not the FP64 parity path) uses a shorter
 * form of the same function: with u = min(2d, 15.5) and g(d) = lookup(d) - d, the
 * result is hi + g, where g is the piece's cubic with the linear coefficient less
 * one, its coefficients scaled so that the cubic takes u (c0/8, c1/4, (c2-1)/2, c3:
 * powers of two, so g(u) rounds as g(d) would).  A 17th entry of zeros serves
 * d > 7.5 and the NaN of two -inf operands (min(NaN, 15.5) = 15.5), so no final
 * compare is needed; the only difference from the reference's piecewise function is
 * a difference of exactly 7.5, which takes the last piece instead of the maximum.
 * The entries are stored in reverse order, entry 16 - ceil(u), so that the index is
 * 16 + floor(-u) (one conversion with a negated operand). */
template <typename T> struct LutEntry;
template <> struct LutEntry<float> { float c[4]; };
template <> struct LutEntry<double> { double c[4]; };
template <typename T> struct LutSize { static constexpr int n = 16; };
template <> struct LutSize<float> { static constexpr int n = 17; };

template <typename T>
__device__ __forceinline__ void lut_fill(LutEntry<T> *lut, int tid) {
    if (tid < 16) {
        const int piece = tid <= 2 ? 0 : tid <= 5 ? 1 : tid <= 9 ? 2 : 3;
        const float C[4][4] = {
            {-0.009350833524763f, 0.130659527668286f, 0.498799810682272f, 0.693203116424741f},
            {-0.014532321752540f, 0.139942324101744f, 0.495635523139337f, 0.692140569840976f},
            {-0.004605031767994f, 0.063427417320019f, 0.695956496475118f, 0.514272634594009f},
            {-0.000458661602210f, 0.009695946122598f, 0.930734667215156f, 0.168037164329057f}};
        for (int i = 0; i < 4; i++) lut[tid].c[i] = (T)C[piece][i];
    }
}

/* the FP32 form's coefficients of piece 0..3 (4: zeros) */
__device__ __forceinline__ void lut_coeffs_f32(int piece, float *c) {
    const float C[4][4] = {
        {-0.009350833524763f, 0.130659527668286f, 0.498799810682272f, 0.693203116424741f},
        {-0.014532321752540f, 0.139942324101744f, 0.495635523139337f, 0.692140569840976f},
        {-0.004605031767994f, 0.063427417320019f, 0.695956496475118f, 0.514272634594009f},
        {-0.000458661602210f, 0.009695946122598f, 0.930734667215156f, 0.168037164329057f}};
    if (piece > 3) { c[0] = c[1] = c[2] = c[3] = 0.0f; return; }
    c[0] = C[piece][0] * 0.125f;
    c[1] = C[piece][1] * 0.25f;
    c[2] = (C[piece][2] - 1.0f) * 0.5f;
    c[3] = C[piece][3];
}
template <>
__device__ __forceinline__ void lut_fill<float>(LutEntry<float> *lut, int tid) {
    if (tid < 17) {
        const int k = 16 - tid; /* entry tid serves ceil(u) = k */
        lut_coeffs_f32(k <= 2 ? 0 : k <= 5 ? 1 : k <= 9 ? 2 : k <= 15 ? 3 : 4, lut[tid].c);
    }
}

/* u = min(2 |a - b|, 15.5) (the doubling is a separate multiply: the output modifier
   x2 is not applied while FP32 denormals are enabled) and the table index
   16 + floor(-u) (one conversion with a negated operand) */
__device__ __forceinline__ float la_u(float t) { return __builtin_fminf(__builtin_fabsf(t) * 2.0f, 15.5f); }
__device__ __forceinline__ int la_index(float u) {
    int m;
    asm("v_cvt_flr_i32_f32_e64 %0, -%1" : "=v"(m) : "v"(u));
    return m + 16;
}

__device__ __forceinline__ float logadd_lut(float a, float b, const LutEntry<float> *lut) {
    const float hi = __builtin_fmaxf(a, b);
    const float u = la_u(a - b);
    const float4 c = *reinterpret_cast<const float4 *>(&lut[la_index(u)]);
    return hi + __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(c.x, u, c.y), u, c.z), u, c.w);
}

/* The FP32 form with its coefficients selected by compares (no table): bit-identical
   to logadd_lut(float).  The constants are the table's, folded at compile time with
   the same float operations as lut_coeffs_f32. */
namespace la32 {
constexpr float C[4][4] = {
    {-0.009350833524763f, 0.130659527668286f, 0.498799810682272f, 0.693203116424741f},
    {-0.014532321752540f, 0.139942324101744f, 0.495635523139337f, 0.692140569840976f},
    {-0.004605031767994f, 0.063427417320019f, 0.695956496475118f, 0.514272634594009f},
    {-0.000458661602210f, 0.009695946122598f, 0.930734667215156f, 0.168037164329057f}};
constexpr float k0(int p) { return C[p][0] * 0.125f; }
constexpr float k1(int p) { return C[p][1] * 0.25f; }
constexpr float k2(int p) { return (C[p][2] - 1.0f) * 0.5f; }
constexpr float k3(int p) { return C[p][3]; }
} // namespace la32
__device__ __forceinline__ float logadd(float a, float b) {
    const float hi = __builtin_fmaxf(a, b);
    const float u = la_u(a - b);
    const bool p0 = u <= 2.0f, p1 = u <= 5.0f, p2 = u <= 9.0f, p3 = u <= 15.0f;
    using namespace la32;
    const float c0 = p0 ? k0(0) : p1 ? k0(1) : p2 ? k0(2) : p3 ? k0(3) : 0.0f;
    const float c1 = p0 ? k1(0) : p1 ? k1(1) : p2 ? k1(2) : p3 ? k1(3) : 0.0f;
    const float c2 = p0 ? k2(0) : p1 ? k2(1) : p2 ? k2(2) : p3 ? k2(3) : 0.0f;
    const float c3 = p0 ? k3(0) : p1 ? k3(1) : p2 ? k3(2) : p3 ? k3(3) : 0.0f;
    return hi + __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(c0, u, c1), u, c2), u, c3);
}

__device__ __forceinline__ double logadd_lut(double a, double b, const LutEntry<double> *lut) {
    const double hi = __builtin_fmax(a, b), lo = __builtin_fmin(a, b);
    const double d = hi - lo;
    const double dd = __builtin_fmin(d, 7.5);
    const int k = (int)__builtin_ceil(dd + dd);
    const double2 c01 = *reinterpret_cast<const double2 *>(&lut[k].c[0]);
    const double2 c23 = *reinterpret_cast<const double2 *>(&lut[k].c[2]);
    const double r = poly3(d, c01.x, c01.y, c23.x, c23.y) + lo;
    return d < 7.5 ? r : hi;
}

__device__ __forceinline__ double flog(double v) { return log(v); }
__device__ __forceinline__ float flog(float v) { return logf(v); }
__device__ __forceinline__ double fexp(double v) { return exp(v); }
/* FP32 posteriors: the hardware exp2 of v*log2(e) (relative error about 4e-7 at the
 * 0.01 threshold's exponent; the FP32 variant's tolerance is 2e-3) */
__device__ __forceinline__ float fexp(float v) { return __expf(v); }

/* a value every lane holds, made wave-uniform for the compiler (scalar registers,
 * scalar branches on it) */
__device__ __forceinline__ float uniform(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ double uniform(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

/* wave priority among the waves sharing a SIMD (higher issues first) */
__device__ __forceinline__ void set_priority(int p) {
    if (p >= 3) __builtin_amdgcn_s_setprio(3);
    else if (p == 2) __builtin_amdgcn_s_setprio(2);
    else if (p == 1) __builtin_amdgcn_s_setprio(1);
}

/* maximum over the wavefront, uniform: within each row of 16 lanes by DPP swaps (pairs,
   quads, half-row and row mirrors; no LDS round trip), then the four row maxima read
   lane by lane.  Exactly the largest value, as any order of comparisons gives. */
__device__ __forceinline__ float dpp_max_step(float v, int ctrl_sel) {
    int u;
    const int b = __float_as_int(v);
    switch (ctrl_sel) {
    case 0: u = __builtin_amdgcn_update_dpp(b, b, 0xb1, 0xf, 0xf, false); break;  /* quad_perm [1,0,3,2] */
    case 1: u = __builtin_amdgcn_update_dpp(b, b, 0x4e, 0xf, 0xf, false); break;  /* quad_perm [2,3,0,1] */
    case 2: u = __builtin_amdgcn_update_dpp(b, b, 0x141, 0xf, 0xf, false); break; /* row_half_mirror */
    default: u = __builtin_amdgcn_update_dpp(b, b, 0x140, 0xf, 0xf, false); break; /* row_mirror */
    }
    const float w = __int_as_float(u);
    return w > v ? w : v;
}
__device__ __forceinline__ float wave_max(float v) {
    v = dpp_max_step(v, 0);
    v = dpp_max_step(v, 1);
    v = dpp_max_step(v, 2);
    v = dpp_max_step(v, 3);
    const float r0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
    const float r1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 16));
    const float r2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32));
    const float r3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 48));
    const float a = r0 > r1 ? r0 : r1, b = r2 > r3 ? r2 : r3;
    return uniform(a > b ? a : b);
}
/* (FP64 sweeps do not renormalise; kept for completeness) */
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        double u = __shfl_xor(v, o, 64);
        v = u > v ? u : v;
    }
    return uniform(v);
}

/* ------------------------------------------------------------------------ */
/* emissions                                                                 */
/* ------------------------------------------------------------------------ */

/* emissions_signal_strawManGetKmerEventMatchProb[WithDescaling]
 * (impl/stateMachine.c:498-589): log N(descaled mean; mu, sd) + log IG(noise; nmu, lambda),
 * rearranged with per-k-mer and per-event constants:
 *   a = (evm - s*mu) / sd,  evm = (mean - shift)/var, s = scale/var (1, 0, 1 undescaled)
 *   e = ca - a^2/2 + evl - nb (noise - nmu)^2 / noise,  evl = -1.5 log(noise). */
template <typename T> __device__ __forceinline__ T fma_t(T x, T y, T z);
template <> __device__ __forceinline__ float fma_t(float x, float y, float z) { return __builtin_fmaf(x, y, z); }
template <> __device__ __forceinline__ double fma_t(double x, double y, double z) { return __builtin_fma(x, y, z); }

/* Every sweep evaluates the emissions through these, with each fused multiply-add
 * written out and no other contraction, so that one k-mer and event give the same
 * bits in every kernel and in every context the compiler inlines them into (the
 * register sweeps, the LDS sweeps and the multi-path sweeps can then meet in one
 * read, and a pair's posterior does not depend on which of them ran). */
template <typename T> struct GaussTerms {
    T a2;    /* a^2 */
    T noise; /* evl - nb (noise - nmu)^2 / noise */
};
template <typename T>
__device__ __forceinline__ GaussTerms<T> gauss_terms(T mu, T isd, T nmu, T nb, T evm, T evn, T evl, T evi, T s) {
#pragma clang fp contract(off)
    const T a = fma_t(-s, mu, evm) * isd;
    const T u = evn - nmu;
    return GaussTerms<T>{a * a, fma_t(-nb, (u * u) * evi, evl)};
}
template <typename T>
__device__ __forceinline__ T gauss_emit(T mu, T isd, T ca, T nmu, T nb, T evm, T evn, T evl, T evi, T s) {
#pragma clang fp contract(off)
    const GaussTerms<T> g = gauss_terms(mu, isd, nmu, nb, evm, evn, evl, evi, s);
    return fma_t((T)-0.5, g.a2, ca) + g.noise;
}
/* The extra-event emission of a model file's tables from the match table's constants:
 * the extra-event table is the match table with sd x 1.75 (impl/stateMachine.c:1410-1412),
 * so a becomes a/1.75 and ca loses log(1.75). */
template <typename T> __device__ __forceinline__ T gauss_ext_std(const GaussTerms<T> &g, T ca) {
#pragma clang fp contract(off)
    return fma_t((T)-0.16326530612244898, g.a2, ca - (T)0.55961578793542266) + g.noise;
}
/* both emissions of a model file's tables at once */
template <typename T>
__device__ __forceinline__ void gauss_both_std(T mu, T isd, T ca, T nmu, T nb, T evm, T evn, T evl, T evi, T s,
                                               T &em, T &ey) {
#pragma clang fp contract(off)
    const GaussTerms<T> g = gauss_terms(mu, isd, nmu, nb, evm, evn, evl, evi, s);
    em = fma_t((T)-0.5, g.a2, ca) + g.noise;
    ey = gauss_ext_std(g, ca);
}

/* emissions_signal_getHdpKmerDensity (impl/stateMachine.c:468-495) ->
 * dir_proc_density (impl/hdp.c:2583-2607) -> grid_spline_interp
 * (impl/hdp_math_utils.c:471-496): Hermite cubic on the row's uniform grid with linear
 * extrapolation, clamped at 0, divided by var, in log space.  Split in two so that a
 * sweep can issue the gathers (hdp_fetch) a diagonal before it needs the value
 * (hdp_finish); hdp_emit_row is the two back to back. */
template <typename T> struct HdpFetch {
    T u;        /* descaled event mean */
    T y0, s0;   /* knot il */
    T y1, s1;   /* knot il + 1 */
    T xl;       /* grid x of knot il */
    int32_t row;
};

/* The knot pair below u (clamped to the grid): below the grid that is knot 0, above it
 * knot G-1 is the right one of the last pair, so the extrapolations read the same
 * registers without a branch around the loads. */
/* (yrow: the row's interleaved (density, slope) pairs, in HBM or staged in LDS;
 * gridx: the knots, likewise) */
template <typename T>
__device__ __forceinline__ HdpFetch<T> hdp_fetch_r(const Model<T> &m, int32_t row, const T *yrow, T hmu, T evm, T s,
                                                   const T *gridx) {
    /* No multiply-add fusion in the HDP emission (here and in hdp_finish): several
       kernels compute it (the sweeps' own gathers, the emission plane's column pass,
       the LDS sweeps), and fusion is decided per kernel, so only unfused arithmetic
       gives every kernel the same bits for the same cell (it is also the reference's
       operation order) */
#pragma clang fp contract(off)
    HdpFetch<T> f;
    f.u = evm + hmu * ((T)1 - s);
    f.row = row;
    const int32_t G = m.grid_length;
    T t = (f.u - m.grid_x0) * m.inv_dx;
    t = t > (T)0 ? t : (T)0;
    t = t < (T)(G - 2) ? t : (T)(G - 2);
    const int32_t il = (int32_t)t;
    const T *ys = yrow + (int64_t)il * 2;
    f.y0 = ys[0]; f.s0 = ys[1]; f.y1 = ys[2]; f.s1 = ys[3];
    f.xl = gridx[il];
    return f;
}
template <typename T>
__device__ __forceinline__ HdpFetch<T> hdp_fetch_g(const Model<T> &m, int32_t row, T hmu, T evm, T s,
                                                   const T *gridx) {
    return hdp_fetch_r(m, row, m.ys + (int64_t)(row < 0 ? 0 : row) * m.grid_length * 2, hmu, evm, s, gridx);
}
/* (gridx: the model's knots, or a copy of them in LDS) */
template <typename T>
__device__ __forceinline__ HdpFetch<T> hdp_fetch(const Model<T> &m, int32_t row, T hmu, T evm, T s) {
    return hdp_fetch_g(m, row, hmu, evm, s, m.gridx);
}

template <typename T>
__device__ __forceinline__ T hdp_finish(const Model<T> &m, const HdpFetch<T> &f, T inv_var) {
#pragma clang fp contract(off)
    if (f.row < 0) return ninf<T>();
    T interp;
    if (f.u <= m.grid_x0) {
        interp = f.y0 - f.s0 * (m.grid_x0 - f.u);
    } else if (f.u >= m.grid_xn) {
        interp = f.y1 + f.s1 * (f.u - m.grid_xn);
    } else {
        const T dx = m.grid_dx;
        const T dy = f.y1 - f.y0;
        const T a = f.s0 * dx - dy;
        const T b = dy - f.s1 * dx;
        const T tl = (f.u - f.xl) * m.inv_dx;
        const T tr = (T)1 - tl;
        interp = tr * f.y0 + tl * f.y1 + tl * tr * (a * tr + b * tl);
    }
    T dens = interp > (T)0 ? interp : (T)0;
    return flog(inv_var * dens);
}

template <typename T>
__device__ __forceinline__ T hdp_emit_row(const Model<T> &m, int32_t row, T hmu, T evm, T s, T inv_var) {
    return hdp_finish(m, hdp_fetch(m, row, hmu, evm, s), inv_var);
}

template <typename T>
__device__ __forceinline__ T hdp_emit(const Model<T> &m, int32_t kid, T evm, T s, T inv_var) {
    return hdp_emit_row(m, m.hrow[kid], m.hmu[kid], evm, s, inv_var);
}

template <typename T>
struct Emit {
    /* match and extra-event emissions of k-mer `kid` against event j */
    __device__ __forceinline__ static void eval(const Model<T> &m, const Args<T> &a, int32_t kid, int64_t j, T s,
                                                T inv_var, T &em, T &ey) {
        T evm = a.evm[j];
        if (m.kind == 2) {
            em = hdp_emit(m, kid, evm, s, inv_var);
            ey = em;
            return;
        }
        T evn = a.evn[j], evl = a.evl[j], evi = a.evi[j];
        if (m.std_y) {
            gauss_both_std(m.mmu[kid], m.misd[kid], m.mca[kid], m.mnmu[kid], m.mnb[kid], evm, evn, evl, evi, s, em, ey);
            return;
        }
        em = gauss_emit(m.mmu[kid], m.misd[kid], m.mca[kid], m.mnmu[kid], m.mnb[kid], evm, evn, evl, evi, s);
        ey = gauss_emit(m.ymu[kid], m.yisd[kid], m.yca[kid], m.ynmu[kid], m.ynb[kid], evm, evn, evl, evi, s);
    }
    __device__ __forceinline__ static T match(const Model<T> &m, const Args<T> &a, int32_t kid, int64_t j, T s,
                                              T inv_var) {
        T evm = a.evm[j];
        if (m.kind == 2) return hdp_emit(m, kid, evm, s, inv_var);
        return gauss_emit(m.mmu[kid], m.misd[kid], m.mca[kid], m.mnmu[kid], m.mnb[kid], evm, a.evn[j], a.evl[j],
                          a.evi[j], s);
    }
    __device__ __forceinline__ static T gapy(const Model<T> &m, const Args<T> &a, int32_t kid, int64_t j, T s,
                                             T inv_var) {
        T evm = a.evm[j];
        if (m.kind == 2) return hdp_emit(m, kid, evm, s, inv_var);
        if (m.std_y)
            return gauss_ext_std(gauss_terms(m.mmu[kid], m.misd[kid], m.mnmu[kid], m.mnb[kid], evm, a.evn[j],
                                             a.evl[j], a.evi[j], s), m.mca[kid]);
        return gauss_emit(m.ymu[kid], m.yisd[kid], m.yca[kid], m.ynmu[kid], m.ynb[kid], evm, a.evn[j], a.evl[j],
                          a.evi[j], s);
    }
    /* gapX: emissions_kmer_getGapProb (impl/stateMachine.c:155-172) or the HDP
     * machine's constant log(0.1) (:1282) */
    __device__ __forceinline__ static T gapx(const Model<T> &m, int32_t kid) {
        if (m.kind == 2) return (T)-2.3025850929940455;
        return m.gx[kid];
    }
};


/* ------------------------------------------------------------------------ */
/* one-lane wave shifts (DPP wave_shr:1 / wave_shl:1, all 64 lanes active)    */
/* ------------------------------------------------------------------------ */

/* lane i receives lane i-1's value; lane 0 receives `fill` */
__device__ __forceinline__ int shr1(int v, int fill) {
    return __builtin_amdgcn_update_dpp(fill, v, 0x138, 0xf, 0xf, false);
}
/* lane i receives lane i+1's value; lane 63 receives `fill` */
__device__ __forceinline__ int shl1(int v, int fill) {
    return __builtin_amdgcn_update_dpp(fill, v, 0x130, 0xf, 0xf, false);
}
__device__ __forceinline__ float shr1(float v, float fill) {
    return __int_as_float(shr1(__float_as_int(v), __float_as_int(fill)));
}
__device__ __forceinline__ float shl1(float v, float fill) {
    return __int_as_float(shl1(__float_as_int(v), __float_as_int(fill)));
}
__device__ __forceinline__ double shr1(double v, double fill) {
    const long long b = __double_as_longlong(v), f = __double_as_longlong(fill);
    const int lo = shr1((int)(b & 0xffffffffll), (int)(f & 0xffffffffll));
    const int hi = shr1((int)(b >> 32), (int)(f >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double shl1(double v, double fill) {
    const long long b = __double_as_longlong(v), f = __double_as_longlong(fill);
    const int lo = shl1((int)(b & 0xffffffffll), (int)(f & 0xffffffffll));
    const int hi = shl1((int)(b >> 32), (int)(f >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

} // namespace sa

#endif
