// nuc3d_device.h -- device-side arithmetic of the per-block RHS path (gfx950).
//
// Two variants are compiled from this header (see nuc3d_kernels.hip); both
// translation units are built with -ffp-contract=off.
//   NUC3D_EXACT=1  reference operation order everywhere; pow(x,2) -> x*x.
//   NUC3D_EXACT=0  inside the WENO reconstructions only: FMA contraction and
//                  a one-division form of the nonlinear weights.  Flux
//                  splitting, differences and the RK update keep the
//                  reference order, because reassociating the large p/J
//                  terms there moves near-zero momentum components by
//                  ~1e-12 of their size.
//
// Reference formulas: schemes.hpp:17-44 (coefficients), weno5js.cpp:75-108,
// weno5z.cpp:74-109, crweno5.cpp:91-218, reconstructionBoundaryScheme.cpp:
// 52-474, centraldifference6th.cpp:43-45, centraldifference8th.cpp:49-52,
// boundaryScheme.cpp:100-126, physicsModel.cpp:258-357 (con2prim),
// :940-979 (local Lax-Friedrichs), :762-938 (AUSM, AUSM+), :1145-1398
// (point splitters), :199-234 (Sutherland), TVD-RK3.cpp:37-137.
#pragma once
#include <hip/hip_runtime.h>

#ifndef NUC3D_EXACT
#define NUC3D_EXACT 0
#endif

namespace nuc3d_dev {

// ---- scheme constants (schemes.hpp) ----
constexpr double A00 = 1.0 / 3.0, A01 = -7.0 / 6.0, A02 = 11.0 / 6.0;
constexpr double A10 = -1.0 / 6.0, A11 = 5.0 / 6.0, A12 = 1.0 / 3.0;
constexpr double A20 = 1.0 / 3.0, A21 = 5.0 / 6.0, A22 = -1.0 / 6.0;
constexpr double C0 = 0.1, C1 = 0.6, C2 = 0.3;          // WENO5 linear weights
constexpr double R0 = 0.2, R1 = 0.5, R2 = 0.3;          // CRWENO5 linear weights
constexpr double G0 = 13.0 / 12.0, G1 = 1.0 / 4.0;
constexpr double EPS = 1.0e-6;
constexpr double CD6_0 = 3.0 / 4.0, CD6_1 = -3.0 / 20.0, CD6_2 = 1.0 / 60.0;
constexpr double CD8_0 = 4.0 / 5.0, CD8_1 = -1.0 / 5.0, CD8_2 = 4.0 / 105.0, CD8_3 = -1.0 / 280.0;
constexpr double CD4_0 = 2.0 / 3.0, CD4_1 = -1.0 / 12.0;

enum { SCH_WENO5JS = 0, SCH_WENO5Z = 1, SCH_CRWENO5 = 2, SCH_CENTRAL8 = 3, SCH_HYBRID = 4 };
enum { RIE_LF = 0, RIE_AUSM = 1, RIE_AUSMP = 2 };
enum { BC_INLET = 0, BC_OUTLET = 1, BC_WALL = 2, BC_SYMM = 3 };

__device__ __forceinline__ double sq(double x) { return x * x; }

// Global-memory access through the block descriptors' field pointers.  Those
// pointers are generic to the compiler, which makes FLAT instructions of their
// loads and stores; a FLAT access also counts on the LDS counter, so a wait
// for an LDS read waits for every global load in flight.  The fields are
// global memory, and these access them as such.
typedef __attribute__((address_space(1))) double gdouble;
typedef __attribute__((address_space(1))) unsigned char gbyte;
typedef __attribute__((address_space(1))) unsigned gword;
__device__ __forceinline__ const gdouble *gp(const double *p) { return (const gdouble *)p; }
__device__ __forceinline__ double gld(const double *p) { return *(const gdouble *)p; }
__device__ __forceinline__ void gst(double *p, double v) { *(gdouble *)p = v; }
__device__ __forceinline__ unsigned char gldb(const unsigned char *p) { return *(const gbyte *)p; }
__device__ __forceinline__ void gstb(unsigned char *p, unsigned char v) { *(gbyte *)p = v; }
__device__ __forceinline__ unsigned gldw(const unsigned *p) { return *(const gword *)p; }
__device__ __forceinline__ void gstw(unsigned *p, unsigned v) { *(gword *)p = v; }

// ---- FP64 division in two halves ----
// The compiler's IEEE division a/b on gfx950 is v_div_scale (of b and of a),
// v_rcp_f64 and two Newton steps for the reciprocal, the quotient a*y with
// one FMA residual correction (v_div_fmas), and v_div_fixup for special
// values.  Where a and b are finite and in the range in which v_div_scale
// leaves them unscaled and v_div_fixup passes the quotient through (b
// normal, 1/b normal, a = +0 or |a| >= 2^-969, a/b normal and not near
// overflow), the same sequence without those steps gives the same correctly
// rounded quotient -- and the reciprocal half can be shared by divisions with
// one denominator.  recip_nr is that half, div_nr the rest.  Callers use them
// only on operands they have shown to be in range (the sweeps check the
// magnitude of every split flux they stage); nuc3d_selftest_division compares
// the pair with a/b on the device.
__device__ __forceinline__ double recip_nr(double b)
{
    double y = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(-b, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-b, y, 1.0);
    return __builtin_fma(y, e, y);
}
__device__ __forceinline__ double div_nr(double a, double b, double y)
{
    const double q = a * y;
    const double r = __builtin_fma(-b, q, a);
    return __builtin_fma(r, y, q);
}
// a / b for a = +-0 as well: the IEEE quotient of a zero is a zero with the
// sign of a times that of b, which a * y has (the reduced sequence alone
// returns +0 for a = -0)
__device__ __forceinline__ double div_nrz(double a, double b, double y)
{
    const double q = div_nr(a, b, y);
    return a == 0.0 ? a * y : q;
}
// a / b for b > 0 and a = +-0 as well, with no extra instruction: the
// correction of div_nr written as -fma(-r, y, -q), which rounds to the same
// value for every a != 0 (round to nearest is symmetric) and, for a zero a
// over a positive b, yields the IEEE zero of a's sign where div_nr's
// fma(r, y, q) returns +0 for a = -0 (for b < 0 and a = -0 it would return
// -0, not the IEEE +0: callers divide by positive b only)
__device__ __forceinline__ double div_nrs(double a, double b, double y)
{
    const double q = a * y;
    const double r = __builtin_fma(-b, q, a);
    return -__builtin_fma(-r, y, -q);
}
// sqrt(x) for x normal, at least 2^-767 and finite: the compiler's FP64
// square root on gfx950 (v_rsq_f64 and the Goldschmidt-style refinement of
// the LLVM AMDGPU lowering) without its steps for tiny arguments (scaled by
// 2^256 before and 2^-128 after) and for 0 and infinity (passed through),
// which leave such an x unchanged: the same instructions on the same values,
// so the same result.  nuc3d_selftest_division compares it with sqrt.
__device__ __forceinline__ double sqrt_nr(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = y * 0.5;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    return __builtin_fma(d, h, g);
}

// Operands of divisions that share a reciprocal.  A positive denominator in
// [2^-200, 2^200] (rd_denominator); numerators zero or of magnitude in
// [2^-600, 2^600], such as p (rd_operand: 0 or in [2^-400, 2^400]) and its
// products with terms that are 0 or in [2^-200, 2^200].  The quotient is then
// zero or normal and far from overflow, and every numerator is far above
// 2^-969.
__device__ __forceinline__ bool rd_operand(double v)
{
    const double a = fabs(v);
    return v == 0.0 || (a >= 0x1p-400 && a <= 0x1p400);
}
__device__ __forceinline__ bool rd_denominator(double v) { return v >= 0x1p-200 && v <= 0x1p200; }
// a metric term multiplying p in such a numerator: 0 or of magnitude in [2^-200, 2^200]
__device__ __forceinline__ bool rd_factor(double v)
{
    const double a = fabs(v);
    return v == 0.0 || (a >= 0x1p-200 && a <= 0x1p200);
}
// a / b, by the reduced sequence (RD) or the compiler's IEEE division
template <bool RD>
__device__ __forceinline__ double ddiv(double a, double b)
{
    if constexpr (RD) return div_nr(a, b, recip_nr(b));
    else return a / b;
}
// Split-flux values for which every division of the WENO5 weights is in the
// range of the reduced sequence: 0, or 2^-100 <= |v| <= 2^60.  Then every
// smoothness indicator is 0 or in [2^-306, 2^127], every weight denominator
// (eps + beta)^2 in [1e-12, 2^254], every WENO5-Z ratio tau / (beta + eps)
// +0 or in [2^-485, 2^147], every unnormalised weight in [2^-256, 2^295] and
// every normalised one in [2^-300, 1].  The fast variant's one-division form
// divides sum(w q) (+-0, or of magnitude between 2^-500 and 2^830) by sum(w)
// (in [2^-84, 2^770]; WENO5-Z included).
__device__ __forceinline__ bool rd_range(double v)
{
    const double a = fabs(v);
    return v == 0.0 || (a >= 0x1p-100 && a <= 0x1p60);
}
// The exact WENO5-JS weights need only the upper bound: their divisors are
// (eps + beta)^2 >= 1e-12 and at most 2^254, their numerators the constants
// C_k and the unnormalised weights (at least C_k / 2^254), so for any finite
// split flux of magnitude at most 2^60 (a NaN fails the test) every quotient
// is normal with a normal numerator, however small the fluxes.  WENO5-Z's
// ratio tau / (beta + eps) and the fast variant's one-division form have
// numerators that small fluxes make small, and keep rd_range.
template <int SCHEME>
__device__ __forceinline__ bool rd_staged(double v)
{
    if constexpr (NUC3D_EXACT && (SCHEME == SCH_WENO5JS || SCHEME == SCH_HYBRID)) return fabs(v) <= 0x1p60;
    else return rd_range(v);
}

// ---- smoothness indicators on an oriented stencil f0..f4 ----
// zq: the +inner beta2 of weno5z.cpp:80 ((f2-2f3+f3)^2 in place of (f2-2f3+f4)^2)
#if NUC3D_EXACT
#define NUC3D_WENO_FP
#else
#define NUC3D_WENO_FP _Pragma("clang fp contract(fast)")
#endif

// Smoothness indicators (weno5js.cpp:75-82); zq: the + inner beta2 of
// weno5z.cpp:80.  smooth5_ref keeps the reference's rounding in every build.
__device__ __forceinline__ void smooth5_ref(double f0, double f1, double f2, double f3, double f4, bool zq,
                                            double &b0, double &b1, double &b2)
{
    b0 = G0 * sq(f0 - 2.0 * f1 + f2) + G1 * sq(f0 - 4.0 * f1 + 3.0 * f2);
    b1 = G0 * sq(f1 - 2.0 * f2 + f3) + G1 * sq(f1 - f3);
    b2 = G0 * sq(zq ? (f2 - 2.0 * f3 + f3) : (f2 - 2.0 * f3 + f4)) + G1 * sq(3.0 * f2 - 4.0 * f3 + f4);
}

__device__ __forceinline__ void smooth5(double f0, double f1, double f2, double f3, double f4, bool zq,
                                        double &b0, double &b1, double &b2)
{
    NUC3D_WENO_FP
    b0 = G0 * sq(f0 - 2.0 * f1 + f2) + G1 * sq(f0 - 4.0 * f1 + 3.0 * f2);
    b1 = G0 * sq(f1 - 2.0 * f2 + f3) + G1 * sq(f1 - f3);
    b2 = G0 * sq(zq ? (f2 - 2.0 * f3 + f3) : (f2 - 2.0 * f3 + f4)) + G1 * sq(3.0 * f2 - 4.0 * f3 + f4);
}

// The same indicators with the exact multiples 2 f and 4 f folded into an
// FMA: fma(-2, f1, f0) rounds f0 - 2 f1 once, as f0 - (2.0 * f1) does when
// 2.0 * f1 is exact, i.e. for every finite product (the reduced-division
// path, whose staged values are at most 2^60).  Five instructions fewer per
// reconstruction, same bits.
__device__ __forceinline__ void smooth5_fma(double f0, double f1, double f2, double f3, double f4, bool zq,
                                            double &b0, double &b1, double &b2)
{
    const double t2 = 3.0 * f2;
    b0 = G0 * sq(__builtin_fma(-2.0, f1, f0) + f2) + G1 * sq(__builtin_fma(-4.0, f1, f0) + t2);
    b1 = G0 * sq(__builtin_fma(-2.0, f2, f1) + f3) + G1 * sq(f1 - f3);
    b2 = G0 * sq(__builtin_fma(-2.0, f3, f2) + (zq ? f3 : f4)) + G1 * sq(__builtin_fma(-4.0, f3, t2) + f4);
}

__device__ __forceinline__ void cand5(double f0, double f1, double f2, double f3, double f4, double &q0, double &q1,
                                      double &q2)
{
    NUC3D_WENO_FP
    q0 = A00 * f0 + A01 * f1 + A02 * f2;
    q1 = A10 * f1 + A11 * f2 + A12 * f3;
    q2 = A20 * f2 + A21 * f3 + A22 * f4;
}

// WENO5-JS value at the face of an oriented 5-point stencil.  RD: the
// divisions by the reduced sequence (same quotients; operands in range).
template <bool RD = false>
__device__ __forceinline__ double weno5js(double f0, double f1, double f2, double f3, double f4)
{
    NUC3D_WENO_FP
    double b0, b1, b2, q0, q1, q2;
    if (NUC3D_EXACT && RD) smooth5_fma(f0, f1, f2, f3, f4, false, b0, b1, b2);
    else smooth5(f0, f1, f2, f3, f4, false, b0, b1, b2);
    cand5(f0, f1, f2, f3, f4, q0, q1, q2);
#if NUC3D_EXACT
    double a0 = ddiv<RD>(C0, sq(EPS + b0));
    double a1 = ddiv<RD>(C1, sq(EPS + b1));
    double a2 = ddiv<RD>(C2, sq(EPS + b2));
    double as = a0 + a1 + a2;
    double o0, o1, o2;
    if constexpr (RD) {
        const double y = recip_nr(as);   // one reciprocal for the three weights
        o0 = div_nr(a0, as, y);
        o1 = div_nr(a1, as, y);
        o2 = div_nr(a2, as, y);
    } else {
        o0 = a0 / as;
        o1 = a1 / as;
        o2 = a2 / as;
    }
    return o0 * q0 + o1 * q1 + o2 * q2;
#else
    // alpha_k = c_k/d_k with d_k=(eps+beta_k)^2; multiplying numerator and
    // denominator of sum(alpha q)/sum(alpha) by d0 d1 d2 leaves one division.
    double d0 = sq(EPS + b0), d1 = sq(EPS + b1), d2 = sq(EPS + b2);
    double w0 = C0 * (d1 * d2), w1 = C1 * (d0 * d2), w2 = C2 * (d0 * d1);
    const double num = w0 * q0 + w1 * q1 + w2 * q2, den = w0 + w1 + w2;
    if constexpr (RD) return div_nrz(num, den, recip_nr(den));
    else return num / den;
#endif
}

// WENO5-Z value (weno5z.cpp:96-109).
template <bool RD = false>
__device__ __forceinline__ double weno5z(double f0, double f1, double f2, double f3, double f4, bool zq)
{
    NUC3D_WENO_FP
    double b0, b1, b2, q0, q1, q2;
    if (NUC3D_EXACT && RD) smooth5_fma(f0, f1, f2, f3, f4, zq, b0, b1, b2);
    else smooth5(f0, f1, f2, f3, f4, zq, b0, b1, b2);
    cand5(f0, f1, f2, f3, f4, q0, q1, q2);
    double tau = fabs(b2 - b0);
#if NUC3D_EXACT
    double a0 = C0 * (1.0 + sq(ddiv<RD>(tau, b0 + EPS)));
    double a1 = C1 * (1.0 + sq(ddiv<RD>(tau, b1 + EPS)));
    double a2 = C2 * (1.0 + sq(ddiv<RD>(tau, b2 + EPS)));
    double as = a0 + a1 + a2;
    double o0, o1, o2;
    if constexpr (RD) {
        const double y = recip_nr(as);
        o0 = div_nr(a0, as, y);
        o1 = div_nr(a1, as, y);
        o2 = div_nr(a2, as, y);
    } else {
        o0 = a0 / as;
        o1 = a1 / as;
        o2 = a2 / as;
    }
    return o0 * q0 + o1 * q1 + o2 * q2;
#else
    double d0 = sq(b0 + EPS), d1 = sq(b1 + EPS), d2 = sq(b2 + EPS), t2 = tau * tau;
    double w0 = C0 * (d0 + t2) * (d1 * d2), w1 = C1 * (d1 + t2) * (d0 * d2), w2 = C2 * (d2 + t2) * (d0 * d1);
    const double num = w0 * q0 + w1 * q1 + w2 * q2, den = w0 + w1 + w2;
    if constexpr (RD) return div_nrz(num, den, recip_nr(den));
    else return num / den;
#endif
}

// third-order closure (reconstructionBoundaryScheme.cpp:327-343)
__device__ __forceinline__ double third_order(double f0, double f1, double f2)
{
    double is0 = sq(f0 - f1), is1 = sq(f1 - f2);
    double q20 = -0.5 * f0 + 1.5 * f1;
    double q21 = 0.5 * f1 + 0.5 * f2;
    double aa0 = 1.0 / 3.0 / sq(1.0e-6 + is0);
    double aa1 = 2.0 / 3.0 / sq(1.0e-6 + is1);
    double w0 = aa0 / (aa0 + aa1);
    double w1 = aa1 / (aa0 + aa1);
    return w0 * q20 + w1 * q21;
}

// Eighth-order central flux of one split part at face m from cells m-4..m+3
// (s[-1..6]): the face point value of the degree-7 polynomial with the eight
// cell averages, (533 (f3+f4) - 139 (f2+f5) + 29 (f1+f6) - 3 (f0+f7)) / 840,
// the division taken as a product with the FP64 rounding of 1/840.
// No counterpart in the reference: the smooth-region branch of the hybrid
// reconstruction.
__device__ __forceinline__ double central8(const double *s)
{
    return ((533.0 * (s[2] + s[3]) - 139.0 * (s[1] + s[4])) + (29.0 * (s[0] + s[5]) - 3.0 * (s[-1] + s[6]))) *
           (1.0 / 840.0);
}

// Interior face value for the + (upwind>0) or - stencil.  s[0..5] holds the
// split flux at cells m-3..m+2 of face m (between cells m-1 and m); the
// central scheme also reads s[-1] and s[6] (cells m-4 and m+3).
template <int SCHEME, bool RD = false>
__device__ __forceinline__ double face_interior(const double *s, bool plus, bool zq)
{
    if (SCHEME == SCH_CENTRAL8) return central8(s);
    if (plus) {
        if (SCHEME == SCH_WENO5Z) return weno5z<RD>(s[0], s[1], s[2], s[3], s[4], zq);
        return weno5js<RD>(s[0], s[1], s[2], s[3], s[4]);
    }
    if (SCHEME == SCH_WENO5Z) return weno5z<RD>(s[5], s[4], s[3], s[2], s[1], false);
    return weno5js<RD>(s[5], s[4], s[3], s[2], s[1]);
}

// One inner row of the CRWENO5 system (crweno5.cpp:130-160 for +,
// :310-340 for -): f points at cell i of face i+1 (element stride st); plus
// uses cells i-2..i+2, minus cells i+3..i-1.  a, b, c are the coefficients of
// faces i, i+1, i+2 (in line order), d the right-hand side.
__device__ __forceinline__ void crweno_row(const double *f, int st, bool plus, double &a, double &b, double &c,
                                           double &d)
{
    // row coefficients in the reference order in both variants: the tridiagonal
    // solve amplifies their rounding
    // f points at cell i of face i+1; plus: cells i-2..i+2, minus: i+3..i-1
    double f0, f1, f2, f3, f4;
    if (plus) { f0 = f[-2 * st]; f1 = f[-st]; f2 = f[0]; f3 = f[st]; f4 = f[2 * st]; }
    else { f0 = f[3 * st]; f1 = f[2 * st]; f2 = f[st]; f3 = f[0]; f4 = f[-st]; }
    double b0, b1, b2;
    smooth5_ref(f0, f1, f2, f3, f4, false, b0, b1, b2);
    const double a0 = R0 / sq(EPS + b0), a1 = R1 / sq(EPS + b1), a2 = R2 / sq(EPS + b2);
    const double as = a0 + a1 + a2;
    const double o0 = a0 / as, o1 = a1 / as, o2 = a2 / as;
    const double lo = 2.0 / 3.0 * o0 + 1.0 / 3.0 * o1;   // coefficient of the upwind-side face
    const double hi = 1.0 / 3.0 * o2;                    // coefficient of the downwind-side face
    b = 1.0 / 3.0 * o0 + 2.0 / 3.0 * (o1 + o2);
    if (plus) { a = lo; c = hi; }
    else { c = lo; a = hi; }
    d = 1.0 / 6.0 * o0 * f1 + (5.0 * (o0 + o1) + o2) * f2 / 6.0 + (o1 + 5.0 * o2) * f3 / 6.0;
}

// Exterior (physical face) closure at face m, k faces from the physical
// face (k = m on the low side, k = n - m on the high side).  s[0..5] holds the
// split flux at cells m-3..m+2, where the slot of ghost cell -1 (+ flux) or n
// (- flux) carries the BC ghost flux.  Mapped onto s, the left and right
// closures of reconstructionBoundaryScheme.cpp:52-312 coincide:
//   +: k=0 s2 (ghost[-1] / c[n-1]); k=1 (s2+s3)/2; k=2 third(s1,s2,s3); k=3 WENO5(s0..s4)
//   -: k=0 s3 (c[0] / ghost[n]);    k=1 (s3+s2)/2; k=2 third(s4,s3,s2); k=3 WENO5(s5..s1)
__device__ __forceinline__ double face_exterior(const double *s, bool plus, int k)
{
    if (plus) {
        switch (k) {
        case 0: return s[2];
        case 1: return 0.5 * (s[2] + s[3]);
        case 2: return third_order(s[1], s[2], s[3]);
        default: return weno5js(s[0], s[1], s[2], s[3], s[4]);
        }
    }
    switch (k) {
    case 0: return s[3];
    case 1: return 0.5 * (s[3] + s[2]);
    case 2: return third_order(s[4], s[3], s[2]);
    default: return weno5js(s[5], s[4], s[3], s[2], s[1]);
    }
}

// ---- physics ----
struct PhysParams {
    double gamma, mach, reynolds, prandtl, sref, twall;
    int riemann;
    // the state read is in the range of state_in_range: the divisions by rho
    // share one reciprocal (same quotients)
    int rdq = 0;
};

// A state whose divisions by rho = Q0 J may share one reciprocal (recip_nr):
// rho in [2^-200, 2^200], the momenta Q1..Q3 (conserved / J) each zero or of
// magnitude in [2^-400, 2^400], and gamma p of magnitude in [2^-400, 2^400]
// (gamma_p_in_range).  With gamma in [1, 8] and Mach numbers in [2^-40, 2^40]
// the temperature's numerator gamma Ma^2 p is then within [2^-500, 2^500];
// every quotient is normal and every numerator far above 2^-969, and c^2 =
// |gamma p / rho| lies in [2^-600, 2^600], where sqrt_nr is the square root.
__device__ __forceinline__ bool momenta_in_range(const double *Q, double rho)
{
    return rd_denominator(rho) && rd_operand(Q[1]) && rd_operand(Q[2]) && rd_operand(Q[3]);
}
__device__ __forceinline__ bool gamma_p_in_range(double gp) { return gp != 0.0 && rd_operand(gp); }

// con2prim (physicsModel.cpp:313-353).  Returns p; sets rho,u,v,w,c.
// P.rdq: the four divisions by rho (the velocities and c^2) through one
// reciprocal, and c by sqrt_nr; yr (may be null) receives the reciprocal for
// further divisions by rho.
__device__ __forceinline__ void con2prim(const double *Q, double jac, const PhysParams &P, double &rho, double &u,
                                         double &v, double &w, double &p, double &c, double *yr = nullptr)
{
    rho = Q[0] * jac;
    if (P.rdq) {
        const double y = recip_nr(rho);
        u = div_nrs(Q[1], rho, y) * jac;
        v = div_nrs(Q[2], rho, y) * jac;
        w = div_nrs(Q[3], rho, y) * jac;
        double E = Q[4] * jac;
        p = (E - 0.5 * rho * (u * u + v * v + w * w)) * (P.gamma - 1.0);
        c = sqrt_nr(fabs(div_nrs(P.gamma * p, rho, y)));
        if (yr) *yr = y;
    } else {
        u = Q[1] / rho * jac;
        v = Q[2] / rho * jac;
        w = Q[3] / rho * jac;
        double E = Q[4] * jac;
        p = (E - 0.5 * rho * (u * u + v * v + w * w)) * (P.gamma - 1.0);
        c = sqrt(fabs(P.gamma * p / rho));
    }
}

// Per-thread counts of a stage: e<0 fix-ups, and states that stay unusable
// after the fix-up (density not positive, pressure negative; NaN counts as
// both).  flush() adds them to counters[0..2] with one atomic each.
struct StageCounts {
    unsigned long long fix = 0;
    unsigned negd = 0, negp = 0;
    __device__ __forceinline__ void flush(unsigned long long *counters) const
    {
        if (fix) atomicAdd(counters, fix);
        if (negd) atomicAdd(counters + 1, (unsigned long long)negd);
        if (negp) atomicAdd(counters + 2, (unsigned long long)negp);
    }
};

// e<0 fix-up of physicsModel.cpp:322-332 applied to a freshly written state,
// then the positivity check the reference makes next: con2prim's pressure
// (E - rho|u|^2/2)(gamma - 1) must not be negative (physicsModel.cpp:1009-1016,
// where it exits).  Counts the fix-up and a non-positive density or negative
// pressure (either one NaN) in cnt.  Returns true if Q[4] was replaced.
// rd: the state passed momenta_in_range (the divisions by rho share one
// reciprocal).  pr (may be null) receives the pressure con2prim computes from
// the final state, the one every later reader of it sees.
__device__ __forceinline__ bool fixup_energy(double *Q, double jac, const PhysParams &P, StageCounts &cnt,
                                             bool rd = false, double *pr = nullptr)
{
    double rho = Q[0] * jac;
    double u, v, w;
    if (rd) {
        const double y = recip_nr(rho);
        u = div_nrs(Q[1], rho, y) * jac;
        v = div_nrs(Q[2], rho, y) * jac;
        w = div_nrs(Q[3], rho, y) * jac;
    } else {
        u = Q[1] / rho * jac;
        v = Q[2] / rho * jac;
        w = Q[3] / rho * jac;
    }
    double E = Q[4] * jac;
    const double ke = 0.5 * rho * (u * u + v * v + w * w);
    double e0 = E - ke;
    bool fixed = false;
    if (pr) *pr = e0 * (P.gamma - 1.0);
    if (e0 < 0.0) {
        double g = P.gamma, m = P.mach;
        E = (pow(rho, g) / ((g * m * m) * (g - 1.0)) + 0.5 * rho * (u * u + v * v + w * w));
        Q[4] = E / jac;
        e0 = E - ke;
        if (pr) *pr = (Q[4] * jac - ke) * (P.gamma - 1.0);
        cnt.fix++;
        fixed = true;
    }
    if (!(rho > 0.0)) cnt.negd++;
    if (!(e0 >= 0.0)) cnt.negp++;
    return fixed;
}

__device__ __forceinline__ double machL(double m) { return fabs(m) < 1.0 ? 0.25 * sq(m + 1.0) + 0.125 * sq(m * m - 1) : 0.50 * (m + fabs(m)); }
__device__ __forceinline__ double machR(double m) { return fabs(m) < 1.0 ? -0.25 * sq(m - 1.0) - 0.125 * sq(m * m - 1) : 0.50 * (m - fabs(m)); }
__device__ __forceinline__ double presL(double m, double p, bool plus_variant)
{
    if (fabs(m) < 1.0)
        return plus_variant ? p * (0.25 * sq(m + 1.0) * (2.0 - m) + 0.1875 * m * sq(m * m - 1.0)) : p * (m + 1.0) / 2.0;
    return 0.50 * p * (m + fabs(m)) / m;
}
__device__ __forceinline__ double presR(double m, double p, bool plus_variant)
{
    if (fabs(m) < 1.0)
        return plus_variant ? p * (0.25 * sq(m - 1.0) * (2.0 + m) - 0.1875 * m * sq(m * m - 1.0)) : -p * (m - 1.0) / 2.0;
    return 0.5 * p * (m - fabs(m)) / m;
}

// Split flux of one cell in one direction (RiemannSolver loop body,
// physicsModel.cpp:557-597 with RiemannLF/AUSM/AUSMp).  Q = conserved/J,
// theta = |grad xi| = sqrt(mx^2 + my^2 + mz^2) (a uniform block's constant).
// yj != 0: the reciprocal half of the division by J (recip_nr(jac)), which the
// Lax-Friedrichs divisions m p / J and p / J share; the caller has checked
// their operands (J and the metric terms per block, p per cell).
// Returns |U| + c|grad xi| (the local eigenvalue of :597).
__device__ __forceinline__ double split_cell(const double *Q, double jac, double mx, double my, double mz,
                                             double theta, double u, double v, double w, double p, double c,
                                             int riemann, double *fp, double *fn, double yj = 0.0)
{
    double U0 = (mx * u + my * v + mz * w);
    double al = c * theta;
    double lam = fabs(U0) + al;
    if (riemann == RIE_LF) {
        double px, py, pz, pj;
        if (yj != 0.0) {
            px = div_nrs(mx * p, jac, yj);
            py = div_nrs(my * p, jac, yj);
            pz = div_nrs(mz * p, jac, yj);
            pj = div_nrs(p, jac, yj);
        } else {
            px = mx * p / jac;
            py = my * p / jac;
            pz = mz * p / jac;
            pj = p / jac;
        }
        double f0 = U0 * Q[0];
        double f1 = U0 * Q[1] + px;
        double f2 = U0 * Q[2] + py;
        double f3 = U0 * Q[3] + pz;
        double f4 = U0 * (Q[4] + pj);
        fp[0] = 0.5 * (f0 + lam * Q[0]);
        fp[1] = 0.5 * (f1 + lam * Q[1]);
        fp[2] = 0.5 * (f2 + lam * Q[2]);
        fp[3] = 0.5 * (f3 + lam * Q[3]);
        fp[4] = 0.5 * (f4 + lam * Q[4]);
        fn[0] = 0.5 * (f0 - lam * Q[0]);
        fn[1] = 0.5 * (f1 - lam * Q[1]);
        fn[2] = 0.5 * (f2 - lam * Q[2]);
        fn[3] = 0.5 * (f3 - lam * Q[3]);
        fn[4] = 0.5 * (f4 - lam * Q[4]);
        return lam;
    }
    bool pv = riemann == RIE_AUSMP;
    double mach = U0 / al;
    double mp = machL(mach), mn = machR(mach);
    double pp = presL(mach, p, pv), pn = presR(mach, p, pv);
    fp[0] = mp * al * Q[0];
    fp[1] = mp * al * Q[1] + mx * pp / jac;
    fp[2] = mp * al * Q[2] + my * pp / jac;
    fp[3] = mp * al * Q[3] + mz * pp / jac;
    fp[4] = mp * al * Q[4] + mp * al * p / jac;
    fn[0] = mn * al * Q[0];
    fn[1] = mn * al * Q[1] + mx * pn / jac;
    fn[2] = mn * al * Q[2] + my * pn / jac;
    fn[3] = mn * al * Q[3] + mz * pn / jac;
    fn[4] = mn * al * Q[4] + mn * al * p / jac;
    return lam;
}

// Point splitter of a ghost primitive state (physicsModel.cpp:1145-1398).
__device__ __forceinline__ void split_point(const double *prim, double jac, double mx, double my, double mz,
                                            const PhysParams &P, double *fl, double *fr)
{
    double rho = prim[0], u = prim[1], v = prim[2], w = prim[3], p = prim[4];
    double g = P.gamma;
    double _rho = rho, _rhou = rho * u, _rhov = rho * v, _rhow = rho * w;
    double _E = p / (g - 1.0) + 0.5 * rho * (u * u + v * v + w * w);
    double _p = (_E - 0.5 * rho * (u * u + v * v + w * w)) * (g - 1.0);
    double _alpha = sqrt(fabs(g * _p / rho));
    double theta = sqrt(mx * mx + my * my + mz * mz);
    double alpha = _alpha * theta;
    double U0 = (mx * u + my * v + mz * w);
    if (P.riemann == RIE_LF) {
        double lam = fabs(U0) + alpha;
        double f0 = U0 * _rho / jac;
        double f1 = U0 * _rhou / jac + mx * p / jac;
        double f2 = U0 * _rhov / jac + my * p / jac;
        double f3 = U0 * _rhow / jac + mz * p / jac;
        double f4 = U0 * (_E / jac + p / jac);
        fl[0] = 0.5 * (f0 + lam * _rho / jac);
        fl[1] = 0.5 * (f1 + lam * _rhou / jac);
        fl[2] = 0.5 * (f2 + lam * _rhov / jac);
        fl[3] = 0.5 * (f3 + lam * _rhow / jac);
        fl[4] = 0.5 * (f4 + lam * _E / jac);
        fr[0] = 0.5 * (f0 - lam * _rho / jac);
        fr[1] = 0.5 * (f1 - lam * _rhou / jac);
        fr[2] = 0.5 * (f2 - lam * _rhov / jac);
        fr[3] = 0.5 * (f3 - lam * _rhow / jac);
        fr[4] = 0.5 * (f4 - lam * _E / jac);
        return;
    }
    bool pv = P.riemann == RIE_AUSMP;
    double mach = U0 / alpha;
    double mp = machL(mach), mn = machR(mach);
    double pp = presL(mach, p, pv), pn = presR(mach, p, pv);
    fl[0] = mp * alpha * _rho / jac;
    fl[1] = mp * alpha * _rhou / jac + mx * pp / jac;
    fl[2] = mp * alpha * _rhov / jac + my * pp / jac;
    fl[3] = mp * alpha * _rhow / jac + mz * pp / jac;
    fl[4] = mp * alpha * _E / jac + mp * alpha * p / jac;
    fr[0] = mn * alpha * _rho / jac;
    fr[1] = mn * alpha * _rhou / jac + mx * pn / jac;
    fr[2] = mn * alpha * _rhov / jac + my * pn / jac;
    fr[3] = mn * alpha * _rhow / jac + mz * pn / jac;
    fr[4] = mn * alpha * _E / jac + mn * alpha * p / jac;
}

// Central differences (centraldifference6th.cpp:43-45, centraldifference8th.cpp:49-52);
// s points at the centre cell, st = element stride along the line (T: double
// for LDS and registers, gdouble for global memory).
template <typename T>
__device__ __forceinline__ double cd6(const T *s, long st)
{
    return CD6_2 * (s[3 * st] - s[-3 * st]) + CD6_1 * (s[2 * st] - s[-2 * st]) + CD6_0 * (s[st] - s[-st]);
}
template <typename T>
__device__ __forceinline__ double cd8(const T *s, long st)
{
    return CD8_3 * (s[4 * st] - s[-4 * st]) + CD8_2 * (s[3 * st] - s[-3 * st]) + CD8_1 * (s[2 * st] - s[-2 * st]) +
           CD8_0 * (s[st] - s[-st]);
}
// Exterior derivative closure (boundaryScheme.cpp:100-126); k = distance from
// the physical face, left = the face is on the low side.
template <typename T>
__device__ __forceinline__ double diff_exterior(const T *s, long st, int k, bool left)
{
    switch (k) {
    case 0: return left ? s[st] - s[0] : s[0] - s[-st];
    case 1: return 0.5 * (s[st] - s[-st]);
    case 2: return CD4_1 * (s[2 * st] - s[-2 * st]) + CD4_0 * (s[st] - s[-st]);
    default: return CD6_2 * (s[3 * st] - s[-3 * st]) + CD6_1 * (s[2 * st] - s[-2 * st]) + CD4_0 * (s[st] - s[-st]);
    }
}

}  // namespace nuc3d_dev
