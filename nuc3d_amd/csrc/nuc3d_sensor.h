// nuc3d_sensor.h -- the learned shock sensor of the hybrid reconstruction
// (device side; no counterpart in the reference).
//
// Per interface the hybrid takes the eighth-order central face value in
// smooth flow and WENO5-JS elsewhere.  A small MLP decides:
//   input   30 features: rho, u, v, w, p on the six cells m-3..m+2 around
//           face m, each variable centred on its stencil mean and divided by
//           (stencil range + 0.01 S), with the acoustic scales S of
//           feature_scales; computed in FP32 from the FP32 roundings of the
//           primitives;
//   layers  30 -> 32 -> 32 -> 1, ReLU, BF16 weights and activations, FP32
//           accumulation on the matrix cores (v_mfma_f32_16x16x32_bf16);
//   output  a logit; WENO5-JS when it reaches the threshold, when the stencil
//           has no positive mean density or pressure, or when the central
//           face value of rho or p is not positive (the positivity fallback).
// The training label (the teacher) is whether the central and WENO5-JS face
// values of any of the five primitives differ by more than tau S.
//
// MFMA operand layout of v_mfma_f32_16x16x32_bf16 for lane l (g = l/16):
//   A (16 x 32): row l%16, k = 8g..8g+7;  B (32 x 16): column l%16, k = 8g..8g+7;
//   C (16 x 16): column l%16, rows 4g..4g+3.
// The network runs transposed (hidden x samples), so the layer-1 output that a
// lane holds (hidden units 4g..4g+3 and 16+4g..16+4g+3 of sample l%16) is
// directly its B operand of layer 2 under the K permutation
//   k = 8g+e  <->  hidden unit e < 4 ? 4g+e : 16+4g+(e-4),
// which the layer-2 weights are loaded with.
#pragma once
#include <hip/hip_runtime.h>

#include "nuc3d_types.h"

namespace nuc3d_sensor {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int NFEAT = NUC3D_SENSOR_NFEAT;   // 30
constexpr int KIN = 32;                     // features padded to one MFMA K step
constexpr int NH = NUC3D_SENSOR_NH;         // 32
constexpr float FEAT_EPS = 0.01f;

// float -> bf16 bits, round to nearest even (finite inputs)
__device__ __forceinline__ unsigned short bf16_bits(float x)
{
    unsigned u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_float(unsigned short b) { return __uint_as_float((unsigned)b << 16); }

// two floats -> two bf16 (round to nearest even, v_cvt_pk_bf16_f32), a in the low half
__device__ __forceinline__ unsigned pack_bf16(float a, float b)
{
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    const bf16x2 v = {(__bf16)a, (__bf16)b};
    return __builtin_bit_cast(unsigned, v);
}

// WENO5-JS face value in the reference's operation order (weno5js.cpp:75-108)
// whatever the kernel variant: the teacher's labels must not depend on it.
__device__ __forceinline__ double weno5js_ref(double f0, double f1, double f2, double f3, double f4)
{
    const double G0 = 13.0 / 12.0, G1 = 1.0 / 4.0, EPS = 1.0e-6;
    double t;
    t = f0 - 2.0 * f1 + f2;
    double u = f0 - 4.0 * f1 + 3.0 * f2;
    const double b0 = G0 * (t * t) + G1 * (u * u);
    t = f1 - 2.0 * f2 + f3;
    u = f1 - f3;
    const double b1 = G0 * (t * t) + G1 * (u * u);
    t = f2 - 2.0 * f3 + f4;
    u = 3.0 * f2 - 4.0 * f3 + f4;
    const double b2 = G0 * (t * t) + G1 * (u * u);
    const double q0 = 1.0 / 3.0 * f0 + -7.0 / 6.0 * f1 + 11.0 / 6.0 * f2;
    const double q1 = -1.0 / 6.0 * f1 + 5.0 / 6.0 * f2 + 1.0 / 3.0 * f3;
    const double q2 = 1.0 / 3.0 * f2 + 5.0 / 6.0 * f3 + -1.0 / 6.0 * f4;
    double e = EPS + b0;
    const double a0 = 0.1 / (e * e);
    e = EPS + b1;
    const double a1 = 0.6 / (e * e);
    e = EPS + b2;
    const double a2 = 0.3 / (e * e);
    const double as = a0 + a1 + a2;
    const double o0 = a0 / as, o1 = a1 / as, o2 = a2 / as;
    return o0 * q0 + o1 * q1 + o2 * q2;
}

// eighth-order central face value from cells m-4..m+3
__device__ __forceinline__ double central8_ref(double c0, double c1, double c2, double c3, double c4, double c5,
                                               double c6, double c7)
{
    return ((533.0 * (c3 + c4) - 139.0 * (c2 + c5)) + (29.0 * (c1 + c6) - 3.0 * (c0 + c7))) * (1.0 / 840.0);
}

// The same in FP32 (the positivity check of the central face values of rho and p)
__device__ __forceinline__ float central8f(float c0, float c1, float c2, float c3, float c4, float c5, float c6,
                                           float c7)
{
    return ((533.0f * (c3 + c4) - 139.0f * (c2 + c5)) + (29.0f * (c1 + c6) - 3.0f * (c0 + c7))) * (1.0f / 840.0f);
}

// ---- features, FP32 from the FP32 roundings of the FP64 primitives ----
// getf(v, j) = (float) primitive v (rho,u,v,w,p) at cell m-3+j, j = 0..5.
template <class Get>
__device__ __forceinline__ float stencil_mean(const Get &getf, int v)
{
    float s = getf(v, 0);
#pragma unroll
    for (int j = 1; j < 6; j++) s = s + getf(v, j);
    return s * (1.0f / 6.0f);
}

// Scales S of the five variables from the stencil's means and a velocity
// scale U = (largest |velocity component| over the stencil) + 0.1 c, with c
// the sound speed of the means: U for the velocities, rho c U for the
// pressure and rho U / c for the density (the acoustic relations
// dp ~ rho c du, drho ~ rho du / c, so that a tolerance tau S is the same
// disturbance in every variable, and low-Mach flows are not measured on
// the scale of the mean pressure).  False when mean rho or mean p is not
// positive (the stencil is invalid; the face takes WENO5-JS).
template <class Get>
__device__ __forceinline__ bool feature_scales(const Get &getf, float gamma, float *S)
{
    const float rb = stencil_mean(getf, 0), pb = stencil_mean(getf, 4);
    if (!(rb > 0.0f && pb > 0.0f)) return false;
    const float cb = sqrtf(gamma * pb / rb);
    float um = 0.0f;
#pragma unroll
    for (int v = 1; v < 4; v++)
#pragma unroll
        for (int j = 0; j < 6; j++) um = fmaxf(um, fabsf(getf(v, j)));
    const float U = um + 0.1f * cb;
    S[0] = rb * U / cb;
    S[1] = U;
    S[2] = U;
    S[3] = U;
    S[4] = rb * cb * U;
    return true;
}

// mean and 1 / (range + 0.01 S) of variable v
template <class Get>
__device__ __forceinline__ void feature_norm(const Get &getf, int v, float Sv, float &mean, float &inv)
{
    mean = stencil_mean(getf, v);
    float lo = getf(v, 0), hi = lo;
#pragma unroll
    for (int j = 1; j < 6; j++) {
        lo = fminf(lo, getf(v, j));
        hi = fmaxf(hi, getf(v, j));
    }
    inv = 1.0f / ((hi - lo) + FEAT_EPS * Sv);
}

// All 30 features as bf16 bits, packed two per word (feature 2i in the low
// half of word i; word 15 is the zero padding of features 30, 31).  Feature
// 6v+j = bf16((P_v(j) - mean_v) / (range_v + 0.01 S_v)), the division taken
// as a product with the reciprocal.  Returns false (and zero features) for an
// invalid stencil.
template <class Get>
__device__ __forceinline__ bool features(const Get &getf, float gamma, unsigned *xw)
{
    float S[5];
    const bool ok = feature_scales(getf, gamma, S);
    float x[KIN];
#pragma unroll
    for (int v = 0; v < 5; v++) {
        float mean = 0.0f, inv = 0.0f;
        if (ok) feature_norm(getf, v, S[v], mean, inv);
#pragma unroll
        for (int j = 0; j < 6; j++) x[6 * v + j] = ok ? (getf(v, j) - mean) * inv : 0.0f;
    }
    x[30] = x[31] = 0.0f;
#pragma unroll
    for (int i = 0; i < 16; i++) xw[i] = pack_bf16(x[2 * i], x[2 * i + 1]);
    return ok;
}

// Teacher label of face m from get8(v, j) = primitive v at cell m-4+j,
// j = 0..7 (FP64): 1 when, for some primitive, the central face value differs
// from the WENO5-JS value of either upwind side by more than tau S, with the
// scales S of feature_scales in FP64 over cells m-3..m+2 (mean rho and p,
// c = sqrt(gamma mean p / mean rho), U = max |velocity component| + 0.1 c;
// S = rho U / c, U, U, U, rho c U).  -1 when mean rho or mean p is not
// positive.
template <class Get8>
__device__ __forceinline__ int teacher(const Get8 &get8, double gamma, double tau)
{
    double mean[5];
#pragma unroll
    for (int v = 0; v < 5; v += 4) {
        double sm = get8(v, 1);
#pragma unroll
        for (int j = 2; j < 7; j++) sm = sm + get8(v, j);
        mean[v] = sm / 6.0;
    }
    const double rb = mean[0], pb = mean[4];
    if (!(rb > 0.0 && pb > 0.0)) return -1;
    const double cb = sqrt(gamma * pb / rb);
    double um = 0.0;
#pragma unroll
    for (int v = 1; v < 4; v++)
#pragma unroll
        for (int j = 1; j < 7; j++) um = fmax(um, fabs(get8(v, j)));
    const double U = um + 0.1 * cb;
    const double S[5] = {rb * U / cb, U, U, U, rb * cb * U};
    double dev = 0.0;
#pragma unroll
    for (int v = 0; v < 5; v++) {
        const double c8 = central8_ref(get8(v, 0), get8(v, 1), get8(v, 2), get8(v, 3), get8(v, 4), get8(v, 5),
                                       get8(v, 6), get8(v, 7));
        const double wp = weno5js_ref(get8(v, 1), get8(v, 2), get8(v, 3), get8(v, 4), get8(v, 5));
        const double wm = weno5js_ref(get8(v, 6), get8(v, 5), get8(v, 4), get8(v, 3), get8(v, 2));
        const double d = fmax(fabs(c8 - wp), fabs(c8 - wm)) / S[v];
        dev = fmax(dev, d);
    }
    return dev > tau ? 1 : 0;
}

// Per-lane weights of the network: the A operands of the three layers and
// the accumulator initial values (the biases).
struct Regs {
    bf16x8 w1[2];      // W1 rows 16mt + l%16, inputs 8g..8g+7
    bf16x8 w2[2];      // W2 rows 16mt + l%16, hidden inputs of the permuted K slots
    bf16x8 w3;         // w3 at the permuted K slots, the same in all 16 rows
    f32x4 b1[2];       // b1 of hidden 16mt + 4g + r, r = 0..3
    f32x4 b2[2];
    float w3f[2][4];   // w3 (bf16 values) of hidden 16mt + 4g + r
    float b3;
};

// hidden unit of permuted K slot e of lane group g
__device__ __forceinline__ int kslot_unit(int g, int e) { return e < 4 ? 4 * g + e : 16 + 4 * g + (e - 4); }

__device__ __forceinline__ void load_regs(const nuc3d_sensor_dev *__restrict__ S, Regs &R)
{
    const int l = threadIdx.x & 63, g = l >> 4, i = l & 15;
    u16x8 c;
#pragma unroll
    for (int e = 0; e < 8; e++) c[e] = bf16_bits(S->w3[kslot_unit(g, e)]);
    R.w3 = __builtin_bit_cast(bf16x8, c);
#pragma unroll
    for (int mt = 0; mt < 2; mt++) {
        const unsigned short *r1 = S->w1 + (16 * mt + i) * KIN;
        const unsigned short *r2 = S->w2 + (16 * mt + i) * NH;
        u16x8 a, b;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            a[e] = r1[8 * g + e];
            b[e] = r2[kslot_unit(g, e)];
        }
        R.w1[mt] = __builtin_bit_cast(bf16x8, a);
        R.w2[mt] = __builtin_bit_cast(bf16x8, b);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            R.b1[mt][r] = S->b1[16 * mt + 4 * g + r];
            R.b2[mt][r] = S->b2[16 * mt + 4 * g + r];
            R.w3f[mt][r] = S->w3[16 * mt + 4 * g + r];
        }
    }
    R.b3 = S->b3;
}

// The inference registers from the lane-arranged copy (eight 16-byte loads per lane).
__device__ __forceinline__ void load_regs_lanes(const nuc3d_sensor_dev *__restrict__ S, Regs &R)
{
    const int l = threadIdx.x & 63;
    R.w3 = __builtin_bit_cast(bf16x8, S->lw3[l]);
#pragma unroll
    for (int mt = 0; mt < 2; mt++) {
        R.w1[mt] = __builtin_bit_cast(bf16x8, S->lw1[mt][l]);
        R.w2[mt] = __builtin_bit_cast(bf16x8, S->lw2[mt][l]);
        const float4 b1 = S->lb1[mt][l], b2 = S->lb2[mt][l];
        R.b1[mt][0] = b1.x; R.b1[mt][1] = b1.y; R.b1[mt][2] = b1.z; R.b1[mt][3] = b1.w;
        R.b2[mt][0] = b2.x; R.b2[mt][1] = b2.y; R.b2[mt][2] = b2.z; R.b2[mt][3] = b2.w;
    }
    R.b3 = S->b3;
}

// Activations of one 16-sample tile kept for the backward pass (per lane:
// sample l%16, hidden units 16mt + 4g + r).
struct Acts {
    f32x4 h1[2];       // layer-1 pre-activations
    f32x4 h2[2];       // layer-2 pre-activations
    bf16x8 a1;         // relu(h1) as bf16, in the permuted K order
    bf16x8 a2;         // relu(h2) as bf16, in the permuted K order
};

// Forward pass of one tile of 16 samples by one wavefront: x is the lane's B
// operand (features 8g..8g+7 of sample l%16).  Each layer is one MFMA
// K step with the bias as the accumulator's initial value; the output
// layer's A operand repeats w3 in all 16 rows, so every lane receives the
// logit of its sample l%16.  Between layers: ReLU, then bf16.
__device__ __forceinline__ float forward16(const Regs &R, bf16x8 x, Acts &A)
{
    A.h1[0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(R.w1[0], x, R.b1[0], 0, 0, 0);
    A.h1[1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(R.w1[1], x, R.b1[1], 0, 0, 0);
    uint4 u;
    u.x = pack_bf16(fmaxf(A.h1[0][0], 0.0f), fmaxf(A.h1[0][1], 0.0f));
    u.y = pack_bf16(fmaxf(A.h1[0][2], 0.0f), fmaxf(A.h1[0][3], 0.0f));
    u.z = pack_bf16(fmaxf(A.h1[1][0], 0.0f), fmaxf(A.h1[1][1], 0.0f));
    u.w = pack_bf16(fmaxf(A.h1[1][2], 0.0f), fmaxf(A.h1[1][3], 0.0f));
    A.a1 = __builtin_bit_cast(bf16x8, u);
    A.h2[0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(R.w2[0], A.a1, R.b2[0], 0, 0, 0);
    A.h2[1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(R.w2[1], A.a1, R.b2[1], 0, 0, 0);
    u.x = pack_bf16(fmaxf(A.h2[0][0], 0.0f), fmaxf(A.h2[0][1], 0.0f));
    u.y = pack_bf16(fmaxf(A.h2[0][2], 0.0f), fmaxf(A.h2[0][3], 0.0f));
    u.z = pack_bf16(fmaxf(A.h2[1][0], 0.0f), fmaxf(A.h2[1][1], 0.0f));
    u.w = pack_bf16(fmaxf(A.h2[1][2], 0.0f), fmaxf(A.h2[1][3], 0.0f));
    A.a2 = __builtin_bit_cast(bf16x8, u);
    const f32x4 b3 = {R.b3, R.b3, R.b3, R.b3};
    const f32x4 z = __builtin_amdgcn_mfma_f32_16x16x32_bf16(R.w3, A.a2, b3, 0, 0, 0);
    return z[0];
}

}  // namespace nuc3d_sensor
