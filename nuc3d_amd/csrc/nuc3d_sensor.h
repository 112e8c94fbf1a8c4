// nuc3d_sensor.h -- the learned shock sensor of the hybrid reconstruction
// (device side; no counterpart in the reference).
//
// Per interface the hybrid takes the eighth-order central face value in
// smooth flow and WENO5-JS elsewhere.  A small MLP decides:
//   input   30 features: rho, u, v, w, p on the six cells m-3..m+2 around
//           face m, each variable centred on its stencil mean and divided by
//           (stencil range + 0.01 S), with S = mean rho for rho, mean p for p
//           and the sound speed of the means for the velocities;
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
constexpr double FEAT_EPS = 0.01;

// float -> bf16 bits, round to nearest even (finite inputs)
__device__ __forceinline__ unsigned short bf16_bits(float x)
{
    unsigned u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_float(unsigned short b) { return __uint_as_float((unsigned)b << 16); }

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
    return ((533.0 * (c3 + c4) - 139.0 * (c2 + c5)) + (29.0 * (c1 + c6) - 3.0 * (c0 + c7))) / 840.0;
}

// Stencil statistics of face m from get(v, j) = primitive v (rho,u,v,w,p) at
// cell m-3+j, j = 0..5: per-variable means and the scales S.  Returns false
// when the mean density or pressure is not positive.
template <class Get>
__device__ __forceinline__ bool stencil_scales(const Get &get, double gamma, double *mean, double *S)
{
#pragma unroll
    for (int v = 0; v < 5; v++) {
        double s = get(v, 0);
#pragma unroll
        for (int j = 1; j < 6; j++) s = s + get(v, j);
        mean[v] = s / 6.0;
    }
    const double rb = mean[0], pb = mean[4];
    if (!(rb > 0.0 && pb > 0.0)) return false;
    const double cb = sqrt(gamma * pb / rb);
    S[0] = rb;
    S[1] = cb;
    S[2] = cb;
    S[3] = cb;
    S[4] = pb;
    return true;
}

// The 30 features as bf16 bits, packed two per word (feature 2i in the low
// half of word i; words 15 is the zero padding of features 30, 31).  Returns
// false (and zero features) for an invalid stencil.
template <class Get>
__device__ __forceinline__ bool features(const Get &get, double gamma, unsigned *xw)
{
    double mean[5], S[5];
    const bool ok = stencil_scales(get, gamma, mean, S);
    unsigned short x[KIN];
#pragma unroll
    for (int v = 0; v < 5; v++) {
        double lo = get(v, 0), hi = lo;
#pragma unroll
        for (int j = 1; j < 6; j++) {
            lo = fmin(lo, get(v, j));
            hi = fmax(hi, get(v, j));
        }
        const double den = (hi - lo) + FEAT_EPS * S[v];
#pragma unroll
        for (int j = 0; j < 6; j++) x[6 * v + j] = ok ? bf16_bits((float)((get(v, j) - mean[v]) / den)) : 0;
    }
    x[30] = x[31] = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) xw[i] = (unsigned)x[2 * i] | ((unsigned)x[2 * i + 1] << 16);
    return ok;
}

// Teacher label of face m from get8(v, j) = primitive v at cell m-4+j,
// j = 0..7: 1 when, for some primitive, the central face value differs from
// the WENO5-JS value of either upwind side by more than tau S.  -1 for an
// invalid stencil.
template <class Get8>
__device__ __forceinline__ int teacher(const Get8 &get8, double gamma, double tau)
{
    double mean[5], S[5];
    auto get6 = [&](int v, int j) { return get8(v, j + 1); };
    if (!stencil_scales(get6, gamma, mean, S)) return -1;
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

// Per-lane weights of the network (the A operands and the per-row terms).
struct Regs {
    bf16x8 w1[2];      // W1 rows 16mt + l%16, inputs 8g..8g+7
    bf16x8 w2[2];      // W2 rows 16mt + l%16, hidden inputs of the permuted K slots
    float b1[2][4];    // b1 of hidden 16mt + 4g + r
    float b2[2][4];
    float w3[2][4];    // w3 (bf16 values) of hidden 16mt + 4g + r
    float b3;
};

__device__ __forceinline__ void load_regs(const nuc3d_sensor_dev *__restrict__ S, Regs &R)
{
    const int l = threadIdx.x & 63, g = l >> 4, i = l & 15;
#pragma unroll
    for (int mt = 0; mt < 2; mt++) {
        const unsigned short *r1 = S->w1 + (16 * mt + i) * KIN;
        const unsigned short *r2 = S->w2 + (16 * mt + i) * NH;
        u16x8 a, b;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            a[e] = r1[8 * g + e];
            b[e] = e < 4 ? r2[4 * g + e] : r2[16 + 4 * g + (e - 4)];
        }
        R.w1[mt] = __builtin_bit_cast(bf16x8, a);
        R.w2[mt] = __builtin_bit_cast(bf16x8, b);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            R.b1[mt][r] = S->b1[16 * mt + 4 * g + r];
            R.b2[mt][r] = S->b2[16 * mt + 4 * g + r];
            R.w3[mt][r] = S->w3[16 * mt + 4 * g + r];
        }
    }
    R.b3 = S->b3;
}

// Activations of one 16-sample tile kept for the backward pass (per lane:
// sample l%16, hidden units 16mt + 4g + r).
struct Acts {
    float h1[2][4];    // layer-1 pre-activations
    float h2[2][4];    // layer-2 pre-activations
    bf16x8 a1;         // relu(h1) as bf16, in the permuted K order
    float a2[2][4];    // relu(h2)
};

// Forward pass of one tile of 16 samples by one wavefront: x is the lane's B
// operand (features 8g..8g+7 of sample l%16).  Returns the logit of sample
// l%16 (the same in the four lanes that share it).
__device__ __forceinline__ float forward16(const Regs &R, bf16x8 x, Acts &A)
{
    const f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
    f32x4 c[2];
    c[0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(R.w1[0], x, zero, 0, 0, 0);
    c[1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(R.w1[1], x, zero, 0, 0, 0);
    u16x8 a1;
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const float h = c[mt][r] + R.b1[mt][r];
            A.h1[mt][r] = h;
            a1[4 * mt + r] = bf16_bits(fmaxf(h, 0.0f));
        }
    A.a1 = __builtin_bit_cast(bf16x8, a1);
    f32x4 d[2];
    d[0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(R.w2[0], A.a1, zero, 0, 0, 0);
    d[1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(R.w2[1], A.a1, zero, 0, 0, 0);
    float z = 0.0f;
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const float h = d[mt][r] + R.b2[mt][r];
            A.h2[mt][r] = h;
            const float a = fmaxf(h, 0.0f);
            A.a2[mt][r] = a;
            z = z + R.w3[mt][r] * a;
        }
    z = z + __shfl_xor(z, 16);
    z = z + __shfl_xor(z, 32);
    return z + R.b3;
}

}  // namespace nuc3d_sensor
