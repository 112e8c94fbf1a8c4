// nuc3d_sensor.hip -- sampling, training and inference kernels of the learned
// shock sensor of the hybrid reconstruction (no counterpart in the
// reference).  The network, its features and its teacher are defined in
// nuc3d_sensor.h; the hybrid sweep (nuc3d_kernels.hip) runs the same
// forward16() on the same features, so a logit computed here for a face is
// the one the sweep computes for it.
//
// Training is plain minibatch Adam on the FP32 master parameters.  Each
// step: k_sensor_fwdbwd (forward and backward of 128 samples per workgroup on
// the matrix cores, BF16 operands, FP32 accumulation; per-workgroup partial
// gradients), k_sensor_reduce (fixed-order sum of the partials), k_sensor_adam.
// Every reduction has a fixed order, so training is reproducible.
#include "nuc3d_device.h"
#include "nuc3d_sensor.h"
#include "nuc3d_types.h"

using namespace nuc3d_dev;
using namespace nuc3d_sensor;

namespace {

constexpr int OW1 = 0, OB1 = 1024, OW2 = 1056, OB2 = 2080, OW3 = 2112, OB3 = 2144, NP = NUC3D_SENSOR_NPARAM;
static_assert(NP == 2145, "parameter layout");

__device__ __forceinline__ unsigned long long mix64(unsigned long long z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ long long cidx(const nuc3d_block_dev &B, int i, int j, int k)
{
    return (long long)i + B.sy * j + B.sz * k;
}

}  // namespace

// ---------------------------------------------------------------------------
// Samples of direction d of block b from state buffer buf: the faces of the
// rotated face array (n_d+1) x n_{d+1} x n_{d+2} (face m fastest) whose linear
// index L satisfies L % stride == offset, stored at out + L / stride.  Each
// sample is the face's 30 features (nuc3d_sensor.h) and its teacher label
// (1 / 0), or -1 when the face's eight-cell stencil leaves the block and its
// ghost frame, meets a physical face, or has no positive mean rho or p.
// The positives are counted into *npos.
__global__ __launch_bounds__(256) void k_sensor_sample(const nuc3d_block_dev *__restrict__ blocks, int b, int d,
                                                       int buf, double gamma, double tau, long long stride,
                                                       long long offset, long long nsamp, unsigned *X, float *Y,
                                                       unsigned long long *npos)
{
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int pos = 0;
    if (s < nsamp) {
        const nuc3d_block_dev &B = blocks[b];
        const int n1 = B.n[d], n2 = B.n[(d + 1) % 3];
        const long long L = offset + s * stride;
        const int m = (int)(L % (n1 + 1));
        const long long r = L / (n1 + 1);
        const int a2 = (int)(r % n2), a3 = (int)(r / n2);
        const int lo = B.ftype[2 * d] == -1 ? 4 : 4 - B.G;
        const int hi = B.ftype[2 * d + 1] == -1 ? n1 - 4 : n1 + B.G - 4;
        unsigned xw[16];
        float y = -1.0f;
        if (m >= lo && m <= hi) {
            PhysParams P;
            P.gamma = gamma;
            // primitives of cells m-4..m+3 along d
            double pr[5][8];
            int c3[3];
            c3[(d + 1) % 3] = a2;
            c3[(d + 2) % 3] = a3;
            for (int j = 0; j < 8; j++) {
                c3[d] = m - 4 + j;
                const long long x = cidx(B, c3[0], c3[1], c3[2]);
                double Q[5];
#pragma unroll
                for (int c = 0; c < 5; c++) Q[c] = B.q[buf][c][x];
                double cs;
                con2prim(Q, B.jac[x], P, pr[0][j], pr[1][j], pr[2][j], pr[3][j], pr[4][j], cs);
            }
            auto getf = [&](int v, int j) { return (float)pr[v][j + 1]; };
            auto get8 = [&](int v, int j) { return pr[v][j]; };
            if (features(getf, (float)gamma, xw)) {
                const int lab = teacher(get8, gamma, tau);
                y = (float)lab;
                pos = lab == 1;
            }
        }
        if (y < 0.0f)
#pragma unroll
            for (int i = 0; i < 16; i++) xw[i] = 0u;
        uint4 *dst = reinterpret_cast<uint4 *>(X + 16 * s);
#pragma unroll
        for (int i = 0; i < 4; i++) dst[i] = make_uint4(xw[4 * i], xw[4 * i + 1], xw[4 * i + 2], xw[4 * i + 3]);
        Y[s] = y;
    }
    const unsigned long long bal = __ballot(pos);
    if ((threadIdx.x & 63) == 0 && bal) atomicAdd(npos, (unsigned long long)__popcll(bal));
}

// ---------------------------------------------------------------------------
// Forward (and backward) of a batch: sample p of the batch is dataset entry
// rnd ? mix(seed, step, p) % count : first + p, for p < nsamp.  Loss per
// sample: wt (softplus(z) - y z) with wt = pos_weight for y = 1, 1 for y = 0,
// 0 for y < 0 or p >= nsamp, times inv_batch.  Writes the workgroup's
// partial gradient and loss to partial[blockIdx.x][0..NP] (loss last) and,
// when logits is set, the logit of every sample.
__global__ __launch_bounds__(128) void k_sensor_fwdbwd(const unsigned *__restrict__ X, const float *__restrict__ Y,
                                                       long long count, long long first, long long nsamp, int rnd,
                                                       unsigned long long seed, long long step,
                                                       const float *__restrict__ params, float pos_weight,
                                                       float inv_batch, int bwd, float *partial, float *logits)
{
    __shared__ nuc3d_sensor_dev SW;
    // per wave: relu(h1), dL/dh2, dL/dh1 and the features, bf16 [64][32];
    // afterwards the per-wave gradient rows
    __shared__ __align__(16) union {
        unsigned short mat[2][4][64][32];
        float red[2][NP + 1];
    } L;

    // bf16 copies of the weights (the inference struct layout)
    for (int e = threadIdx.x; e < 32 * 32; e += blockDim.x) {
        SW.w1[e] = bf16_bits(params[OW1 + e]);
        SW.w2[e] = bf16_bits(params[OW2 + e]);
    }
    for (int e = threadIdx.x; e < 32; e += blockDim.x) {
        SW.b1[e] = params[OB1 + e];
        SW.b2[e] = params[OB2 + e];
        SW.w3[e] = bf16_float(bf16_bits(params[OW3 + e]));
    }
    if (threadIdx.x == 0) {
        SW.b3 = params[OB3];
        SW.thr = 0.0f;
    }
    __syncthreads();

    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, g = lane >> 4, j = lane & 15;
    Regs R;
    load_regs(&SW, R);  // the same operands as load_regs_lanes gives the sweep
    // A operand of dA1^T = W2^T dH2^T: row = input unit 16mt + j, k = output unit 8g+e
    bf16x8 w2t[2];
#pragma unroll
    for (int mt = 0; mt < 2; mt++) {
        u16x8 a;
#pragma unroll
        for (int e = 0; e < 8; e++) a[e] = SW.w2[(8 * g + e) * 32 + 16 * mt + j];
        w2t[mt] = __builtin_bit_cast(bf16x8, a);
    }
    unsigned short(*A1)[32] = L.mat[wv][0];
    unsigned short(*DH2)[32] = L.mat[wv][1];
    unsigned short(*DH1)[32] = L.mat[wv][2];
    unsigned short(*XS)[32] = L.mat[wv][3];

    float gb1[2][4] = {}, gb2[2][4] = {}, gw3[2][4] = {};
    float gb3 = 0.0f, loss = 0.0f;
    const f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int q = 0; q < 4; q++) {
        const int s = 64 * wv + 16 * q + j;                 // sample within the workgroup
        const long long p = (long long)blockIdx.x * 128 + s;  // position in the batch
        const bool inb = p < nsamp;
        long long idx = 0;
        if (inb) idx = rnd ? (long long)(mix64(seed ^ mix64((unsigned long long)step * 0x100000001B3ull +
                                                             (unsigned long long)p)) % (unsigned long long)count)
                           : first + p;
        uint4 v4 = make_uint4(0u, 0u, 0u, 0u);
        float y = -1.0f;
        if (inb) {
            v4 = *reinterpret_cast<const uint4 *>(X + 16 * idx + 4 * g);
            y = Y[idx];
        }
        const unsigned wds[4] = {v4.x, v4.y, v4.z, v4.w};
        u16x8 xv;
#pragma unroll
        for (int e = 0; e < 4; e++) {
            xv[2 * e] = (unsigned short)(wds[e] & 0xffffu);
            xv[2 * e + 1] = (unsigned short)(wds[e] >> 16);
        }
        Acts A;
        const float z = forward16(R, __builtin_bit_cast(bf16x8, xv), A);
        if (logits && inb && g == 0) logits[p] = z;
        if (!bwd) continue;
        const float wt = (inb && y >= 0.0f) ? (y > 0.5f ? pos_weight : 1.0f) * inv_batch : 0.0f;
        const float yy = y > 0.5f ? 1.0f : 0.0f;
        const float sig = 1.0f / (1.0f + expf(-z));
        const float dz = wt * (sig - yy);
        if (g == 0) {
            // softplus(z) - y z, written stably
            const float sp = fmaxf(z, 0.0f) + log1pf(expf(-fabsf(z)));
            loss += wt * (sp - yy * z);
            gb3 += dz;
        }
        u16x8 dh2b;
#pragma unroll
        for (int mt = 0; mt < 2; mt++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const float dh2 = A.h2[mt][r] > 0.0f ? dz * R.w3f[mt][r] : 0.0f;
                gw3[mt][r] += dz * bf16_float(__builtin_bit_cast(u16x8, A.a2)[4 * mt + r]);
                gb2[mt][r] += dh2;
                const unsigned short hb = bf16_bits(dh2);
                dh2b[4 * mt + r] = hb;
                DH2[16 * q + j][16 * mt + 4 * g + r] = hb;
                A1[16 * q + j][16 * mt + 4 * g + r] = __builtin_bit_cast(u16x8, A.a1)[4 * mt + r];
            }
#pragma unroll
        for (int e = 0; e < 8; e++) XS[16 * q + j][8 * g + e] = xv[e];
        // dA1^T (inputs x samples) = W2^T dH2^T; the lane's B operand is dH2 of
        // sample j at output units 8g..8g+7 (natural order)
        __syncthreads();
        u16x8 bop;
#pragma unroll
        for (int e = 0; e < 8; e++) bop[e] = DH2[16 * q + j][8 * g + e];
        f32x4 da[2];
        da[0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w2t[0], __builtin_bit_cast(bf16x8, bop), zero, 0, 0, 0);
        da[1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w2t[1], __builtin_bit_cast(bf16x8, bop), zero, 0, 0, 0);
#pragma unroll
        for (int mt = 0; mt < 2; mt++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const float dh1 = A.h1[mt][r] > 0.0f ? da[mt][r] : 0.0f;
                gb1[mt][r] += dh1;
                DH1[16 * q + j][16 * mt + 4 * g + r] = bf16_bits(dh1);
            }
        (void)dh2b;
    }
    if (!bwd) return;
    __syncthreads();
    // weight gradients over the wave's 64 samples: dW2 = dH2^T A1, dW1 = dH1^T X
    f32x4 gw2[2][2], gw1[2][2];
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
        for (int nt = 0; nt < 2; nt++) {
            f32x4 a2c = zero, a1c = zero;
#pragma unroll
            for (int ks = 0; ks < 2; ks++) {
                u16x8 ao2, bo2, ao1, bo1;
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    const int s = 32 * ks + 8 * g + e;
                    ao2[e] = DH2[s][16 * mt + j];
                    bo2[e] = A1[s][16 * nt + j];
                    ao1[e] = DH1[s][16 * mt + j];
                    bo1[e] = XS[s][16 * nt + j];
                }
                a2c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, ao2),
                                                              __builtin_bit_cast(bf16x8, bo2), a2c, 0, 0, 0);
                a1c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, ao1),
                                                              __builtin_bit_cast(bf16x8, bo1), a1c, 0, 0, 0);
            }
            gw2[mt][nt] = a2c;
            gw1[mt][nt] = a1c;
        }
    // per-unit sums over the 16 samples of a lane group (butterfly, fixed order)
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) {
                gb1[mt][r] += __shfl_xor(gb1[mt][r], o);
                gb2[mt][r] += __shfl_xor(gb2[mt][r], o);
                gw3[mt][r] += __shfl_xor(gw3[mt][r], o);
            }
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        gb3 += __shfl_xor(gb3, o);
        loss += __shfl_xor(loss, o);
    }
    __syncthreads();  // the bf16 matrices are dead; the union becomes the gradient rows
    float *red = L.red[wv];
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
        for (int nt = 0; nt < 2; nt++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = 16 * mt + 4 * g + r, col = 16 * nt + j;
                red[OW2 + row * 32 + col] = gw2[mt][nt][r];
                red[OW1 + row * 32 + col] = gw1[mt][nt][r];
            }
    if (j == 0)
#pragma unroll
        for (int mt = 0; mt < 2; mt++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int u = 16 * mt + 4 * g + r;
                red[OB1 + u] = gb1[mt][r];
                red[OB2 + u] = gb2[mt][r];
                red[OW3 + u] = gw3[mt][r];
            }
    if (lane == 0) {
        red[OB3] = gb3;
        red[NP] = loss;
    }
    __syncthreads();
    float *dst = partial + (long long)blockIdx.x * (NP + 1);
    for (int e = threadIdx.x; e < NP + 1; e += blockDim.x) dst[e] = L.red[0][e] + L.red[1][e];
}

// grad[e] = sum over workgroups w of partial[w][e], in order of w
__global__ __launch_bounds__(256) void k_sensor_reduce(const float *__restrict__ partial, int nwg, float *grad)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e > NP) return;
    float s = 0.0f;
    for (int w = 0; w < nwg; w++) s += partial[(long long)w * (NP + 1) + e];
    grad[e] = s;
}

// Adam (beta1 0.9, beta2 0.999, eps 1e-8) on the master parameters; the two
// padding inputs of W1 stay zero.  grad[NP] (the loss) is copied to loss_out.
__global__ __launch_bounds__(256) void k_sensor_adam(float *params, float *mom, float *vel, const float *grad,
                                                     float lr, int t, float *loss_out)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e == 0 && loss_out) *loss_out = grad[NP];
    if (e >= NP) return;
    if (e < OB1 && (e & 31) >= NUC3D_SENSOR_NFEAT) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float gr = grad[e];
    const float m = b1 * mom[e] + (1.0f - b1) * gr;
    const float v = b2 * vel[e] + (1.0f - b2) * gr * gr;
    mom[e] = m;
    vel[e] = v;
    const float mh = m / (1.0f - powf(b1, (float)t));
    const float vh = v / (1.0f - powf(b2, (float)t));
    params[e] -= lr * mh / (sqrtf(vh) + eps);
}

// ---------------------------------------------------------------------------
// host-side launchers
extern "C" hipError_t nuc3d_launch_sensor_sample(const nuc3d_block_dev *blocks, int b, int d, int buf, double gamma,
                                                 double tau, long long stride, long long offset, long long nsamp,
                                                 unsigned *X, float *Y, unsigned long long *npos, hipStream_t st)
{
    if (nsamp <= 0) return hipSuccess;
    const long long nwg = (nsamp + 255) / 256;
    hipLaunchKernelGGL(k_sensor_sample, dim3((unsigned)nwg), dim3(256), 0, st, blocks, b, d, buf, gamma, tau, stride,
                       offset, nsamp, X, Y, npos);
    return hipGetLastError();
}

// one forward/backward pass; returns the number of workgroups (partials)
extern "C" hipError_t nuc3d_launch_sensor_fwdbwd(const unsigned *X, const float *Y, long long count, long long first,
                                                 long long nsamp, int rnd, unsigned long long seed, long long step,
                                                 const float *params, float pos_weight, float inv_batch, int bwd,
                                                 float *partial, float *grad, float *logits, hipStream_t st)
{
    if (nsamp <= 0) return hipSuccess;
    const int nwg = (int)((nsamp + 127) / 128);
    hipLaunchKernelGGL(k_sensor_fwdbwd, dim3(nwg), dim3(128), 0, st, X, Y, count, first, nsamp, rnd, seed, step,
                       params, pos_weight, inv_batch, bwd, partial, logits);
    if (bwd) hipLaunchKernelGGL(k_sensor_reduce, dim3((NP + 1 + 255) / 256), dim3(256), 0, st, partial, nwg, grad);
    return hipGetLastError();
}

extern "C" hipError_t nuc3d_launch_sensor_adam(float *params, float *mom, float *vel, const float *grad, float lr,
                                               int t, float *loss_out, hipStream_t st)
{
    hipLaunchKernelGGL(k_sensor_adam, dim3((NP + 255) / 256), dim3(256), 0, st, params, mom, vel, grad, lr, t,
                       loss_out);
    return hipGetLastError();
}
