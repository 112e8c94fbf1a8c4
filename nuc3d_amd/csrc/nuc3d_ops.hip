// Operator-level kernels: one reconstruction, derivative or TVD-RK3 update
// of whole Fields, with the semantics of the reference's plugin interface
// (interoplation / differential / integration, fieldOperator.h:15-80).
// A Field is nx*ny*nz float64, i-fastest; the operator acts along x on the
// ny*nz lines.  A ghost Field ("boundary") holds w cells per line: the left
// one cells -w..-1 at slots 0..w-1, the right one cells n..n+w-1
// (weno5js.cpp:255-266, :470-485).  These kernels keep the reference's
// operation order (the same device functions as the stage kernels' exact
// variant); they are the parity path of the drop-in operators, not the
// performance path.
#define NUC3D_EXACT 1
#include "nuc3d_device.h"
#include "nuc3d_types.h"

using namespace nuc3d_dev;

namespace {

// cell a of line l, -w <= a < n + w: interior, left ghost or right ghost.  A
// part that has no ghost Field on a side passes NULL there; its stencils never
// use those cells, and the value returned for them is 0.
__device__ __forceinline__ double cell_at(const double *in, const double *gl, const double *gr, int n, int w, int l,
                                          int a)
{
    if (a < 0) return gl ? gl[(long long)l * w + (a + w)] : 0.0;
    if (a >= n) return gr ? gr[(long long)l * w + (a - n)] : 0.0;
    return in[(long long)l * n + a];
}

// Faces [m0, m0+nf) of every line.  kind 0: scheme face (interior formula,
// with the WENO5-Z quirk on + inner faces when zq); kind 1: exterior closure
// on the low side; kind 2: exterior closure on the high side
// (reconstructionBoundaryScheme.cpp:52-312).
template <int SCHEME>
__global__ __launch_bounds__(256) void k_op_interp(int kind, int plus, int zq, int n, int w, int nlines, int m0,
                                                   int nf, const double *__restrict__ in,
                                                   const double *__restrict__ gl, const double *__restrict__ gr,
                                                   double *__restrict__ out)
{
    const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (e >= (long long)nf * nlines) return;
    const int l = (int)(e / nf), m = m0 + (int)(e % nf);
    double s[6];
#pragma unroll
    for (int o = 0; o < 6; o++) {
        const int a = m - 3 + o;
        // slots a stencil never reads may lie beyond the ghost frame
        s[o] = (a >= -w && a < n + w) ? cell_at(in, gl, gr, n, w, l, a) : 0.0;
    }
    double v;
    if (kind == 1)
        v = face_exterior(s, plus != 0, m);
    else if (kind == 2)
        v = face_exterior(s, plus != 0, n - m);
    else
        v = face_interior<SCHEME>(s, plus != 0, zq != 0);
    out[(long long)l * (n + 1) + m] = v;
}

// CRWENO5 inner faces w..n-w of every line (crweno5.cpp:44-222 for +,
// :224-401 for -): rows 0 and M-1 are explicit WENO5-JS, the inner rows the
// compact system; Thomas forward sweep and back substitution
// (crweno5.cpp:207-218, :387-398).  One thread per line; c' and d' in scratch.
__global__ __launch_bounds__(64) void k_op_crweno(int plus, int n, int w, int nlines, const double *__restrict__ in,
                                                  double *__restrict__ out, double *__restrict__ scratch)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nlines) return;
    const int M = n - 2 * w + 1;
    const double *p = in + (long long)l * n;
    double *C = scratch + (long long)l * 2 * M, *D = C + M;
    double cprev = 0.0, dprev = 0.0;
    for (int r = 0; r < M; r++) {
        const int i = w - 1 + r;  // reference loop index, face i+1
        double a, b, c, d;
        if (r == 0 || r == M - 1) {
            const double *s6 = p + i - 2;  // cells i-2..i+3
            a = 0.0;
            b = 1.0;
            c = 0.0;
            d = plus ? weno5js(s6[0], s6[1], s6[2], s6[3], s6[4]) : weno5js(s6[5], s6[4], s6[3], s6[2], s6[1]);
        } else {
            crweno_row(p + i, 1, plus != 0, a, b, c, d);
        }
        if (r > 0) {
            c = c / (b - cprev * a);
            d = (d - dprev * a) / (b - cprev * a);
        }
        C[r] = c;
        D[r] = d;
        cprev = c;
        dprev = d;
    }
    double *o = out + (long long)l * (n + 1) + w;
    double x = D[M - 1];
    o[M - 1] = x;
    for (int r = M - 2; r >= 0; r--) {
        x = D[r] - C[r] * x;
        o[r] = x;
    }
}

// Derivative along x of cells [t0, t0+nt) (centraldifference6th.cpp:16-183,
// centraldifference8th.cpp:22-192); kind 1 / 2: exterior closure on the low /
// high side (boundaryScheme.cpp:31-126).
__global__ __launch_bounds__(256) void k_op_diff(int scheme, int kind, int n, int w, int nlines, int t0, int nt,
                                                 const double *__restrict__ in, const double *__restrict__ gl,
                                                 const double *__restrict__ gr, double *__restrict__ out)
{
    const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (e >= (long long)nt * nlines) return;
    const int l = (int)(e / nt), t = t0 + (int)(e % nt);
    // stencil radius: 4 for cd8, 3 for cd6 and the exterior closure
    const int R = (kind == 0 && scheme == 1) ? 4 : 3;
    double s[9];
#pragma unroll
    for (int o = 0; o < 9; o++) {
        const int a = t - 4 + o;
        s[o] = (a >= t - R && a <= t + R && a >= -w && a < n + w) ? cell_at(in, gl, gr, n, w, l, a) : 0.0;
    }
    double v;
    if (kind == 1)
        v = diff_exterior(&s[4], 1, t, true);
    else if (kind == 2)
        v = diff_exterior(&s[4], 1, n - 1 - t, false);
    else
        v = scheme == 1 ? cd8(&s[4], 1) : cd6(&s[4], 1);
    out[(long long)l * n + t] = v;
}

// TVD-RK3 stage over one field (TVD-RK3.cpp:37-137)
__global__ __launch_bounds__(256) void k_op_rk3(int stage, double dt, long long n, const double *__restrict__ rhs,
                                                const double *__restrict__ un, double *__restrict__ ui)
{
    const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (e >= n) return;
    const double r = -rhs[e];
    const double u = ui[e];
    double v;
    if (stage == 0)
        v = 1.0 * u + 1.0 * r * dt;
    else if (stage == 1)
        v = 3.0 / 4.0 * un[e] + 1.0 / 4.0 * (r * dt + u);
    else
        v = 1.0 / 3.0 * un[e] + 2.0 / 3.0 * (r * dt + u);
    ui[e] = v;
}

// Cell centres, coordinate derivatives, J and metric terms of one block from
// its mesh nodes (block::getXYZ_center / getJacobians, block.cpp:147-477):
// within each tile of 4x4x4 cells the five nodes per direction carry a
// fourth-degree Lagrange polynomial, evaluated (LAG) or differentiated per
// unit index (DLAG) at the cell centres 0.5..3.5 -- the weights of the
// reference's tables (block.h:36-47).  One thread per cell, the reference's
// accumulation order (zeta outermost, block.cpp:212-234), no contraction.
__constant__ double LAG[4][5] = {
    {35.0 / 128.0, 35.0 / 32.0, -35.0 / 64.0, 7.0 / 32.0, -5.0 / 128.0},
    {-5.0 / 128.0, 15.0 / 32.0, 45.0 / 64.0, -5.0 / 32.0, 3.0 / 128.0},
    {3.0 / 128.0, -5.0 / 32.0, 45.0 / 64.0, 15.0 / 32.0, -5.0 / 128.0},
    {-5.0 / 128.0, 7.0 / 32.0, -35.0 / 64.0, 35.0 / 32.0, 35.0 / 128.0}};
__constant__ double DLAG[4][5] = {
    {-11.0 / 12.0, 17.0 / 24.0, 3.0 / 8.0, -5.0 / 24.0, 1.0 / 24.0},
    {1.0 / 24.0, -9.0 / 8.0, 9.0 / 8.0, -1.0 / 24.0, 0.0},
    {0.0, 1.0 / 24.0, -9.0 / 8.0, 9.0 / 8.0, -1.0 / 24.0},
    {-1.0 / 24.0, 5.0 / 24.0, -3.0 / 8.0, -17.0 / 24.0, 11.0 / 12.0}};

__device__ __forceinline__ double tile_sum(const double *f, long long sy, long long sz, const double *wx,
                                           const double *wy, const double *wz)
{
    double vz = 0.0;
    for (int k = 0; k <= 4; k++) {
        double ve = 0.0;
        for (int j = 0; j <= 4; j++) {
            double vx = 0.0;
#pragma unroll
            for (int i = 0; i <= 4; i++) vx += wx[i] * f[i + j * sy + k * sz];
            ve += wy[j] * vx;
        }
        vz += wz[k] * ve;
    }
    return vz;
}

__global__ __launch_bounds__(256) void k_mesh_metrics(nuc3d_block_dev B, const double *__restrict__ xn,
                                                      const double *__restrict__ yn, const double *__restrict__ zn,
                                                      double *__restrict__ centre, int *__restrict__ bad)
{
    const int nx = B.n[0], ny = B.n[1], nz = B.n[2];
    const long long N = (long long)nx * ny * nz;
    const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (e >= N) return;
    const int i = (int)(e % nx), j = (int)((e / nx) % ny), k = (int)(e / ((long long)nx * ny));
    const long long sy = nx + 1, sz = (long long)(nx + 1) * (ny + 1);
    const long long base = (i / 4) * 4 + ((j / 4) * 4) * sy + ((k / 4) * 4) * sz;
    const int li = i % 4, lj = j % 4, lk = k % 4;
    const double *node[3] = {xn, yn, zn};
    double dv[9];
    for (int c = 0; c < 3; c++) {
        if (centre) centre[c * N + e] = tile_sum(node[c] + base, sy, sz, LAG[li], LAG[lj], LAG[lk]);
        dv[3 * c + 0] = tile_sum(node[c] + base, sy, sz, DLAG[li], LAG[lj], LAG[lk]);
        dv[3 * c + 1] = tile_sum(node[c] + base, sy, sz, LAG[li], DLAG[lj], LAG[lk]);
        dv[3 * c + 2] = tile_sum(node[c] + base, sy, sz, LAG[li], LAG[lj], DLAG[lk]);
    }
    const double x_xi = dv[0], x_eta = dv[1], x_zeta = dv[2];
    const double y_xi = dv[3], y_eta = dv[4], y_zeta = dv[5];
    const double z_xi = dv[6], z_eta = dv[7], z_zeta = dv[8];
    const double jacob = 1.0 / (x_xi * (y_eta * z_zeta - y_zeta * z_eta) + x_eta * (y_zeta * z_xi - y_xi * z_zeta) +
                                x_zeta * (y_xi * z_eta - y_eta * z_xi));
    const long long x = i + B.sy * j + B.sz * k;
    B.met[0][x] = (y_eta * z_zeta - y_zeta * z_eta) * jacob;
    B.met[1][x] = (x_zeta * z_eta - x_eta * z_zeta) * jacob;
    B.met[2][x] = (x_eta * y_zeta - x_zeta * y_eta) * jacob;
    B.met[3][x] = (y_zeta * z_xi - y_xi * z_zeta) * jacob;
    B.met[4][x] = (x_xi * z_zeta - x_zeta * z_xi) * jacob;
    B.met[5][x] = (x_zeta * y_xi - x_xi * y_zeta) * jacob;
    B.met[6][x] = (y_xi * z_eta - y_eta * z_xi) * jacob;
    B.met[7][x] = (x_eta * z_xi - x_xi * z_eta) * jacob;
    B.met[8][x] = (x_xi * y_eta - x_eta * y_xi) * jacob;
#pragma unroll
    for (int c = 0; c < 9; c++) B.dxyz[c][x] = dv[c];
    B.jac[x] = jacob;
    if (!(jacob > 0.0)) atomicOr(bad, 1);  // block.cpp:310-314 exits on J < 0
}

unsigned grid_of(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

extern "C" hipError_t nuc3d_launch_op_interp(int scheme, int kind, int plus, int zq, int n, int w, int nlines, int m0,
                                             int nf, const double *in, const double *gl, const double *gr, double *out,
                                             hipStream_t st)
{
    const unsigned g = grid_of((long long)nf * nlines, 256);
    if (scheme == SCH_WENO5Z)
        hipLaunchKernelGGL(k_op_interp<SCH_WENO5Z>, dim3(g), dim3(256), 0, st, kind, plus, zq, n, w, nlines, m0, nf,
                           in, gl, gr, out);
    else
        hipLaunchKernelGGL(k_op_interp<SCH_WENO5JS>, dim3(g), dim3(256), 0, st, kind, plus, zq, n, w, nlines, m0, nf,
                           in, gl, gr, out);
    return hipGetLastError();
}

extern "C" hipError_t nuc3d_launch_op_crweno(int plus, int n, int w, int nlines, const double *in, double *out,
                                             double *scratch, hipStream_t st)
{
    hipLaunchKernelGGL(k_op_crweno, dim3(grid_of(nlines, 64)), dim3(64), 0, st, plus, n, w, nlines, in, out, scratch);
    return hipGetLastError();
}

extern "C" hipError_t nuc3d_launch_op_diff(int scheme, int kind, int n, int w, int nlines, int t0, int nt,
                                           const double *in, const double *gl, const double *gr, double *out,
                                           hipStream_t st)
{
    hipLaunchKernelGGL(k_op_diff, dim3(grid_of((long long)nt * nlines, 256)), dim3(256), 0, st, scheme, kind, n, w,
                       nlines, t0, nt, in, gl, gr, out);
    return hipGetLastError();
}

extern "C" hipError_t nuc3d_launch_op_rk3(int stage, double dt, long long n, const double *rhs, const double *un,
                                          double *ui, hipStream_t st)
{
    hipLaunchKernelGGL(k_op_rk3, dim3(grid_of(n, 256)), dim3(256), 0, st, stage, dt, n, rhs, un, ui);
    return hipGetLastError();
}

// Uniform-mesh check of a block (no counterpart in the reference): bad[0]
// becomes 1 if J or a metric term of a cell that the stage kernels read (the
// interior, and the ghost slabs of the faces joined to a neighbour) differs,
// bit for bit, from cell (0,0,0)'s.  Edge and corner ghosts and the ghost
// slabs of physical faces are not read by them.
__global__ __launch_bounds__(256) void k_uniform_check(nuc3d_block_dev B, int *bad)
{
    const int G = B.G;
    const long long px = B.n[0] + 2LL * G, py = B.n[1] + 2LL * G, pz = B.n[2] + 2LL * G;
    const long long total = px * py * pz;
    const unsigned long long j0 = __double_as_longlong(B.jac[0]);
    int mism = 0;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long long)gridDim.x * 256) {
        const int c3[3] = {(int)(e % px) - G, (int)((e / px) % py) - G, (int)(e / (px * py)) - G};
        int out = 0, face = 0;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            if (c3[d] < 0) {
                out++;
                face = 2 * d;
            } else if (c3[d] >= B.n[d]) {
                out++;
                face = 2 * d + 1;
            }
        }
        if (out > 1 || (out == 1 && B.ftype[face] < 0)) continue;
        const long long x = c3[0] + B.sy * c3[1] + B.sz * c3[2];
        if ((unsigned long long)__double_as_longlong(B.jac[x]) != j0) mism = 1;
#pragma unroll
        for (int m = 0; m < 9; m++)
            if (__double_as_longlong(B.met[m][x]) != __double_as_longlong(B.met[m][0])) mism = 1;
    }
    if (mism) atomicOr(bad, 1);
}

extern "C" hipError_t nuc3d_launch_uniform_check(nuc3d_block_dev B, int *bad, hipStream_t st)
{
    hipLaunchKernelGGL(k_uniform_check, dim3(1024), dim3(256), 0, st, B, bad);
    return hipGetLastError();
}

// Constants of a uniform block, computed here so that they are the values the
// stage kernels would compute per cell: |grad xi_d| = sqrt(mx^2 + my^2 + mz^2)
// of each direction's metric terms (physicsModel.cpp:557-597), the reciprocal
// half of a division by J (recip_nr), and whether J and every metric term lie
// in the range where the split's divisions p/J and m p/J may share that
// reciprocal (J positive in [2^-200, 2^200], each term 0 or of magnitude there).
// out: theta[3], recip, flag (1.0 or 0.0).
struct nuc3d_consts_in {
    double jac, met[9];
};
__global__ void k_block_consts(nuc3d_consts_in in, double *out)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    bool ok = nuc3d_dev::rd_denominator(in.jac);
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double mx = in.met[3 * d], my = in.met[3 * d + 1], mz = in.met[3 * d + 2];
        out[d] = sqrt(mx * mx + my * my + mz * mz);
        ok = ok && nuc3d_dev::rd_factor(mx) && nuc3d_dev::rd_factor(my) && nuc3d_dev::rd_factor(mz);
    }
    out[3] = nuc3d_dev::recip_nr(in.jac);
    out[4] = ok ? 1.0 : 0.0;
}

extern "C" hipError_t nuc3d_launch_block_consts(double jac0, const double *met0, double *out5, hipStream_t st)
{
    nuc3d_consts_in in;
    in.jac = jac0;
    for (int m = 0; m < 9; m++) in.met[m] = met0[m];
    hipLaunchKernelGGL(k_block_consts, dim3(1), dim3(64), 0, st, in, out5);
    return hipGetLastError();
}

extern "C" hipError_t nuc3d_launch_mesh_metrics(nuc3d_block_dev B, const double *xn, const double *yn,
                                                const double *zn, double *centre, int *bad, hipStream_t st)
{
    const long long N = (long long)B.n[0] * B.n[1] * B.n[2];
    hipLaunchKernelGGL(k_mesh_metrics, dim3(grid_of(N, 256)), dim3(256), 0, st, B, xn, yn, zn, centre, bad);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Self-test of the reduced FP64 division (recip_nr / div_nr of
// nuc3d_device.h) against the compiler's IEEE a / b on this device: n
// operand pairs from a hash of (seed, index), b = +-m * 2^eb and a = +0 (one
// in 16) or +-m * 2^ea with eb, ea in [-600, 600] and |ea - eb| < 700, the
// mantissa m random or an edge pattern (1, 2 - ulp, 1 + ulp, 1.5); each pair
// also divides a - ulp and 2a - ulp by the same reciprocal (the shared form),
// and a, 2a - ulp and a signed zero over |b| by the sign-keeping form div_nrs.
// Counts quotients whose bits differ.
namespace {
__device__ __forceinline__ unsigned long long mix64(unsigned long long x)
{
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33;
    return x;
}

__device__ __forceinline__ double make_operand(unsigned long long h, int e)
{
    unsigned long long mant;
    switch ((h >> 60) & 7) {
    case 0: mant = 0; break;                       // 1.0
    case 1: mant = 0xfffffffffffffULL; break;      // 2 - ulp
    case 2: mant = 1; break;                       // 1 + ulp
    case 3: mant = 0x8000000000000ULL; break;      // 1.5
    default: mant = h & 0xfffffffffffffULL; break;
    }
    const unsigned long long sign = (h >> 59) & 1;
    return __longlong_as_double((long long)((sign << 63) | ((unsigned long long)(e + 1023) << 52) | mant));
}
}  // namespace

__global__ __launch_bounds__(256) void k_selftest_division(long long n, unsigned long long seed,
                                                            unsigned long long *mismatch)
{
    unsigned long long bad = 0;
    for (long long t = blockIdx.x * 256LL + threadIdx.x; t < n; t += (long long)gridDim.x * 256) {
        const unsigned long long h1 = mix64(seed ^ (unsigned long long)(2 * t + 1) * 0x9e3779b97f4a7c15ULL);
        const unsigned long long h2 = mix64(h1 + 0x632be59bd9b4e019ULL);
        const int eb = (int)(h1 % 1201ULL) - 600;
        int ea = (int)(h2 % 1201ULL) - 600;
        if (ea - eb >= 700) ea = eb + 699;
        if (eb - ea >= 700) ea = eb - 699;
        const double b = make_operand(h1, eb);
        const double a = ((h2 >> 40) & 15) == 0 ? 0.0 : make_operand(h2, ea);
        const double a2 = a == 0.0 ? 0.0 : __longlong_as_double(__double_as_longlong(a) - 1);
        const double a3 = a2 + a2;
        const double y = nuc3d_dev::recip_nr(b);
        // and a signed zero over b by div_nrz
        const double z = ((h2 >> 44) & 1) ? -0.0 : 0.0;
        // and the same over |b| through div_nrs, whose correction keeps the
        // IEEE sign of a zero over a positive denominator
        const double bp = fabs(b), yp = nuc3d_dev::recip_nr(bp);
        // and the square root of |a / b| (in [2^-700, 2^700] for a != 0) by sqrt_nr
        const double sq = fabs(a / b);
        const double q[8] = {nuc3d_dev::div_nr(a, b, y),    nuc3d_dev::div_nr(a2, b, y),
                             nuc3d_dev::div_nr(a3, b, y),   nuc3d_dev::div_nrz(z, b, y),
                             nuc3d_dev::div_nrs(a, bp, yp), nuc3d_dev::div_nrs(a3, bp, yp),
                             nuc3d_dev::div_nrs(z, bp, yp), a == 0.0 ? 0.0 : nuc3d_dev::sqrt_nr(sq)};
        const double r[8] = {a / b, a2 / b, a3 / b, z / b, a / bp, a3 / bp, z / bp, sqrt(sq)};
#pragma unroll
        for (int k = 0; k < 8; k++)
            if (__double_as_longlong(q[k]) != __double_as_longlong(r[k])) bad++;
    }
    if (bad) atomicAdd(mismatch, bad);
}

extern "C" hipError_t nuc3d_launch_selftest_division(long long n, unsigned long long seed,
                                                     unsigned long long *mismatch, hipStream_t st)
{
    hipLaunchKernelGGL(k_selftest_division, dim3(4096), dim3(256), 0, st, n, seed, mismatch);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Sums of the diagnostics' per-workgroup partials (pairs: enstrophy, kinetic
// energy) in index order by one thread: the order, and so the rounding, of
// the host loop of nuc3d_diagnostics.  out2 receives the two sums.
__global__ void k_sum_pairs_ordered(const double *__restrict__ part, long long n, double *__restrict__ out2)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double e = 0.0, k = 0.0;
    for (long long w = 0; w < n; w++) {
        e += part[2 * w];
        k += part[2 * w + 1];
    }
    out2[0] = e;
    out2[1] = k;
}

extern "C" hipError_t nuc3d_launch_sum_pairs(const double *part, long long n, double *out2, hipStream_t st)
{
    hipLaunchKernelGGL(k_sum_pairs_ordered, dim3(1), dim3(64), 0, st, part, n, out2);
    return hipGetLastError();
}
