// nuc3d_kernels.hip -- HIP kernels of the per-block RHS + TVD-RK3 path (gfx950).
//
// Compiled twice: with -DNUC3D_EXACT=1 -ffp-contract=off (kernel names end in
// _exact) and with -DNUC3D_EXACT=0 (names end in _fast).  The host side picks
// one set per case.
//
// One RK stage is three direction sweeps (k_sweep<0>, <1>, <2>): each
// workgroup stages the split fluxes of 16 lines x 21 cells in LDS
// (physicsModel.cpp:528-754), reconstructs 16 faces per line with one face
// per thread (WENO5-JS / WENO5-Z or the physical closures) and differences
// them into the RHS (euler3d.cpp:143-241).  The z sweep also subtracts the
// viscous divergence for Navier-Stokes (NaiverStokes3d.cpp:401-513) and
// applies the TVD-RK3 stage (TVD-RK3.cpp:37-137) into a second state buffer,
// pushing the new values into the ghost layers of neighbouring blocks (the
// on-device form of MPICommunicator.cpp:55-115).  Stage 2 also accumulates
// the maximum eigenvalues of the new state, from which the next step's dt is
// taken (euler3d.cpp:429-435, PDEData3d.cpp:74-80).
#include "nuc3d_device.h"
#include "nuc3d_types.h"

#if NUC3D_EXACT
#define KNAME(name) name##_exact
#else
#define KNAME(name) name##_fast
#endif

using namespace nuc3d_dev;

namespace {

__device__ __forceinline__ PhysParams phys_of(const nuc3d_kparams &K)
{
    PhysParams P;
    P.gamma = K.gamma;
    P.mach = K.mach;
    P.reynolds = K.reynolds;
    P.prandtl = K.prandtl;
    P.sref = K.sref;
    P.twall = K.twall;
    P.riemann = K.riemann;
    return P;
}

__device__ __forceinline__ long long cidx(const nuc3d_block_dev &B, int i, int j, int k)
{
    return (long long)i + B.sy * j + B.sz * k;
}

__device__ __forceinline__ long long dstride(const nuc3d_block_dev &B, int d)
{
    return d == 0 ? 1LL : (d == 1 ? B.sy : B.sz);
}

// split fluxes of cell x of buffer qb in direction d; returns the local eigenvalue
__device__ __forceinline__ double cell_split(const nuc3d_block_dev &B, double *const *qb, long long x, int d,
                                             const PhysParams &P, double *fp, double *fn)
{
    double Q[5];
#pragma unroll
    for (int c = 0; c < 5; c++) Q[c] = qb[c][x];
    double jac = B.jac[x];
    double rho, u, v, w, p, cs;
    con2prim(Q, jac, P, rho, u, v, w, p, cs);
    return split_cell(Q, jac, B.met[3 * d][x], B.met[3 * d + 1][x], B.met[3 * d + 2][x], u, v, w, p, cs, P.riemann,
                      fp, fn);
}

// BC ghost split flux of physical face f from its boundary cell xb
// (boundaryConditions.cpp:140-1086; only the layer next to the face enters
// the closures of reconstructionBoundaryScheme.cpp).
__device__ __forceinline__ void bc_ghost(const nuc3d_block_dev &B, double *const *qb, int f, long long xb,
                                         const PhysParams &P, double *fl, double *fr)
{
    const int d = f / 2, id = B.fid[f];
    double prim[5];
    const double jac = B.jac[xb];
    if (id == BC_INLET) {
#pragma unroll
        for (int c = 0; c < 5; c++) prim[c] = B.fval[f][c];
    } else {
        double Q[5];
#pragma unroll
        for (int c = 0; c < 5; c++) Q[c] = qb[c][xb];
        double cs;
        con2prim(Q, jac, P, prim[0], prim[1], prim[2], prim[3], prim[4], cs);
    }
    if (id == BC_WALL) {
        prim[0] = prim[4] * P.gamma * P.mach * P.mach / P.twall;
        prim[1] = 0.0;
        prim[2] = 0.0;
        prim[3] = 0.0;
    } else if (id == BC_SYMM) {
        const double U = B.met[0][xb] * prim[1] + B.met[1][xb] * prim[2] + B.met[2][xb] * prim[3];
        const double V = B.met[3][xb] * prim[1] + B.met[4][xb] * prim[2] + B.met[5][xb] * prim[3];
        const double W = B.met[6][xb] * prim[1] + B.met[7][xb] * prim[2] + B.met[8][xb] * prim[3];
#pragma unroll
        for (int m = 0; m < 3; m++) {
            const double xx = B.dxyz[3 * m][xb], xe = B.dxyz[3 * m + 1][xb], xz = B.dxyz[3 * m + 2][xb];
            if (d == 0)
                prim[1 + m] = -xx * U + xe * V + xz * W;
            else if (d == 1)
                prim[1 + m] = xx * U - xe * V + xz * W;
            else
                prim[1 + m] = xx * U + xe * V - xz * W;
        }
    }
    split_point(prim, jac, B.met[3 * d][xb], B.met[3 * d + 1][xb], B.met[3 * d + 2][xb], P, fl, fr);
}

// Split fluxes of cell (i,j,k) in direction d, where the cell may lie in the
// ghost frame.  Beyond a physical face only the slot next to the face is
// defined (+ flux at -1, - flux at n); other slots are zero and never read.
__device__ __forceinline__ void split_any(const nuc3d_block_dev &B, double *const *qb, int i, int j, int k, int d,
                                          const PhysParams &P, double *fp, double *fn)
{
    const int c3[3] = {i, j, k};
    const int t = c3[d], n = B.n[d];
    if (t < 0 && B.ftype[2 * d] == -1) {
#pragma unroll
        for (int c = 0; c < 5; c++) fp[c] = fn[c] = 0.0;
        if (t == -1) {
            int cb[3] = {i, j, k};
            cb[d] = 0;
            double tmp[5];
            bc_ghost(B, qb, 2 * d, cidx(B, cb[0], cb[1], cb[2]), P, fp, tmp);
        }
        return;
    }
    if (t >= n && B.ftype[2 * d + 1] == -1) {
#pragma unroll
        for (int c = 0; c < 5; c++) fp[c] = fn[c] = 0.0;
        if (t == n) {
            int cb[3] = {i, j, k};
            cb[d] = n - 1;
            double tmp[5];
            bc_ghost(B, qb, 2 * d + 1, cidx(B, cb[0], cb[1], cb[2]), P, tmp, fn);
        }
        return;
    }
    cell_split(B, qb, cidx(B, i, j, k), d, P, fp, fn);
}

// Face value of direction d at face m from the six-slot stencils sp (F+) and
// sn (F-), with the physical closures and the WENO5-Z quirk of weno5z.cpp:80.
template <int SCHEME>
__device__ __forceinline__ double face_value(const nuc3d_block_dev &B, const nuc3d_kparams &K, int d, int m,
                                             const double *sp, const double *sn)
{
    const int n = B.n[d], w = K.width[d];
    if (B.ftype[2 * d] == -1 && m < w) return face_exterior(sp, true, m) + face_exterior(sn, false, m);
    if (B.ftype[2 * d + 1] == -1 && m > n - w) return face_exterior(sp, true, n - m) + face_exterior(sn, false, n - m);
    const bool zq = SCHEME == SCH_WENO5Z && K.zquirk && m >= w && m <= n - w;
    return face_interior<SCHEME>(sp, true, zq) + face_interior<SCHEME>(sn, false, false);
}

// Derivative of field f along d at cell t of a line (cd6 / exterior closure),
// NaiverStokes3d.cpp:122-168 with fieldOperator.cpp:505-542.
__device__ __forceinline__ double line_diff(const nuc3d_block_dev &B, const nuc3d_kparams &K, const double *f,
                                            int d, int t)
{
    const long long st = dstride(B, d);
    const int n = B.n[d], w = K.width[d];
    if (B.ftype[2 * d] == -1 && t < w) return diff_exterior(f, st, t, true);
    if (B.ftype[2 * d + 1] == -1 && t >= n - w) return diff_exterior(f, st, n - 1 - t, false);
    return cd6(f, st);
}

// Same derivative on a line held in LDS (element stride st along the line).
__device__ __forceinline__ double line_diff_lds(const nuc3d_block_dev &B, const nuc3d_kparams &K, const double *f,
                                                int st, int d, int t)
{
    const int n = B.n[d], w = K.width[d];
    if (B.ftype[2 * d] == -1 && t < w) return diff_exterior(f, st, t, true);
    if (B.ftype[2 * d + 1] == -1 && t >= n - w) return diff_exterior(f, st, n - 1 - t, false);
    return cd6(f, st);
}

// Ghost cell of the neighbour that receives cell (i,j,k) through face f: the
// raw slab copy of bufferData::setBufferSend + MPI_Irecv (bufferData.cpp:56-192,
// MPICommunicator.cpp:55-81).  Face f packs cells [0,G) (even f) or [n-G,n)
// (odd f) in order g; the neighbour stores layer g as ghost g-G (even nf) or
// n+g (odd nf).
__device__ __forceinline__ bool ghost_target(const nuc3d_block_dev *blocks, const nuc3d_block_dev &B, int f, int i,
                                             int j, int k, int &nb, long long &x)
{
    const int d = f / 2;
    const int c3[3] = {i, j, k};
    const int t = c3[d], n = B.n[d], G = B.G;
    int g;
    if (f & 1) {
        if (t < n - G) return false;
        g = t - (n - G);
    } else {
        if (t >= G) return false;
        g = t;
    }
    if (B.ftype[f] < 0) return false;
    nb = B.ftype[f];
    const nuc3d_block_dev &N = blocks[nb];
    const int nf = B.fid[f];
    int o[3] = {i, j, k};
    o[d] = (nf & 1) ? N.n[d] + g : g - G;
    x = cidx(N, o[0], o[1], o[2]);
    return true;
}

__device__ __forceinline__ void push_q(const nuc3d_block_dev *blocks, const nuc3d_block_dev &B, int buf, int i, int j,
                                       int k, const double *vals)
{
#pragma unroll
    for (int f = 0; f < 6; f++) {
        int nb;
        long long x;
        if (ghost_target(blocks, B, f, i, j, k, nb, x)) {
#pragma unroll
            for (int c = 0; c < 5; c++) blocks[nb].q[buf][c][x] = vals[c];
        }
    }
}

__device__ __forceinline__ double warp_max(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

__device__ __forceinline__ void block_max3_atomic(double *m3, unsigned long long *dst)
{
    __shared__ double red[3][4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double v = warp_max(m3[d]);
        if (lane == 0) red[d][wv] = v;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double v = fmax(fmax(red[threadIdx.x][0], red[threadIdx.x][1]), fmax(red[threadIdx.x][2], red[threadIdx.x][3]));
        if (v > 0.0) atomicMax(dst + threadIdx.x, (unsigned long long)__double_as_longlong(v));
    }
}

// dt_global = cfl * MIN_b min_d 1/lambda_{b,d} (euler3d.cpp:429-435, PDEData3d.cpp:78-79)
__device__ __forceinline__ double global_dt(const unsigned long long *lam, int nb, double cfl)
{
    double m = 0.0;
    for (int b = 0; b < nb; b++) {
        double a = 1.0 / __longlong_as_double((long long)lam[3 * b]);
        double bb = 1.0 / __longlong_as_double((long long)lam[3 * b + 1]);
        double cc = 1.0 / __longlong_as_double((long long)lam[3 * b + 2]);
        double mm = (bb < a) ? bb : a;
        double dl = (cc < mm) ? cc : mm;
        m = (b == 0 || dl < m) ? dl : m;
    }
    return m * cfl;
}

}  // namespace

// ---------------------------------------------------------------------------
// Direction sweep.  A workgroup owns 16 lines of one block along direction
// DIR and 15 consecutive cells of each line (16 faces).  It computes the split
// fluxes of the 21 cells of every line that its faces reach (3-cell halo,
// physicsModel.cpp:528-754) into LDS, reconstructs the 16 faces per line
// (one face per thread: WENO5-JS / WENO5-Z or the physical closures), and
// differences them into the RHS (euler3d.cpp:143-241, :368-427):
//   DIR 0: rhs  = dF/dxi
//   DIR 1: rhs += dG/deta
//   DIR 2: rhs += dH/dzeta, then for Navier-Stokes
//          rhs -= dFv/dxi, dGv/deta, dHv/dzeta (NaiverStokes3d.cpp:401-513),
//          then (MODE 1) the TVD-RK3 stage (TVD-RK3.cpp:37-137) into buffer
//          bout with the e<0 fix-up and the ghost push to neighbours.
// Lines run along j at fixed k (DIR 0) and along i otherwise, so that global
// loads and stores of a wavefront touch contiguous i.
// tile: x = block, y = first cell a0 along DIR, z = first line l0, w = third index.
template <int DIR, int SCHEME, bool NS, int MODE>
__global__ __launch_bounds__(256) void KNAME(k_sweep)(const nuc3d_block_dev *__restrict__ blocks,
                                                      const int4 *__restrict__ tiles, nuc3d_kparams K, int stage,
                                                      int bin, int bun, int bout, double *dtbuf,
                                                      const unsigned long long *lam_cur,
                                                      unsigned long long *lam_next, unsigned long long *counters)
{
    constexpr int NL = 16, NF = 16, NC = NF - 1, H = NUC3D_HALO, NSL = NC + 2 * H;  // 21 slots per line
    __shared__ double S[10][NSL * NL];
    __shared__ double HF[5][NF * NL];
    // viscous flux components 1..4 of this direction along the lines (fast
    // variant: each sweep subtracts its own viscous divergence)
    constexpr bool VLDS = NS && !NUC3D_EXACT;
    __shared__ double V[VLDS ? 4 : 1][VLDS ? NSL * NL : 1];

    const int4 t = tiles[blockIdx.x];
    const nuc3d_block_dev &B = blocks[t.x];
    const PhysParams P = phys_of(K);
    const int a0 = t.y, l0 = t.z, cc = t.w;
    const int nc = min(NC, B.n[DIR] - a0);
    const int nlines = min(NL, (DIR == 0 ? B.n[1] : B.n[0]) - l0);
    double *const *qin = B.q[bin];

    // (line, slot) -> LDS offset: slot-fastest for DIR 0, line-fastest otherwise
    auto sidx = [](int line, int slot) { return DIR == 0 ? line * NSL + slot : slot * NL + line; };
    auto hidx = [](int line, int f) { return DIR == 0 ? line * NF + f : f * NL + line; };
    auto cell = [&](int line, int a, int &i, int &j, int &k) {
        if (DIR == 0) { i = a; j = l0 + line; k = cc; }
        else if (DIR == 1) { i = l0 + line; j = a; k = cc; }
        else { i = l0 + line; j = cc; k = a; }
    };

    const int nslots = nc + 2 * H;
    for (int e = threadIdx.x; e < NSL * NL; e += 256) {
        const int line = DIR == 0 ? e / NSL : e % NL;
        const int sl = DIR == 0 ? e % NSL : e / NL;
        if (line < nlines && sl < nslots) {
            int i, j, k;
            cell(line, a0 - H + sl, i, j, k);
            double fp[5], fn[5];
            split_any(B, qin, i, j, k, DIR, P, fp, fn);
            const int o = sidx(line, sl);
#pragma unroll
            for (int c = 0; c < 5; c++) {
                S[c][o] = fp[c];
                S[5 + c][o] = fn[c];
            }
            if (VLDS) {
                const long long x = cidx(B, i, j, k);
#pragma unroll
                for (int c = 0; c < 4; c++) V[c][o] = B.fv[4 * DIR + c][x];
            }
        }
    }
    __syncthreads();

    const int line = DIR == 0 ? threadIdx.x / NF : threadIdx.x % NL;
    const int f = DIR == 0 ? threadIdx.x % NF : threadIdx.x / NL;
    if (line < nlines && f <= nc) {
        const int m = a0 + f;
#pragma unroll
        for (int c = 0; c < 5; c++) {
            double sp[6], sn[6];
#pragma unroll
            for (int q = 0; q < 6; q++) {
                sp[q] = S[c][sidx(line, f + q)];
                sn[q] = S[5 + c][sidx(line, f + q)];
            }
            HF[c][hidx(line, f)] = face_value<SCHEME>(B, K, DIR, m, sp, sn);
        }
    }
    __syncthreads();

    double lmax[3] = {0.0, 0.0, 0.0};
    unsigned long long nfix = 0;
    double dt = 0.0;
    if (DIR == 2 && MODE == 1) {
        if (stage == 0) {
            dt = global_dt(lam_cur, K.nblocks, K.cfl);
            if (blockIdx.x == 0 && threadIdx.x == 0) dtbuf[0] = dt;
            if (blockIdx.x == 0)
                for (int e = threadIdx.x; e < 3 * K.nblocks; e += blockDim.x) lam_next[e] = 0ULL;
        } else {
            dt = dtbuf[0];
        }
    }
    if (line < nlines && f >= 1 && f <= nc) {
        int i, j, k;
        cell(line, a0 + f - 1, i, j, k);
        const long long x = cidx(B, i, j, k);
        double r[5];
#pragma unroll
        for (int c = 0; c < 5; c++) {
            double d = HF[c][hidx(line, f)] - HF[c][hidx(line, f - 1)];
            if (VLDS && c > 0) {
                const int st = DIR == 0 ? 1 : NL;
                d -= line_diff_lds(B, K, &V[c - 1][sidx(line, f + 2)], st, DIR, a0 + f - 1);
            }
            r[c] = DIR == 0 ? d : B.rhs[c][x] + d;
        }
        if (DIR < 2) {
#pragma unroll
            for (int c = 0; c < 5; c++) B.rhs[c][x] = r[c];
        } else {
            if (NS && !VLDS) {  // reference order: rhs = df + dg + dh - dfv - dgv - dhv
#pragma unroll
                for (int c = 1; c < 5; c++) {
                    const double gx = line_diff(B, K, B.fv[c - 1] + x, 0, i);
                    const double gy = line_diff(B, K, B.fv[4 + c - 1] + x, 1, j);
                    const double gz = line_diff(B, K, B.fv[8 + c - 1] + x, 2, k);
                    r[c] = ((r[c] - gx) - gy) - gz;
                }
            }
            if (MODE == 0) {
#pragma unroll
                for (int c = 0; c < 5; c++) B.rhs[c][x] = r[c];
            } else {
                double qn[5];
#pragma unroll
                for (int c = 0; c < 5; c++) {
                    const double rhs = -r[c];
                    const double u = qin[c][x];
                    if (stage == 0)
                        qn[c] = 1.0 * u + 1.0 * rhs * dt;
                    else if (stage == 1)
                        qn[c] = 3.0 / 4.0 * B.q[bun][c][x] + 1.0 / 4.0 * (rhs * dt + u);
                    else
                        qn[c] = 1.0 / 3.0 * B.q[bun][c][x] + 2.0 / 3.0 * (rhs * dt + u);
                }
                const double jac = B.jac[x];
                if (fixup_energy(qn, jac, P)) nfix++;
#pragma unroll
                for (int c = 0; c < 5; c++) B.q[bout][c][x] = qn[c];
                push_q(blocks, B, bout, i, j, k, qn);
                if (stage == 2) {
                    double rho, u, v, w, pr, cs;
                    con2prim(qn, jac, P, rho, u, v, w, pr, cs);
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        const double mx = B.met[3 * d][x], my = B.met[3 * d + 1][x], mz = B.met[3 * d + 2][x];
                        const double theta = sqrt(mx * mx + my * my + mz * mz);
                        const double U0 = (mx * u + my * v + mz * w);
                        const double lam = fabs(U0) + cs * theta;
                        lmax[d] = (lmax[d] < lam) ? lam : lmax[d];
                    }
                }
            }
        }
    }
    if (DIR == 2 && MODE == 1) {
        if (nfix) atomicAdd(counters, nfix);
        if (stage == 2) block_max3_atomic(lmax, lam_next + 3 * t.x);
    }
}

// ---------------------------------------------------------------------------
// Navier-Stokes viscous fluxes (NaiverStokes3d.cpp:63-326): gradients of u, v,
// w, T by cd6 along each direction, Sutherland viscosity, stresses and heat
// flux, contravariant viscous fluxes divided by J.  Writes fv of the interior
// and pushes them into the neighbours' ghost layers.
__global__ __launch_bounds__(256) void KNAME(k_prim)(const nuc3d_block_dev *__restrict__ blocks, int nblocks,
                                                     nuc3d_kparams K, int bin)
{
    const int b = blockIdx.y;
    if (b >= nblocks) return;
    const nuc3d_block_dev &B = blocks[b];
    const PhysParams P = phys_of(K);
    const int G = B.G;
    const long long px = B.n[0] + 2 * G, py = B.n[1] + 2 * G, pz = B.n[2] + 2 * G;
    const long long tot = px * py * pz;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot; e += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(e % px) - G, j = (int)((e / px) % py) - G, k = (int)(e / (px * py)) - G;
        const int outside = (i < 0 || i >= B.n[0]) + (j < 0 || j >= B.n[1]) + (k < 0 || k >= B.n[2]);
        if (outside > 1) continue;  // corner/edge ghosts are never read
        const long long x = cidx(B, i, j, k);
        double Q[5];
#pragma unroll
        for (int c = 0; c < 5; c++) Q[c] = B.q[bin][c][x];
        const double jac = B.jac[x];
        if (jac == 0.0) continue;
        double rho, u, v, w, p, cs;
        con2prim(Q, jac, P, rho, u, v, w, p, cs);
        B.prim[0][x] = u;
        B.prim[1][x] = v;
        B.prim[2][x] = w;
        B.prim[3][x] = P.gamma * P.mach * P.mach * p / rho;  // physicsModel.cpp:1029
    }
}

__global__ __launch_bounds__(256) void KNAME(k_visc_flux)(const nuc3d_block_dev *__restrict__ blocks,
                                                          const int4 *__restrict__ tiles, nuc3d_kparams K)
{
    // one thread per cell on the tiles of the y sweep: 16 i (lines) x 15 j
    const int4 t = tiles[blockIdx.x];
    const nuc3d_block_dev &B = blocks[t.x];
    const int i = t.z + (threadIdx.x & 15), j = t.y + (threadIdx.x >> 4), k = t.w;
    if (i >= B.n[0] || j >= B.n[1] || (threadIdx.x >> 4) >= 15) return;
    const double S = K.sref;
    const long long x = cidx(B, i, j, k);
    double g[4][3];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const double a = line_diff(B, K, B.prim[v] + x, 0, i);
        const double bb = line_diff(B, K, B.prim[v] + x, 1, j);
        const double cc = line_diff(B, K, B.prim[v] + x, 2, k);
        g[v][0] = a * B.met[0][x] + bb * B.met[3][x] + cc * B.met[6][x];
        g[v][1] = a * B.met[1][x] + bb * B.met[4][x] + cc * B.met[7][x];
        g[v][2] = a * B.met[2][x] + bb * B.met[5][x] + cc * B.met[8][x];
    }
    const double T0 = B.prim[3][x];
    const double miu0 = (1.0 + S) / (T0 + S) * pow(T0, 1.5) / K.reynolds;  // physicsModel.cpp:226
    const double coeff0 = miu0 / ((K.gamma - 1.0) * K.mach * K.mach * K.prandtl);
    const double ux = g[0][0], uy = g[0][1], uz = g[0][2];
    const double vx = g[1][0], vy = g[1][1], vz = g[1][2];
    const double wx = g[2][0], wy = g[2][1], wz = g[2][2];
    const double grad = ux + vy + wz;
    const double txx = miu0 * (2.0 * ux - 2.0 / 3.0 * grad);
    const double txy = miu0 * (uy + vx - 2.0 / 3.0 * grad);
    const double txz = miu0 * (uz + wx - 2.0 / 3.0 * grad);
    const double tyy = miu0 * (2.0 * vy - 2.0 / 3.0 * grad);
    const double tyz = miu0 * (vz + wy - 2.0 / 3.0 * grad);
    const double tzz = miu0 * (2.0 * wz - 2.0 / 3.0 * grad);
    const double ttx = coeff0 * g[3][0], tty = coeff0 * g[3][1], ttz = coeff0 * g[3][2];
    const double u = B.prim[0][x], vv = B.prim[1][x], w = B.prim[2][x];
    const double fv[4] = {txx, txy, txz, u * txx + vv * txy + w * txz + ttx};
    const double gv[4] = {txy, tyy, tyz, u * txy + vv * tyy + w * tyz + tty};
    const double hv[4] = {txz, tyz, tzz, u * txz + vv * tyz + w * tzz + ttz};
    const double jac = B.jac[x];
    double out[12];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double mx = B.met[3 * d][x], my = B.met[3 * d + 1][x], mz = B.met[3 * d + 2][x];
#pragma unroll
        for (int c = 0; c < 4; c++) out[4 * d + c] = (mx * fv[c] + my * gv[c] + mz * hv[c]) / jac;
    }
#pragma unroll
    for (int e = 0; e < 12; e++) B.fv[e][x] = out[e];
#pragma unroll
    for (int f = 0; f < 6; f++) {
        int nb;
        long long y;
        if (ghost_target(blocks, B, f, i, j, k, nb, y)) {
#pragma unroll
            for (int e = 0; e < 12; e++) blocks[nb].fv[e][y] = out[e];
        }
    }
}

// ---------------------------------------------------------------------------
// Maximum eigenvalue per block and direction of buffer `buf`
// (physicsModel.cpp:597,673,750), accumulated with atomicMax on the bit
// patterns of positive doubles.
__global__ __launch_bounds__(256) void KNAME(k_eigen)(const nuc3d_block_dev *__restrict__ blocks,
                                                      const nuc3d_tile *__restrict__ tiles, nuc3d_kparams K, int buf,
                                                      unsigned long long *lam)
{
    constexpr int T = NUC3D_TILE;
    const nuc3d_tile tl = tiles[blockIdx.x];
    const nuc3d_block_dev &B = blocks[tl.b];
    const PhysParams P = phys_of(K);
    const int i = tl.i0 + (threadIdx.x & (T - 1)), j = tl.j0 + threadIdx.x / T;
    double lmax[3] = {0.0, 0.0, 0.0};
    if (i < B.n[0] && j < B.n[1]) {
        for (int k = tl.k0; k < tl.k0 + tl.kn; k++) {
            const long long x = cidx(B, i, j, k);
            double Q[5];
#pragma unroll
            for (int c = 0; c < 5; c++) Q[c] = B.q[buf][c][x];
            double rho, u, v, w, p, cs;
            con2prim(Q, B.jac[x], P, rho, u, v, w, p, cs);
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const double mx = B.met[3 * d][x], my = B.met[3 * d + 1][x], mz = B.met[3 * d + 2][x];
                const double theta = sqrt(mx * mx + my * my + mz * mz);
                const double U0 = (mx * u + my * v + mz * w);
                const double l = fabs(U0) + cs * theta;
                lmax[d] = (lmax[d] < l) ? l : lmax[d];
            }
        }
    }
    block_max3_atomic(lmax, lam + 3 * tl.b);
}

// e<0 fix-up of an uploaded state (physicsModel.cpp:322-332)
__global__ __launch_bounds__(256) void KNAME(k_fixup)(const nuc3d_block_dev *__restrict__ blocks,
                                                      const nuc3d_tile *__restrict__ tiles, nuc3d_kparams K, int buf,
                                                      unsigned long long *counters)
{
    constexpr int T = NUC3D_TILE;
    const nuc3d_tile tl = tiles[blockIdx.x];
    const nuc3d_block_dev &B = blocks[tl.b];
    const PhysParams P = phys_of(K);
    const int i = tl.i0 + (threadIdx.x & (T - 1)), j = tl.j0 + threadIdx.x / T;
    if (i >= B.n[0] || j >= B.n[1]) return;
    unsigned long long nfix = 0;
    for (int k = tl.k0; k < tl.k0 + tl.kn; k++) {
        const long long x = cidx(B, i, j, k);
        double Q[5];
#pragma unroll
        for (int c = 0; c < 5; c++) Q[c] = B.q[buf][c][x];
        if (fixup_energy(Q, B.jac[x], P)) {
            B.q[buf][4][x] = Q[4];
            nfix++;
        }
    }
    if (nfix) atomicAdd(counters, nfix);
}

// Pull-style ghost fill of the face slabs of every non-physical face:
// kind 0 = the five fields of state buffer `buf`, kind 1 = J and the nine
// metric terms.  Grid: x over slab cells, y = face, z = block.
__global__ __launch_bounds__(256) void KNAME(k_ghost_fill)(const nuc3d_block_dev *__restrict__ blocks, int kind,
                                                           int buf)
{
    const int f = blockIdx.y, b = blockIdx.z;
    const nuc3d_block_dev &B = blocks[b];
    if (B.ftype[f] < 0) return;
    const int d = f / 2, d2 = (d + 1) % 3, d3 = (d + 2) % 3, G = B.G;
    const long long n2 = B.n[d2], n3 = B.n[d3];
    const long long tot = (long long)G * n2 * n3;
    const nuc3d_block_dev &N = blocks[B.ftype[f]];
    const int nf = B.fid[f];
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot; e += (long long)gridDim.x * blockDim.x) {
        const int g = (int)(e % G), a2 = (int)((e / G) % n2), a3 = (int)(e / (G * n2));
        int o[3], s[3];
        o[d2] = s[d2] = a2;
        o[d3] = s[d3] = a3;
        o[d] = (f & 1) ? B.n[d] + g : g - G;            // receiving ghost cell
        s[d] = (nf & 1) ? N.n[d] - G + g : g;           // sending cell of the neighbour face
        const long long xo = cidx(B, o[0], o[1], o[2]), xs = cidx(N, s[0], s[1], s[2]);
        if (kind == 0) {
#pragma unroll
            for (int c = 0; c < 5; c++) B.q[buf][c][xo] = N.q[buf][c][xs];
        } else {
            B.jac[xo] = N.jac[xs];
#pragma unroll
            for (int c = 0; c < 9; c++) B.met[c][xo] = N.met[c][xs];
        }
    }
}

// ---------------------------------------------------------------------------
// host-side launchers (the only symbols the C-ABI layer uses)
template <int DIR, int SCHEME, bool NS, int MODE>
static void launch_sweep(const nuc3d_block_dev *blocks, const int4 *tiles, int ntiles, nuc3d_kparams K, int stage,
                         int bin, int bun, int bout, double *dtbuf, const unsigned long long *lam_cur,
                         unsigned long long *lam_next, unsigned long long *counters, hipStream_t st)
{
    hipLaunchKernelGGL((KNAME(k_sweep)<DIR, SCHEME, NS, MODE>), dim3(ntiles), dim3(256), 0, st, blocks, tiles, K,
                       stage, bin, bun, bout, dtbuf, lam_cur, lam_next, counters);
}

template <int SCHEME, bool NS>
static void sweep_dispatch(int dir, int mode, const nuc3d_block_dev *blocks, const int4 *tiles, int ntiles,
                           nuc3d_kparams K, int stage, int bin, int bun, int bout, double *dtbuf,
                           const unsigned long long *lam_cur, unsigned long long *lam_next,
                           unsigned long long *counters, hipStream_t st)
{
#define NUC3D_A blocks, tiles, ntiles, K, stage, bin, bun, bout, dtbuf, lam_cur, lam_next, counters, st
    if (dir == 0)
        launch_sweep<0, SCHEME, NS, 0>(NUC3D_A);
    else if (dir == 1)
        launch_sweep<1, SCHEME, NS, 0>(NUC3D_A);
    else if (mode == 0)
        launch_sweep<2, SCHEME, NS, 0>(NUC3D_A);
    else
        launch_sweep<2, SCHEME, NS, 1>(NUC3D_A);
#undef NUC3D_A
}

// dir 0/1 accumulate the RHS; dir 2 completes it (mode 0) or applies the RK stage (mode 1)
extern "C" hipError_t KNAME(nuc3d_launch_sweep)(int dir, int scheme, int ns, int mode, const nuc3d_block_dev *blocks,
                                                const int4 *tiles, int ntiles, nuc3d_kparams K, int stage, int bin,
                                                int bun, int bout, double *dtbuf,
                                                const unsigned long long *lam_cur, unsigned long long *lam_next,
                                                unsigned long long *counters, hipStream_t st)
{
#define NUC3D_A dir, mode, blocks, tiles, ntiles, K, stage, bin, bun, bout, dtbuf, lam_cur, lam_next, counters, st
    if (scheme == SCH_WENO5JS) {
        if (ns) sweep_dispatch<SCH_WENO5JS, true>(NUC3D_A);
        else sweep_dispatch<SCH_WENO5JS, false>(NUC3D_A);
    } else if (scheme == SCH_WENO5Z) {
        if (ns) sweep_dispatch<SCH_WENO5Z, true>(NUC3D_A);
        else sweep_dispatch<SCH_WENO5Z, false>(NUC3D_A);
    } else {
        return hipErrorInvalidValue;
    }
#undef NUC3D_A
    return hipGetLastError();
}

extern "C" hipError_t KNAME(nuc3d_launch_visc)(const nuc3d_block_dev *blocks, int nblocks, const int4 *ytiles,
                                               int nytiles, nuc3d_kparams K, int bin, hipStream_t st)
{
    hipLaunchKernelGGL(KNAME(k_prim), dim3(1024, nblocks), dim3(256), 0, st, blocks, nblocks, K, bin);
    hipLaunchKernelGGL(KNAME(k_visc_flux), dim3(nytiles), dim3(256), 0, st, blocks, ytiles, K);
    return hipGetLastError();
}

extern "C" hipError_t KNAME(nuc3d_launch_eigen)(const nuc3d_block_dev *blocks, const nuc3d_tile *tiles, int ntiles,
                                                nuc3d_kparams K, int buf, unsigned long long *lam, hipStream_t st)
{
    hipLaunchKernelGGL(KNAME(k_eigen), dim3(ntiles), dim3(256), 0, st, blocks, tiles, K, buf, lam);
    return hipGetLastError();
}

extern "C" hipError_t KNAME(nuc3d_launch_fixup)(const nuc3d_block_dev *blocks, const nuc3d_tile *tiles, int ntiles,
                                                nuc3d_kparams K, int buf, unsigned long long *counters,
                                                hipStream_t st)
{
    hipLaunchKernelGGL(KNAME(k_fixup), dim3(ntiles), dim3(256), 0, st, blocks, tiles, K, buf, counters);
    return hipGetLastError();
}

extern "C" hipError_t KNAME(nuc3d_launch_ghost_fill)(const nuc3d_block_dev *blocks, int nblocks, int kind, int buf,
                                                     hipStream_t st)
{
    hipLaunchKernelGGL(KNAME(k_ghost_fill), dim3(256, 6, nblocks), dim3(256), 0, st, blocks, kind, buf);
    return hipGetLastError();
}
