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
#include <algorithm>
#include <type_traits>

#include "nuc3d_device.h"
#include "nuc3d_sensor.h"
#include "nuc3d_types.h"

#if NUC3D_EXACT
#define KNAME(name) name##_exact
#else
#define KNAME(name) name##_fast
#endif

using namespace nuc3d_dev;

namespace {

// physics of block B: the case's, with the block's (its case's) Mach and
// Reynolds numbers
__device__ __forceinline__ PhysParams phys_of(const nuc3d_kparams &K, const nuc3d_block_dev &B)
{
    PhysParams P;
    P.gamma = K.gamma;
    P.mach = B.mach;
    P.reynolds = B.reynolds;
    P.prandtl = K.prandtl;
    P.sref = K.sref;
    P.twall = K.twall;
    P.riemann = K.riemann;
    return P;
}

// The same for a kernel reading the state of buffer bin: the divisions by rho
// of its con2prim share one reciprocal where every cell of that buffer passed
// momenta_in_range when it was written (K.qok[bin], kept by the RK stages)
__device__ __forceinline__ PhysParams phys_of(const nuc3d_kparams &K, const nuc3d_block_dev &B, int bin)
{
    PhysParams P = phys_of(K, B);
    P.rdq = K.qok && !K.ieee_div && gldw(K.qok + bin) != 0u;
    return P;
}

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X, each
// with its own L2.  Tile lists are ordered so that neighbouring tiles (whose
// halos overlap) are adjacent; this gives XCD x the contiguous range of tiles
// [x*per, ...) instead of every eighth tile, so halo lines are shared in L2.
__device__ __forceinline__ int xcd_tile(int w, int nw)
{
    constexpr int NXCD = 8;
    const int per = nw / NXCD, rem = nw % NXCD, x = w % NXCD, l = w / NXCD;
    return x < rem ? x * (per + 1) + l : rem * (per + 1) + (x - rem) * per + l;
}

// Work items of a direction sweep.  A launch of one workgroup per tile
// (K.pq null) takes tile xcd_tile(blockIdx.x) once.  A persistent launch
// (K.pq: eight queue counters and an exit counter, zero at launch; a grid of
// as many workgroups as the GPU holds at once) takes tiles until none is
// left: the workgroups of XCD x draw from x's contiguous range of the tile
// list first (neighbouring tiles share that XCD's L2), then from the other
// ranges in turn.  Thread 0 keeps the raw counter value of the next draw
// in flight while the current tile is computed (the atomic's round trip is
// not waited for per tile); a range found drained is left for the next one.
// Every thread of the workgroup calls next(); it returns the same tile (or
// -1) to all of them.  sw: one word of LDS.
struct WorkQueue {
    int q = 0, raw = 0;  // thread 0: ranges drained so far, counter value of the pending draw
    __device__ __forceinline__ int next(const nuc3d_kparams &K, int it, int *sw)
    {
        if (!K.pq) return it == 0 ? xcd_tile(blockIdx.x, gridDim.x) : -1;
        constexpr int NXCD = 8;
        if (threadIdx.x == 0) {
            const int n = K.pq_n, per = n / NXCD, rem = n % NXCD, x0 = blockIdx.x % NXCD;
            if (it == 0) raw = (int)atomicAdd(K.pq + x0, 1u);
            int w = -1;
            while (w < 0 && q < NXCD) {
                const int x = (x0 + q) % NXCD;
                const int cnt = per + (x < rem ? 1 : 0);
                const int start = x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per;
                if (raw < cnt) {
                    w = start + raw;
                } else if (++q < NXCD) {
                    raw = (int)atomicAdd(K.pq + (x0 + q) % NXCD, 1u);
                }
            }
            if (w >= 0) raw = (int)atomicAdd(K.pq + (x0 + q) % NXCD, 1u);  // the next draw, in flight
            sw[0] = w;
        }
        __syncthreads();
        return sw[0];
    }
};

// End of a persistent launch: the last workgroup to finish (every other one
// has drawn its last item) zeroes the queue and exit counters for the next
// launch on the stream.
__device__ __forceinline__ void work_done(const nuc3d_kparams &K)
{
    if (!K.pq || threadIdx.x != 0) return;
    __threadfence();
    if (atomicAdd(K.pq + 8, 1u) == gridDim.x - 1)
        for (int x = 0; x < 9; x++) K.pq[x] = 0u;
}

__device__ __forceinline__ long long cidx(const nuc3d_block_dev &B, int i, int j, int k)
{
    return (long long)i + B.sy * j + B.sz * k;
}

__device__ __forceinline__ long long dstride(const nuc3d_block_dev &B, int d)
{
    return d == 0 ? 1LL : (d == 1 ? B.sy : B.sz);
}

// J and metric term e of cell x (uniform blocks: the block's constants, the
// same values as the fields hold)
__device__ __forceinline__ double jac_at(const nuc3d_block_dev &B, long long x)
{
    return B.uniform ? B.jac0 : gld(B.jac + x);
}

__device__ __forceinline__ double met_at(const nuc3d_block_dev &B, int e, long long x)
{
    return B.uniform ? B.met0[e] : gld(B.met[e] + x);
}

// split fluxes of cell x of buffer qb in direction d; returns the local
// eigenvalue.  prim (may be null) receives rho, u, v, w, p.  A uniform
// block's |grad xi| is its descriptor's constant, and where J, the metric
// terms and p allow it the divisions by J share one reciprocal (same bits).
__device__ __forceinline__ double cell_split(const nuc3d_block_dev &B, double *const *qb, long long x, int d,
                                             const PhysParams &P, double *fp, double *fn, double *prim = nullptr,
                                             double *qout = nullptr)
{
    double Q[5];
#pragma unroll
    for (int c = 0; c < 5; c++) Q[c] = gld(qb[c] + x);
    if (qout)
#pragma unroll
        for (int c = 0; c < 5; c++) qout[c] = Q[c];
    double jac = jac_at(B, x);
    double rho, u, v, w, p, cs;
    con2prim(Q, jac, P, rho, u, v, w, p, cs);
    if (prim) {
        prim[0] = rho;
        prim[1] = u;
        prim[2] = v;
        prim[3] = w;
        prim[4] = p;
    }
    const double mx = met_at(B, 3 * d, x), my = met_at(B, 3 * d + 1, x), mz = met_at(B, 3 * d + 2, x);
    const double theta = B.uniform ? B.theta0[d] : sqrt(mx * mx + my * my + mz * mz);
    const double yj = (B.uniform && B.rdsplit && rd_operand(p)) ? B.rjac0 : 0.0;
    return split_cell(Q, jac, mx, my, mz, theta, u, v, w, p, cs, P.riemann, fp, fn, yj);
}

// BC ghost split flux of physical face f from its boundary cell xb
// (boundaryConditions.cpp:140-1086; only the layer next to the face enters
// the closures of reconstructionBoundaryScheme.cpp).
__device__ __forceinline__ void bc_ghost(const nuc3d_block_dev &B, double *const *qb, int f, long long xb,
                                         const PhysParams &P, double *fl, double *fr)
{
    const int d = f / 2, id = B.fid[f];
    double prim[5];
    const double jac = jac_at(B, xb);
    if (id == BC_INLET) {
#pragma unroll
        for (int c = 0; c < 5; c++) prim[c] = B.fval[f][c];
    } else {
        double Q[5];
#pragma unroll
        for (int c = 0; c < 5; c++) Q[c] = gld(qb[c] + xb);
        double cs;
        con2prim(Q, jac, P, prim[0], prim[1], prim[2], prim[3], prim[4], cs);
    }
    if (id == BC_WALL) {
        prim[0] = prim[4] * P.gamma * P.mach * P.mach / P.twall;
        prim[1] = 0.0;
        prim[2] = 0.0;
        prim[3] = 0.0;
    } else if (id == BC_SYMM) {
        const double U = met_at(B, 0, xb) * prim[1] + met_at(B, 1, xb) * prim[2] + met_at(B, 2, xb) * prim[3];
        const double V = met_at(B, 3, xb) * prim[1] + met_at(B, 4, xb) * prim[2] + met_at(B, 5, xb) * prim[3];
        const double W = met_at(B, 6, xb) * prim[1] + met_at(B, 7, xb) * prim[2] + met_at(B, 8, xb) * prim[3];
#pragma unroll
        for (int m = 0; m < 3; m++) {
            const double xx = gld(B.dxyz[3 * m] + xb), xe = gld(B.dxyz[3 * m + 1] + xb), xz = gld(B.dxyz[3 * m + 2] + xb);
            if (d == 0)
                prim[1 + m] = -xx * U + xe * V + xz * W;
            else if (d == 1)
                prim[1 + m] = xx * U - xe * V + xz * W;
            else
                prim[1 + m] = xx * U + xe * V - xz * W;
        }
    }
    split_point(prim, jac, met_at(B, 3 * d, xb), met_at(B, 3 * d + 1, xb), met_at(B, 3 * d + 2, xb), P, fl, fr);
}

// Split fluxes of cell (i,j,k) in direction d, where the cell may lie in the
// ghost frame.  Beyond a physical face only the slot next to the face is
// defined (+ flux at -1, - flux at n); other slots are zero and never read.
// prim (may be null) receives the cell's rho, u, v, w, p (zero beyond a
// physical face).
__device__ __forceinline__ void split_any(const nuc3d_block_dev &B, double *const *qb, int i, int j, int k, int d,
                                          const PhysParams &P, double *fp, double *fn, double *prim = nullptr,
                                          double *qout = nullptr)
{
    const int c3[3] = {i, j, k};
    const int t = c3[d], n = B.n[d];
    if (prim && ((t < 0 && B.ftype[2 * d] == -1) || (t >= n && B.ftype[2 * d + 1] == -1)))
#pragma unroll
        for (int c = 0; c < 5; c++) prim[c] = 0.0;
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
    cell_split(B, qb, cidx(B, i, j, k), d, P, fp, fn, prim, qout);
}

// Face value of direction d at face m from the six-slot stencils sp (F+) and
// sn (F-), with the physical closures and the WENO5-Z quirk of weno5z.cpp:80.
template <int SCHEME, bool RD = false>
__device__ __forceinline__ double face_value(const nuc3d_block_dev &B, const nuc3d_kparams &K, int d, int m,
                                             const double *sp, const double *sn)
{
    const int n = B.n[d], w = K.width[d];
    if (B.ftype[2 * d] == -1 && m < w) return face_exterior(sp, true, m) + face_exterior(sn, false, m);
    if (B.ftype[2 * d + 1] == -1 && m > n - w) return face_exterior(sp, true, n - m) + face_exterior(sn, false, n - m);
    const bool zq = SCHEME == SCH_WENO5Z && K.zquirk && m >= w && m <= n - w;
    return face_interior<SCHEME, RD>(sp, true, zq) + face_interior<SCHEME, RD>(sn, false, false);
}

// Derivative of field f along d at cell t of a line (cd6 or cd8 of the
// direction, or the exterior closure), NaiverStokes3d.cpp:122-168 with
// fieldOperator.cpp:505-542; cd8 (centraldifference8th.cpp:49-52) reads four
// cells each side, which the width-4 cases hold.
// WIDE: the kernel is built for cases with cd8 directions (kernels of cd6-only
// cases leave the cd8 code out, which keeps their registers down).
template <bool WIDE = true>
__device__ __forceinline__ double line_diff(const nuc3d_block_dev &B, const nuc3d_kparams &K, const double *f,
                                            int d, int t)
{
    const long long st = dstride(B, d);
    const int n = B.n[d], w = K.width[d];
    const gdouble *g = gp(f);
    if (B.ftype[2 * d] == -1 && t < w) return diff_exterior(g, st, t, true);
    if (B.ftype[2 * d + 1] == -1 && t >= n - w) return diff_exterior(g, st, n - 1 - t, false);
    if constexpr (WIDE)
        if (K.diff[d]) return cd8(g, st);
    return cd6(g, st);
}

// Same derivative on a line held in LDS (element stride st along the line).
template <bool WIDE = true>
__device__ __forceinline__ double line_diff_lds(const nuc3d_block_dev &B, const nuc3d_kparams &K, const double *f,
                                                int st, int d, int t)
{
    const int n = B.n[d], w = K.width[d];
    if (B.ftype[2 * d] == -1 && t < w) return diff_exterior(f, st, t, true);
    if (B.ftype[2 * d + 1] == -1 && t >= n - w) return diff_exterior(f, st, n - 1 - t, false);
    if constexpr (WIDE)
        if (K.diff[d]) return cd8(f, st);
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
            for (int c = 0; c < 5; c++) gst(blocks[nb].q[buf][c] + x, vals[c]);
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

// Per-workgroup sum of v, stored at *dst by thread 0 (the residual partials
// of one work item; the host adds them per block in a fixed order).
__device__ __forceinline__ void block_sum_store(double v, double *dst)
{
    __shared__ double red[4];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) red[wv] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += red[w];
        *dst = t;
    }
}

// dt of group g = cfl * MIN over its blocks b of min_d 1/lambda_{b,d}
// (euler3d.cpp:429-435, PDEData3d.cpp:78-79: the MPI_MIN over the ranks of one
// case; with one group, every block)
__device__ __forceinline__ double group_dt(const nuc3d_block_dev *blocks, const unsigned long long *lam, int nb, int g,
                                          double cfl)
{
    double m = 0.0;
    bool first = true;
    for (int b = 0; b < nb; b++) {
        if (blocks[b].group != g) continue;
        double a = 1.0 / __longlong_as_double((long long)lam[3 * b]);
        double bb = 1.0 / __longlong_as_double((long long)lam[3 * b + 1]);
        double cc = 1.0 / __longlong_as_double((long long)lam[3 * b + 2]);
        double mm = (bb < a) ? bb : a;
        double dl = (cc < mm) ? cc : mm;
        m = (first || dl < m) ? dl : m;
        first = false;
    }
    return m * cfl;
}

// dt of the stage for a block of group g: stage 0 takes cfl * min 1/lambda
// over the group from the eigenvalue maxima of the stage-0 state (and
// workgroup 0 publishes every group's dt and clears the maxima the coming
// stage 2 accumulates); stages 1 and 2 reuse it.
__device__ __forceinline__ double stage_dt(const nuc3d_block_dev *blocks, const nuc3d_kparams &K, int stage, int g,
                                          double *dtbuf, const unsigned long long *lam_cur,
                                          unsigned long long *lam_next)
{
    // the buffer the next stage writes (stage 0 writes 1, 1 writes 2, 2 writes
    // 0) is not read before that stage has rewritten all of it: its range flag
    // is set here and cleared there by any cell out of range
    if (K.qok && blockIdx.x == 0 && threadIdx.x == 0) gstw(K.qok + (stage == 0 ? 2 : stage - 1), 1u);
    if (stage != 0) return dtbuf[g];
    const double dt = group_dt(blocks, lam_cur, K.nblocks, g, K.cfl);
    if (blockIdx.x == 0)
        for (int g2 = threadIdx.x; g2 < K.ngroups; g2 += blockDim.x)
            dtbuf[g2] = g2 == g ? dt : group_dt(blocks, lam_cur, K.nblocks, g2, K.cfl);
    if (blockIdx.x == 0)
        for (int e = threadIdx.x; e < 3 * K.nblocks; e += blockDim.x) lam_next[e] = 0ULL;
    return dt;
}

// TVD-RK3 stage of cell x from RHS r (TVD-RK3.cpp:37-137) with the cell's
// current state u5, u_n (un5, stages 1 and 2) and J already loaded; e<0
// fix-up (physicsModel.cpp:322-332), ghost push, and the eigenvalue maxima of
// the new state at stage 2 (physicsModel.cpp:557-597).  qok (may be null):
// the range flag of buffer bout, cleared if the new state is out of the range
// of momenta_in_range or its gamma p out of gamma_p_in_range's; rd: this kernel
// may use the reduced division (then so do the fix-up and the eigenvalues of
// a state in range).
__device__ __forceinline__ void rk_update_vals(const nuc3d_block_dev *blocks, const nuc3d_block_dev &B,
                                               const PhysParams &P, int i, int j, int k, long long x, const double *r,
                                               const double *u5, const double *un5, double jac, int stage, int bout,
                                               double dt, double *lmax, StageCounts &nfix, double &r2,
                                               unsigned *qok, bool rd)
{
    double qn[5];
#pragma unroll
    for (int c = 0; c < 5; c++) {
        const double rhs = -r[c];
        const double u = u5[c];
        if (stage == 0)
            qn[c] = 1.0 * u + 1.0 * rhs * dt;
        else if (stage == 1)
            qn[c] = 3.0 / 4.0 * un5[c] + 1.0 / 4.0 * (rhs * dt + u);
        else
            qn[c] = 1.0 / 3.0 * un5[c] + 2.0 / 3.0 * (rhs * dt + u);
    }
    // residual of the step: the new state before the fix-up of the next
    // con2prim, against u_n (PDEData3d.cpp:70-122)
    if (stage == 2)
#pragma unroll
        for (int c = 0; c < 5; c++) r2 += (qn[c] - un5[c]) * (qn[c] - un5[c]);
    const bool inr = qok && momenta_in_range(qn, qn[0] * jac);
    double pw = 0.0;
    fixup_energy(qn, jac, P, nfix, rd && inr, &pw);
    PhysParams Pn = P;
    Pn.rdq = inr && gamma_p_in_range(P.gamma * pw);
    if (qok && !Pn.rdq) gstw(qok, 0u);
    Pn.rdq = Pn.rdq && rd;
#pragma unroll
    for (int c = 0; c < 5; c++) gst(B.q[bout][c] + x, qn[c]);
    push_q(blocks, B, bout, i, j, k, qn);
    if (stage == 2) {
        double rho, u, v, w, pr, cs;
        con2prim(qn, jac, Pn, rho, u, v, w, pr, cs);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const double mx = met_at(B, 3 * d, x), my = met_at(B, 3 * d + 1, x), mz = met_at(B, 3 * d + 2, x);
            const double theta = B.uniform ? B.theta0[d] : sqrt(mx * mx + my * my + mz * mz);
            const double U0 = (mx * u + my * v + mz * w);
            const double lam = fabs(U0) + cs * theta;
            lmax[d] = (lmax[d] < lam) ? lam : lmax[d];
        }
    }
}

// The same, loading the cell's state from buffers bin / bun.
__device__ __forceinline__ void rk_update(const nuc3d_block_dev *blocks, const nuc3d_block_dev &B,
                                          const nuc3d_kparams &K, const PhysParams &P, int i, int j, int k,
                                          long long x, const double *r, int stage, int bin, int bun, int bout,
                                          double dt, double *lmax, StageCounts &nfix, double &r2)
{
    double u5[5], un5[5];
#pragma unroll
    for (int c = 0; c < 5; c++) {
        u5[c] = gld(B.q[bin][c] + x);
        un5[c] = stage == 0 ? 0.0 : gld(B.q[bun][c] + x);
    }
    rk_update_vals(blocks, B, P, i, j, k, x, r, u5, un5, jac_at(B, x), stage, bout, dt, lmax, nfix, r2,
                   K.qok ? K.qok + bout : nullptr, !K.ieee_div);
}

// Cell (i,j,k) receives the flux difference dF of direction DIR:
//   DIR 0: rhs = dF;  DIR 1, 2: rhs += dF (euler3d.cpp:383-422).  For Euler
//   the z sweep then writes the RHS (MODE 0) or applies the RK stage; for
//   Navier-Stokes k_visc_div finishes the cell of a single RHS evaluation and
//   of the CRWENO5 z lines (the stages of the other z sweeps finish their
//   cells in finish_cell_visc_z).
template <int DIR, bool NS, int MODE>
__device__ __forceinline__ void finish_cell(const nuc3d_block_dev *blocks, const nuc3d_block_dev &B,
                                            const nuc3d_kparams &K, const PhysParams &P, int i, int j, int k,
                                            const double *dF, int stage, int bin, int bun, int bout, double dt,
                                            double *lmax, StageCounts &nfix, double &r2)
{
    const long long x = cidx(B, i, j, k);
    double r[5];
#pragma unroll
    for (int c = 0; c < 5; c++) r[c] = DIR == 0 ? dF[c] : gld(B.rhs[c] + x) + dF[c];
    if (DIR < 2 || NS || MODE == 0) {  // Navier-Stokes: k_visc_div completes the cell
#pragma unroll
        for (int c = 0; c < 5; c++) gst(B.rhs[c] + x, r[c]);
        return;
    }
    rk_update(blocks, B, K, P, i, j, k, x, r, stage, bin, bun, bout, dt, lmax, nfix, r2);
}

// Navier-Stokes stages: the z sweep also finishes its cells with the viscous
// divergence and the RK stage.  Cell (i,j,k) of the sweep's tile (16 lines
// i = l0..l0+15, nc <= 15 cells k = a0..a0+nc-1, j fixed) takes
//   r = rhs + dH/dzeta;  r[1+v] = ((r[1+v] - dFv/dxi) - dGv/deta) - dHv/dzeta
// (NaiverStokes3d.cpp:401-513, the order of the separate divergence pass), then
// the TVD-RK3 stage.  The viscous fluxes are read along i, j and k from memory
// (neighbouring threads and tiles share them in L1 and L2); the cell's state
// u5 comes from the thread's own staging of that cell.
template <bool WIDE>
__device__ __forceinline__ void finish_cell_visc_z(const nuc3d_block_dev *blocks, const nuc3d_block_dev &B,
                                                   const nuc3d_kparams &K, const PhysParams &P, const double *u5,
                                                   int i, int j, int k, const double *dF, int stage, int bun, int bout,
                                                   double dt, double *lmax, StageCounts &nfix, double &r2)
{
    const long long x = cidx(B, i, j, k);
    double r[5];
#pragma unroll
    for (int c = 0; c < 5; c++) r[c] = gld(B.rhs[c] + x) + dF[c];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const double gx = line_diff<WIDE>(B, K, B.fv[v] + x, 0, i);
        const double gy = line_diff<WIDE>(B, K, B.fv[4 + v] + x, 1, j);
        const double gz = line_diff<WIDE>(B, K, B.fv[8 + v] + x, 2, k);
        r[1 + v] = ((r[1 + v] - gx) - gy) - gz;
    }
    double un5[5];
#pragma unroll
    for (int c = 0; c < 5; c++) un5[c] = stage == 0 ? 0.0 : gld(B.q[bun][c] + x);
    rk_update_vals(blocks, B, P, i, j, k, x, r, u5, un5, jac_at(B, x), stage, bout, dt, lmax, nfix, r2,
                   K.qok ? K.qok + bout : nullptr, !K.ieee_div);
}

}  // namespace

// ---------------------------------------------------------------------------
// Direction sweep.  A workgroup owns 16 lines of one block along direction
// DIR and 15 consecutive cells of each line (16 faces).  It computes the split
// fluxes of the 21 cells of every line that its faces reach (3-cell halo,
// physicsModel.cpp:528-754; 23 cells and a 4-cell halo for the eighth-order
// central scheme) into LDS, reconstructs the 16 faces per line (one face per
// thread: WENO5-JS / WENO5-Z / central or the physical closures), and
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
// NF faces per line and 256 / NF lines per workgroup: 16 x 16 (the default)
// or 32 faces x 8 lines (fewer halo cells staged per face; the autotuner picks
// per block shape and direction).
// The Navier-Stokes z sweep (viscous divergence and RK stage) is built for
// five waves per SIMD: 91 VGPRs, where the compiler's own choice (94-97)
// would drop it to four once the RK stage's range check is added.
#define NUC3D_SWEEP_WAVES(NS, DIR, MODE) amdgpu_waves_per_eu((NS) && (DIR) == 2 && (MODE) == 1 ? 5 : 1)
template <int DIR, int SCHEME, bool NS, int MODE, int NF = 16, bool WIDE = false, bool PERS = false>
__global__ __launch_bounds__(256) __attribute__((NUC3D_SWEEP_WAVES(NS, DIR, MODE))) void KNAME(k_sweep)(
    const nuc3d_block_dev *__restrict__ blocks, const int4 *__restrict__ tiles, nuc3d_kparams K, int stage, int bin,
    int bun, int bout, double *dtbuf, const unsigned long long *lam_cur, unsigned long long *lam_next,
    unsigned long long *counters)
{
    // 3-cell halo (21 slots per line) for WENO5; the central scheme reaches 4 cells (23 slots)
    constexpr int H = SCHEME == SCH_CENTRAL8 ? 4 : NUC3D_HALO;
    constexpr int NL = 256 / NF, NC = NF - 1, NSL = NC + 2 * H, NST = 2 * H;
    static_assert(NL * NF == 256, "one face per thread");
    constexpr bool VISC = NS && DIR == 2 && MODE == 1;  // Navier-Stokes stage: viscous divergence here
    __shared__ double S[10][NSL * NL];
    // face fluxes: once every face is computed the split fluxes are dead and
    // their area takes them (27 KB of LDS per workgroup instead of 37: six
    // workgroups fit a CU's LDS instead of four)
    static_assert(10 * NSL >= 5 * NF, "the face fluxes fit the split-flux area");
    double(*HF)[NF * NL] = reinterpret_cast<double(*)[NF * NL]>(&S[0][0]);

    // one tile: witem of the tile list, tid = threadIdx.x
    auto tile = [&](const int witem, const int tid) __attribute__((always_inline)) {
        const int4 t = tiles[witem];
        const nuc3d_block_dev &B = blocks[t.x];
        const PhysParams P = phys_of(K, B, bin);
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
        // whether every staged split flux is in the range of the reduced division
        // (then the workgroup's WENO weights use it; same quotients either way)
        bool rd_ok = true;
        auto stage_slot = [&](int line, int sl, double *qout) {
            int i, j, k;
            cell(line, a0 - H + sl, i, j, k);
            double fp[5], fn[5];
            split_any(B, qin, i, j, k, DIR, P, fp, fn, nullptr, qout);
            const int o = sidx(line, sl);
    #pragma unroll
            for (int c = 0; c < 5; c++) {
                S[c][o] = fp[c];
                S[5 + c][o] = fn[c];
                rd_ok = rd_ok && rd_staged<SCHEME>(fp[c]) && rd_staged<SCHEME>(fn[c]);
            }
        };
        double u5[5] = {0.0, 0.0, 0.0, 0.0, 0.0};  // viscous z sweep: the state of the thread's own cell
        if constexpr (VISC) {
            // the thread that finishes cell f-1 of a line stages it (and keeps its
            // state for the RK stage); then the 2H halo slots of every line
            const int line = tid % NL, f = tid / NL;
            if (f >= 1 && f <= nc && line < nlines) stage_slot(line, H + f - 1, u5);
            for (int e = tid; e < 2 * H * NL; e += 256) {
                const int ln = e % NL, r = e / NL, sl = r < H ? r : nc + r;
                if (ln < nlines) stage_slot(ln, sl, nullptr);
            }
        } else {
            for (int e = tid; e < NSL * NL; e += 256) {
                const int line = DIR == 0 ? e / NSL : e % NL;
                const int sl = DIR == 0 ? e % NSL : e / NL;
                if (line < nlines && sl < nslots) stage_slot(line, sl, nullptr);
            }
        }
        const bool rd = __syncthreads_and(rd_ok) != 0 && !K.ieee_div;

        const int line = DIR == 0 ? tid / NF : tid % NL;
        const int f = DIR == 0 ? tid % NF : tid / NL;
        const bool has_face = line < nlines && f <= nc;
        double hv[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (has_face) {
            const int m = a0 + f;
    #pragma unroll
            for (int c = 0; c < 5; c++) {
                // slots f..f+NST-1 = cells m-H..m+H-1; the six-cell WENO stencil starts at cell m-3
                double sp[NST], sn[NST];
    #pragma unroll
                for (int q = 0; q < NST; q++) {
                    sp[q] = S[c][sidx(line, f + q)];
                    sn[q] = S[5 + c][sidx(line, f + q)];
                }
                hv[c] = rd ? face_value<SCHEME, true>(B, K, DIR, m, sp + (H - 3), sn + (H - 3))
                           : face_value<SCHEME, false>(B, K, DIR, m, sp + (H - 3), sn + (H - 3));
            }
        }
        __syncthreads();  // every face has read the split fluxes
        if (has_face)
    #pragma unroll
            for (int c = 0; c < 5; c++) HF[c][hidx(line, f)] = hv[c];
        __syncthreads();

        double lmax[3] = {0.0, 0.0, 0.0};
        StageCounts nfix;
        double r2 = 0.0;  // residual partial (stage 2)
        const double dt = (DIR == 2 && MODE == 1) ? stage_dt(blocks, K, stage, B.group, dtbuf, lam_cur, lam_next) : 0.0;
        if (line < nlines && f >= 1 && f <= nc) {
            int i, j, k;
            cell(line, a0 + f - 1, i, j, k);
            double dF[5];
    #pragma unroll
            for (int c = 0; c < 5; c++) dF[c] = HF[c][hidx(line, f)] - HF[c][hidx(line, f - 1)];
            if constexpr (VISC)
                finish_cell_visc_z<WIDE>(blocks, B, K, P, u5, i, j, k, dF, stage, bun, bout, dt, lmax, nfix, r2);
            else
                finish_cell<DIR, NS, MODE>(blocks, B, K, P, i, j, k, dF, stage, bin, bun, bout, dt, lmax, nfix, r2);
        }
        if (DIR == 2 && MODE == 1) {
            nfix.flush(counters);
            if (stage == 2) block_max3_atomic(lmax, lam_next + 3 * t.x);
            if (stage == 2 && K.resid) block_sum_store(r2, K.resid + witem);
        }    };
    if constexpr (PERS) {
        // a persistent launch over the device work queue, one tile per
        // iteration; the thread index is opaque to the compiler per tile, so
        // values derived from it (LDS addresses of the staging and the faces)
        // are not kept live across tiles
        __shared__ int SW[1];
        WorkQueue wq;
        for (int it = 0;; it++) {
            const int witem = wq.next(K, it, SW);
            if (witem < 0) break;
            int tid = threadIdx.x;
            asm volatile("" : "+v"(tid));
            tile(witem, tid);
            __syncthreads();  // the next tile reuses the LDS
        }
        work_done(K);
    } else {
        tile(xcd_tile(blockIdx.x, gridDim.x), threadIdx.x);
    }
}


// ---------------------------------------------------------------------------
// Hybrid direction sweep (no counterpart in the reference): the tile, staging
// and differencing of k_sweep with a 4-cell halo, and per interface a choice
// between the eighth-order central face value and WENO5-JS made by the
// learned sensor (nuc3d_sensor.h), one decision per face for all five
// components and both split parts, so the flux of a face is still computed
// once and differenced (conservation is kept).
//   1. split fluxes and primitives of 23 cells x 16 lines into LDS;
//   2. every thread: the 30 features of its face, the positivity check of
//      the central face values of rho and p;
//   3. every wavefront: the network on its 64 faces, four MFMA tiles of 16;
//   (2 and 3 run where the decisions are made (DECIDE): in a single RHS
//   evaluation and in stage 0 of a step; stages 1 and 2 read the step's
//   decisions back from B.hflag[DIR]);
//   4. central faces and physical-face closures by their own thread; the
//      WENO5-JS faces are compacted into a list and their (face, component)
//      items spread over the whole workgroup;
//   5. flux differences, and the RK stage as in k_sweep.
// Decisions are counted when made (K.hcount: central, WENO5-JS by the
// network, WENO5-JS by the fallback; one atomic per workgroup and count,
// spread over slots) and kept per face in B.hflag[DIR]: 0 central, 1 network,
// 2 fallback, 3 physical-face closure.
template <int DIR, bool NS, int MODE, bool DECIDE, bool WIDE = false, int NFH = 16>
__global__ __launch_bounds__(256) __attribute__((NUC3D_SWEEP_WAVES(NS, DIR, MODE))) void KNAME(k_sweep_hybrid)(
    const nuc3d_block_dev *__restrict__ blocks, const int4 *__restrict__ tiles, nuc3d_kparams K, int stage, int bin,
    int bun, int bout, double *dtbuf, const unsigned long long *lam_cur, unsigned long long *lam_next,
    unsigned long long *counters)
{
    using namespace nuc3d_sensor;
    constexpr int H = 4;
    // NFH faces per line and 256 / NFH lines per workgroup: 16 x 16, or 32 x 8
    // for the x and y sweeps (fewer halo cells staged per face)
    constexpr int NF = NFH, NL = 256 / NF, NC = NF - 1, NSL = NC + 2 * H, NST = 2 * H;
    static_assert(NL * NF == 256, "one face per thread");
    __shared__ double S[10][NSL * NL];
    // Where decisions are made, one area has three lives: the FP32 roundings
    // of rho, u, v, w, p of the staged cells; then the bf16 features of the
    // 256 faces, 15 words (30 features) per face, rows of 60 B (an odd number
    // of words: consecutive rows start on different banks), whose first word
    // then takes the face's logit (a wavefront reads and writes only the rows
    // of its own faces); then the face fluxes and the list of WENO5-JS faces.
    // 32-face tiles: 40.4 KB of LDS per workgroup, four per CU.  Where
    // decisions are reused, the face fluxes go to the split-flux area (below).
    struct LdsDecide {
        union {
            float pr[5][NSL * NL];
            unsigned xf[256 * 15];
            struct {
                double hf[5][NF * NL];
                unsigned short lst[256];
            } fl;
        } u;
    };
    struct LdsNone {
        union {
            double hf[1][1];
        } u;
    };
    // Where decisions are reused, the face fluxes take the split-flux area
    // once every face is computed (five workgroups per CU), after the staged
    // fv_xi in the viscous z sweep.
    constexpr bool VISC = NS && DIR == 2 && MODE == 1;  // Navier-Stokes stage: viscous divergence here
    constexpr bool HF_IN_S = !DECIDE;
    static_assert(10 * NSL >= 5 * NF, "the face fluxes fit the split-flux area");
    __shared__ __align__(16) std::conditional_t<DECIDE, LdsDecide, LdsNone> U;
    double(*HF)[NF * NL] = nullptr;
    float(*PR)[NSL * NL] = nullptr;
    unsigned *XF = nullptr;
    if constexpr (DECIDE) {
        HF = U.u.fl.hf;
        PR = U.u.pr;
        XF = U.u.xf;
    } else {
        HF = reinterpret_cast<double(*)[NF * NL]>(&S[0][0]);
    }
    // Where decisions are reused, the FP32 roundings of rho and p of the
    // staged cells, for the positivity check every stage repeats; the list of
    // WENO5-JS faces takes the same area once that check is made.
    struct LdsReuse {
        union {
            float prp[2][NSL * NL];
            unsigned short lst[256];
        } u;
    };
    struct LdsNoReuse {
        unsigned pad[4];
    };
    __shared__ std::conditional_t<DECIDE, LdsNoReuse, LdsReuse> LR;
    unsigned short *LST = nullptr;
    if constexpr (DECIDE) LST = U.u.fl.lst;
    else LST = reinterpret_cast<LdsReuse &>(LR).u.lst;
    __shared__ int WCNT[4];
    __shared__ int WSTAT[4][3];

    const int witem = xcd_tile(blockIdx.x, gridDim.x);
    const int4 t = tiles[witem];
    const nuc3d_block_dev &B = blocks[t.x];
    const PhysParams P = phys_of(K, B, bin);
    const int a0 = t.y, l0 = t.z, cc = t.w;
    const int nc = min(NC, B.n[DIR] - a0);
    const int nlines = min(NL, (DIR == 0 ? B.n[1] : B.n[0]) - l0);
    double *const *qin = B.q[bin];

    auto sidx = [](int line, int slot) { return DIR == 0 ? line * NSL + slot : slot * NL + line; };
    auto hidx = [](int line, int f) { return DIR == 0 ? line * NF + f : f * NL + line; };
    auto cell = [&](int line, int a, int &i, int &j, int &k) {
        if (DIR == 0) { i = a; j = l0 + line; k = cc; }
        else if (DIR == 1) { i = l0 + line; j = a; k = cc; }
        else { i = l0 + line; j = cc; k = a; }
    };

    const int nslots = nc + 2 * H;
    const int line = DIR == 0 ? threadIdx.x / NF : threadIdx.x % NL;
    const int f = DIR == 0 ? threadIdx.x % NF : threadIdx.x / NL;
    const int n = B.n[DIR], w = K.width[DIR];
    const int m = a0 + f;
    const bool active = line < nlines && f <= nc;
    const bool closure = (B.ftype[2 * DIR] == -1 && m < w) || (B.ftype[2 * DIR + 1] == -1 && m > n - w);
    const bool decided = active && !closure;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    // the face's index in the decision array of direction DIR (a single RHS
    // evaluation keeps its decisions in the second half)
    long long fidx = 0;
    {
        int i, j, k;
        cell(line, m, i, j, k);
        if (DIR == 0) fidx = m + (long long)(B.n[0] + 1) * (j + (long long)B.n[1] * k);
        else if (DIR == 1) fidx = m + (long long)(B.n[1] + 1) * (k + (long long)B.n[2] * i);
        else fidx = m + (long long)(B.n[2] + 1) * (i + (long long)B.n[0] * j);
        if (K.hrhs) fidx += (long long)(B.n[DIR] + 1) * B.n[(DIR + 1) % 3] * B.n[(DIR + 2) % 3];
    }
    // the stored decision of the face, loaded before the staging so that the
    // load's latency overlaps it
    unsigned char fl0 = 0;
    if constexpr (!DECIDE) fl0 = decided ? gldb(B.hflag[DIR] + fidx) : (unsigned char)0;
    // whether every staged split flux is in the range of the reduced division
    // (the WENO5-JS faces of the workgroup then use it; same quotients)
    bool rd_ok = true;
    auto stage_slot = [&](int line, int sl, double *qout) {
        int i, j, k;
        cell(line, a0 - H + sl, i, j, k);
        double fp[5], fn[5], pr[5];
        split_any(B, qin, i, j, k, DIR, P, fp, fn, pr, qout);
        const int o = sidx(line, sl);
#pragma unroll
        for (int c = 0; c < 5; c++) {
            S[c][o] = fp[c];
            S[5 + c][o] = fn[c];
            if constexpr (DECIDE) PR[c][o] = (float)pr[c];
            rd_ok = rd_ok && rd_staged<SCH_HYBRID>(fp[c]) && rd_staged<SCH_HYBRID>(fn[c]);
        }
        if constexpr (!DECIDE) {
            reinterpret_cast<LdsReuse &>(LR).u.prp[0][o] = (float)pr[0];
            reinterpret_cast<LdsReuse &>(LR).u.prp[1][o] = (float)pr[4];
        }
    };
    double u5[5] = {0.0, 0.0, 0.0, 0.0, 0.0};  // viscous z sweep: the state of the thread's own cell
    if constexpr (VISC) {
        // the thread that finishes cell f-1 of a line stages it (and keeps its
        // state for the RK stage); then the 2H halo slots of every line
        const int line = threadIdx.x % NL, f = threadIdx.x / NL;
        if (f >= 1 && f <= nc && line < nlines) stage_slot(line, H + f - 1, u5);
        for (int e = threadIdx.x; e < 2 * H * NL; e += 256) {
            const int ln = e % NL, r = e / NL, sl = r < H ? r : nc + r;
            if (ln < nlines) stage_slot(ln, sl, nullptr);
        }
    } else {
        for (int e = threadIdx.x; e < NSL * NL; e += 256) {
            const int line = DIR == 0 ? e / NSL : e % NL;
            const int sl = DIR == 0 ? e % NSL : e / NL;
            if (line < nlines && sl < nslots) stage_slot(line, sl, nullptr);
        }
    }
    // the workgroup-wide AND of rd_ok through WCNT (free until the decisions
    // are counted, after further barriers): no LDS of its own, which keeps
    // the reusing sweeps at five workgroups per CU
    {
        const unsigned long long all = __ballot(rd_ok);
        if ((threadIdx.x & 63) == 0) WCNT[threadIdx.x >> 6] = all == ~0ull ? 1 : 0;
    }
    __syncthreads();
    const bool rd = WCNT[0] && WCNT[1] && WCNT[2] && WCNT[3] && !K.ieee_div;

    bool posok = false, weno = false;
    if constexpr (DECIDE) {
        // every face: its 30 features and the positivity fallback (the stencil's
        // mean rho and p and the central face values of rho and p over cells
        // m-4..m+3 = slots f..f+7 must be positive)
        unsigned xw[16];
        if (decided) {
            auto getf = [&](int v, int j) { return PR[v][sidx(line, f + 1 + j)]; };
            const bool valid = features(getf, (float)K.gamma, xw);
            auto c8 = [&](int v) {
                return central8f(PR[v][sidx(line, f)], PR[v][sidx(line, f + 1)], PR[v][sidx(line, f + 2)],
                                 PR[v][sidx(line, f + 3)], PR[v][sidx(line, f + 4)], PR[v][sidx(line, f + 5)],
                                 PR[v][sidx(line, f + 6)], PR[v][sidx(line, f + 7)]);
            };
            posok = valid && c8(0) > 0.0f && c8(4) > 0.0f;
        } else {
#pragma unroll
            for (int i = 0; i < 16; i++) xw[i] = 0u;
        }
        __syncthreads();  // the primitives are dead: the area takes the features
        {
            // features 30 and 31 (word 15) are zero and not stored
            unsigned *row = XF + 15 * threadIdx.x;
#pragma unroll
            for (int i = 0; i < 15; i++) row[i] = xw[i];
        }
        __syncthreads();
        // the network: wavefront wv evaluates its 64 faces in four MFMA tiles of
        // 16; lane group g takes features 8g..8g+7 (words 4g..4g+3 of a row);
        // the logit of face tf then takes the first word of its row
        {
            Regs R;
            load_regs_lanes(K.sensor, R);
            const int g = lane >> 4;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int tf = 64 * wv + 16 * q + (lane & 15);
                const unsigned *row = XF + 15 * tf + 4 * g;
                const unsigned w3 = row[g < 3 ? 3 : 2];  // lane group 3: word 15 is zero (not stored)
                const uint4 v4 = make_uint4(row[0], row[1], row[2], g < 3 ? w3 : 0u);
                Acts A;
                const float z = forward16(R, __builtin_bit_cast(bf16x8, v4), A);
                if (g == 0) XF[15 * tf] = __float_as_uint(z);
            }
        }
        __syncthreads();  // the features are dead: the area takes the face fluxes
        weno = decided && (!posok || __uint_as_float(XF[15 * threadIdx.x]) >= K.sensor->thr);
    } else {
        // stages 1 and 2 of a step (and every stage of a step that reuses the
        // previous step's decisions) take the stored network decisions, but
        // the positivity fallback is evaluated again on this stage's state: a
        // central face whose stencil means or central face values of rho or p
        // are no longer positive takes WENO5-JS and is recorded as a fallback
        unsigned char fl = fl0;
        if (decided && fl == 0) {  // only central faces can fall back
            const float(*PP)[NSL * NL] = reinterpret_cast<LdsReuse &>(LR).u.prp;
            auto getf = [&](int v, int j) { return PP[v == 0 ? 0 : 1][sidx(line, f + 1 + j)]; };
            auto c8 = [&](int v) {
                return central8f(PP[v][sidx(line, f)], PP[v][sidx(line, f + 1)], PP[v][sidx(line, f + 2)],
                                 PP[v][sidx(line, f + 3)], PP[v][sidx(line, f + 4)], PP[v][sidx(line, f + 5)],
                                 PP[v][sidx(line, f + 6)], PP[v][sidx(line, f + 7)]);
            };
            const bool pos = stencil_mean(getf, 0) > 0.0f && stencil_mean(getf, 4) > 0.0f && c8(0) > 0.0f &&
                             c8(1) > 0.0f;
            if (fl == 0 && !pos) {
                fl = 2;
                gstb(B.hflag[DIR] + fidx, 2);
            }
        }
        weno = decided && (fl == 1 || fl == 2);
        posok = fl != 2;
        __syncthreads();  // rho and p are dead: the area takes the list of WENO5-JS faces
    }
    const unsigned long long bal = __ballot(weno);
    // a tile owns faces 1..nc (and face 0 of the first tile): the face shared
    // with the previous tile is counted there
    const bool owned = f >= 1 || a0 == 0;
    const unsigned long long bcen = __ballot(owned && decided && !weno);
    const unsigned long long bnet = __ballot(owned && weno && posok);
    const unsigned long long bfb = __ballot(owned && decided && !posok);
    if (lane == 0) {
        WCNT[wv] = __popcll(bal);
        WSTAT[wv][0] = __popcll(bcen);
        WSTAT[wv][1] = __popcll(bnet);
        WSTAT[wv][2] = __popcll(bfb);
    }
    if (DECIDE && active) gstb(B.hflag[DIR] + fidx, closure ? 3 : (weno ? (posok ? 1 : 2) : 0));
    __syncthreads();
    int base = 0, total = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        base += q < wv ? WCNT[q] : 0;
        total += WCNT[q];
    }
    if (DECIDE && threadIdx.x < 3 && K.hcount) {
        const int v = WSTAT[0][threadIdx.x] + WSTAT[1][threadIdx.x] + WSTAT[2][threadIdx.x] + WSTAT[3][threadIdx.x];
        if (v) atomicAdd(K.hcount + 16 * (witem % NUC3D_HCOUNT_SLOTS) + threadIdx.x, (unsigned long long)v);
    }
    if (weno) LST[base + __popcll(bal & ((1ull << lane) - 1ull))] = (unsigned short)threadIdx.x;
    // slots f..f+7 = cells m-4..m+3; the six-cell stencil starts at m-3.
    // Closure faces (next to physical faces only) and central faces are
    // separate branches, so waves without closure faces skip that code.
    // Where the face fluxes take the split-flux area, every value is kept in
    // registers until no thread reads the split fluxes any more.
    double hv[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    const bool own_face = active && (closure || !weno);
    if (active && closure) {
#pragma unroll
        for (int c = 0; c < 5; c++) {
            double sp[NST], sn[NST];
#pragma unroll
            for (int q = 0; q < NST; q++) {
                sp[q] = S[c][sidx(line, f + q)];
                sn[q] = S[5 + c][sidx(line, f + q)];
            }
            hv[c] = rd ? face_value<SCH_WENO5JS, true>(B, K, DIR, m, sp + 1, sn + 1)
                       : face_value<SCH_WENO5JS, false>(B, K, DIR, m, sp + 1, sn + 1);
        }
    } else if (decided && !weno) {
#pragma unroll
        for (int c = 0; c < 5; c++) {
            double sp[NST], sn[NST];
#pragma unroll
            for (int q = 0; q < NST; q++) {
                sp[q] = S[c][sidx(line, f + q)];
                sn[q] = S[5 + c][sidx(line, f + q)];
            }
            hv[c] = central8(sp + 1) + central8(sn + 1);
        }
    }
    if constexpr (!HF_IN_S) {
        if (own_face)
#pragma unroll
            for (int c = 0; c < 5; c++) HF[c][hidx(line, f)] = hv[c];
    }
    __syncthreads();
    // items component-major, so that neighbouring lanes read one component's
    // LDS array at neighbouring faces (no bank conflicts between components);
    // at most five per thread (5 x 256 faces)
    double wv5[5];
    int wo5[5];
#pragma unroll
    for (int r = 0; r < 5; r++) {
        const int e = threadIdx.x + 256 * r;
        wo5[r] = -1;
        wv5[r] = 0.0;
        if (e < 5 * total) {
            const int c = e / total, tf = LST[e - c * total];
            const int ln = DIR == 0 ? tf / NF : tf % NL, ff = DIR == 0 ? tf % NF : tf / NL;
            double sp[6], sn[6];
#pragma unroll
            for (int q = 0; q < 6; q++) {
                sp[q] = S[c][sidx(ln, ff + 1 + q)];
                sn[q] = S[5 + c][sidx(ln, ff + 1 + q)];
            }
            wv5[r] = rd ? face_interior<SCH_WENO5JS, true>(sp, true, false) +
                              face_interior<SCH_WENO5JS, true>(sn, false, false)
                        : face_interior<SCH_WENO5JS, false>(sp, true, false) +
                              face_interior<SCH_WENO5JS, false>(sn, false, false);
            wo5[r] = c * (NF * NL) + hidx(ln, ff);
            if constexpr (!HF_IN_S) (&HF[0][0])[wo5[r]] = wv5[r];
        }
    }
    if constexpr (HF_IN_S) {
        __syncthreads();  // no thread reads the split fluxes any more
        if (own_face)
#pragma unroll
            for (int c = 0; c < 5; c++) HF[c][hidx(line, f)] = hv[c];
#pragma unroll
        for (int r = 0; r < 5; r++)
            if (wo5[r] >= 0) (&HF[0][0])[wo5[r]] = wv5[r];
    }
    __syncthreads();

    double lmax[3] = {0.0, 0.0, 0.0};
    StageCounts nfix;
    double r2 = 0.0;
    const double dt = (DIR == 2 && MODE == 1) ? stage_dt(blocks, K, stage, B.group, dtbuf, lam_cur, lam_next) : 0.0;
    if (line < nlines && f >= 1 && f <= nc) {
        int i, j, k;
        cell(line, a0 + f - 1, i, j, k);
        double dF[5];
#pragma unroll
        for (int c = 0; c < 5; c++) dF[c] = HF[c][hidx(line, f)] - HF[c][hidx(line, f - 1)];
        if constexpr (VISC)
            finish_cell_visc_z<WIDE>(blocks, B, K, P, u5, i, j, k, dF, stage, bun, bout, dt, lmax, nfix, r2);
        else
            finish_cell<DIR, NS, MODE>(blocks, B, K, P, i, j, k, dF, stage, bin, bun, bout, dt, lmax, nfix, r2);
    }
    if (DIR == 2 && MODE == 1) {
        nfix.flush(counters);
        if (stage == 2) block_max3_atomic(lmax, lam_next + 3 * t.x);
        if (stage == 2 && K.resid) block_sum_store(r2, K.resid + witem);
    }
}

// ---------------------------------------------------------------------------
// CRWENO5 along direction DIR (crweno5.cpp:44-401): NUC3D_CRW_LINES lines of
// one block per wavefront, one lane per tridiagonal system.  Faces w..n-w of
// a line form a system per component and upwind side whose end rows are
// explicit WENO5-JS and whose inner rows carry the compact weights
// (crweno5.cpp:130-160, :310-340); the ten systems of each of the six lines
// are solved by 60 lanes side by side with the Thomas recurrence of
// crweno5.cpp:207-218, in the reference's operation order, the row
// coefficients computed on the fly from split fluxes staged in LDS in chunks
// of CRW_R rows.  c' and d' go to global scratch (one 64-wide row per system
// row: coalesced), the back substitution walks the chunks down again and
// stages the face solutions in LDS, where the cells of the chunk are
// finished.  Faces outside the system use the WENO5-JS face variants or the
// physical closures.  Line list: groups of NUC3D_CRW_LINES lines of one
// block (x = block, y, z = transverse coordinates j,k / i,k / i,j), padded
// with x = -1.
#ifndef NUC3D_CRW_R
#define NUC3D_CRW_R 32
#endif
#ifndef NUC3D_CRW_BATCH
#define NUC3D_CRW_BATCH 4
#endif
constexpr int CRW_R = NUC3D_CRW_R;       // rows per staged chunk
constexpr int CRW_B = NUC3D_CRW_BATCH;   // rows whose coefficients a lane computes together
constexpr int CRW_HQ = 9;                // staging slot of the high end cells (after the w + 5 <= 9 low ones)

template <int DIR, bool NS, int MODE>
__global__ __launch_bounds__(64) void KNAME(k_crweno)(const nuc3d_block_dev *__restrict__ blocks,
                                                      const int4 *__restrict__ lines, nuc3d_kparams K, int stage,
                                                      int bin, int bun, int bout, double *dtbuf,
                                                      const unsigned long long *lam_cur,
                                                      unsigned long long *lam_next, unsigned long long *counters,
                                                      double *scratch, int nmax)
{
    constexpr int LW = NUC3D_CRW_LINES, R = CRW_R, NQ = R + 5;
    static_assert(LW * 10 <= 64, "one lane per system");
    static_assert(NQ >= CRW_HQ + 9, "end cells of both line ends fit one chunk");
    __shared__ union {
        double sf[10][LW][NQ + 1];  // split fluxes F+ (0..4), F- (5..9) of the staged cells
        double xl[64][R + 1];       // face solutions of one chunk, per system lane
    } U;
    __shared__ double hfe[LW][5][8];  // explicit faces: 0..w-1 at 0..w-1, n-w+1..n at 4..3+w
    __shared__ int4 lsh[LW];
    const int witem = xcd_tile(blockIdx.x, gridDim.x);
    const int4 *lg = lines + (long long)witem * LW;
    if (threadIdx.x < LW) lsh[threadIdx.x] = lg[threadIdx.x];
    const nuc3d_block_dev &B = blocks[lg[0].x];
    const PhysParams P = phys_of(K, B, bin);
    const int n = B.n[DIR], w = K.width[DIR], M = n - 2 * w + 1;
    double *const *qin = B.q[bin];
    __syncthreads();
    auto cell = [&](int l, int a, int &i, int &j, int &k) {
        const int4 Ll = lsh[l];
        if (DIR == 0) { i = a; j = Ll.y; k = Ll.z; }
        else if (DIR == 1) { i = Ll.y; j = a; k = Ll.z; }
        else { i = Ll.y; j = Ll.z; k = a; }
    };
    // split fluxes of cells cb..cb+nc-1 of every line into U.sf[.][l][q0 + q]
    // (consecutive lanes take consecutive cells along x, consecutive lines
    // across: the lines of a group are neighbours in i for DIR 1 and 2)
    auto stage_cells = [&](int cb, int nc, int q0) {
        for (int e = threadIdx.x; e < LW * nc; e += 64) {
            const int l = DIR == 0 ? e / nc : e % LW, q = DIR == 0 ? e % nc : e / LW;
            if (lsh[l].x < 0) continue;
            int i, j, k;
            cell(l, cb + q, i, j, k);
            double fp[5], fn[5];
            split_any(B, qin, i, j, k, DIR, P, fp, fn);
#pragma unroll
            for (int c = 0; c < 5; c++) {
                U.sf[c][l][q0 + q] = fp[c];
                U.sf[5 + c][l][q0 + q] = fn[c];
            }
        }
    };

    // explicit faces 0..w-1 (cells -3..w+1 at q 0..w+4) and n-w+1..n
    // (cells n-w-2..n+2 at q CRW_HQ..CRW_HQ+w+4): halo face variants or physical closures
    stage_cells(-3, w + 5, 0);
    stage_cells(n - w - 2, w + 5, CRW_HQ);
    __syncthreads();
    for (int e = threadIdx.x; e < LW * 5 * 2 * w; e += 64) {
        const int l = e / (10 * w), c = (e / (2 * w)) % 5, t = e % (2 * w);
        if (lsh[l].x < 0) continue;
        const int m = t < w ? t : n - w + 1 + (t - w);
        const int q = t < w ? m : CRW_HQ + (t - w);  // q of cell m-3
        double sp[6], sn[6];
#pragma unroll
        for (int o = 0; o < 6; o++) {
            sp[o] = U.sf[c][l][q + o];
            sn[o] = U.sf[5 + c][l][q + o];
        }
        hfe[l][c][t < w ? t : 4 + (t - w)] = face_value<SCH_WENO5JS>(B, K, DIR, m, sp, sn);
    }

    // forward elimination, CRW_R rows per chunk; row r = face w+r, reference
    // loop index i = w-1+r, stencil cells i-2..i+3 (staged at q = rr..rr+5)
    const int s = threadIdx.x, ls = s / 10, sys = s % 10, c = sys % 5;
    const bool plus = sys < 5;
    const bool active = s < LW * 10 && lsh[ls < LW ? ls : 0].x >= 0;
    double *Cg = scratch + (long long)witem * 128 * nmax;  // c'[r][64]
    double *Dg = Cg + 64 * (long long)nmax;                // d'[r][64]
    double cprev = 0.0, dprev = 0.0;
    for (int r0 = 0; r0 < M; r0 += R) {
        const int nr = min(R, M - r0);
        __syncthreads();
        stage_cells(w - 3 + r0, nr + 5, 0);
        __syncthreads();
        if (active) {
            const double *f = &U.sf[plus ? c : 5 + c][ls][0];
            int rr = 0;
            if (r0 == 0) {  // row 0: explicit face w
                const double d = plus ? weno5js(f[0], f[1], f[2], f[3], f[4])
                                      : weno5js(f[5], f[4], f[3], f[2], f[1]);
                Cg[s] = 0.0;
                Dg[s] = d;
                cprev = 0.0;
                dprev = d;
                rr = 1;
            }
            const bool last = r0 + nr == M && M > 1;
            const int rin = last ? nr - 1 : nr;
            // the row coefficients do not depend on the recurrence: four rows
            // at a time give the lane four independent chains to interleave
            for (; rr + CRW_B <= rin; rr += CRW_B) {
                double a[CRW_B], b[CRW_B], cc[CRW_B], d[CRW_B];
#pragma unroll
                for (int u = 0; u < CRW_B; u++) crweno_row(f + rr + u + 2, 1, plus, a[u], b[u], cc[u], d[u]);
#pragma unroll
                for (int u = 0; u < CRW_B; u++) {
                    cc[u] = cc[u] / (b[u] - cprev * a[u]);
                    d[u] = (d[u] - dprev * a[u]) / (b[u] - cprev * a[u]);
                    Cg[(long long)(r0 + rr + u) * 64 + s] = cc[u];
                    Dg[(long long)(r0 + rr + u) * 64 + s] = d[u];
                    cprev = cc[u];
                    dprev = d[u];
                }
            }
            for (; rr < rin; rr++) {
                double a, b, cc, d;
                crweno_row(f + rr + 2, 1, plus, a, b, cc, d);
                cc = cc / (b - cprev * a);
                d = (d - dprev * a) / (b - cprev * a);
                Cg[(long long)(r0 + rr) * 64 + s] = cc;
                Dg[(long long)(r0 + rr) * 64 + s] = d;
                cprev = cc;
                dprev = d;
            }
            if (last) {  // row M-1: explicit face n-w
                const double a = 0.0, b = 1.0;
                double cc = 0.0;
                double d = plus ? weno5js(f[rr], f[rr + 1], f[rr + 2], f[rr + 3], f[rr + 4])
                                : weno5js(f[rr + 5], f[rr + 4], f[rr + 3], f[rr + 2], f[rr + 1]);
                cc = cc / (b - cprev * a);
                d = (d - dprev * a) / (b - cprev * a);
                Cg[(long long)(M - 1) * 64 + s] = cc;
                Dg[(long long)(M - 1) * 64 + s] = d;
            }
        }
    }

    double lmax[3] = {0.0, 0.0, 0.0};
    StageCounts nfix;
    double r2 = 0.0;  // residual partial (stage 2)
    const double dt = (DIR == 2 && MODE == 1) ? stage_dt(blocks, K, stage, B.group, dtbuf, lam_cur, lam_next) : 0.0;
    // back substitution (out[n-w] = d'[M-1]; out[w+r] = d'[r] - c'[r] out[w+r+1])
    // by chunks of cells c0..c1-1 from the top; faces c0..c1 at slots c0-c1+R..R
    double x = 0.0;
    for (int c1 = n; c1 > 0; c1 -= R) {
        const int c0 = max(0, c1 - R);
        __syncthreads();
        if (active) {
            if (c1 < n && c1 >= w && c1 <= n - w) U.xl[s][R] = x;  // face c1: the row of the chunk above
            int fhi = min(c1 - 1, n - w);
            const int flo = max(c0, w);
            if (fhi == n - w && fhi >= flo) {
                x = Dg[(long long)(M - 1) * 64 + s];
                U.xl[s][fhi - c1 + R] = x;
                fhi--;
            }
#pragma unroll 4
            for (int f = fhi; f >= flo; f--) {
                const long long o = (long long)(f - w) * 64 + s;
                x = Dg[o] - Cg[o] * x;
                U.xl[s][f - c1 + R] = x;
            }
        }
        __syncthreads();
        auto hf = [&](int l, int cc, int f) -> double {
            if (f < w) return hfe[l][cc][f];
            if (f > n - w) return hfe[l][cc][4 + f - (n - w + 1)];
            const int sl = f - c1 + R;
            return U.xl[l * 10 + cc][sl] + U.xl[l * 10 + 5 + cc][sl];
        };
        const int nc = c1 - c0;
        for (int e = threadIdx.x; e < LW * nc; e += 64) {
            const int l = DIR == 0 ? e / nc : e % LW, a = c0 + (DIR == 0 ? e % nc : e / LW);
            if (lsh[l].x < 0) continue;
            int i, j, k;
            cell(l, a, i, j, k);
            double dF[5];
#pragma unroll
            for (int cc = 0; cc < 5; cc++) dF[cc] = hf(l, cc, a + 1) - hf(l, cc, a);
            finish_cell<DIR, NS, MODE>(blocks, B, K, P, i, j, k, dF, stage, bin, bun, bout, dt, lmax, nfix, r2);
        }
    }
    if (DIR == 2 && MODE == 1) {
        nfix.flush(counters);
        if (stage == 2) {
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const double v = warp_max(lmax[d]);
                if (threadIdx.x == 0 && v > 0.0)
                    atomicMax(lam_next + 3 * lg[0].x + d, (unsigned long long)__double_as_longlong(v));
            }
            if (!NS && K.resid) block_sum_store(r2, K.resid + witem);
        }
    }
}

// ---------------------------------------------------------------------------
// Navier-Stokes viscous terms (NaiverStokes3d.cpp:51-513) in two passes
// over 16x16 column tiles (NUC3D_TILE_X x NUC3D_TILE_Y) marching in k.
//
// k_visc_flux: u, v, w, T from Q/J and J as each plane arrives (con2prim,
//              physicsModel.cpp:313-353, T = gamma Ma^2 p / rho,
//              physicsModel.cpp:1029; ghost Q and J are the neighbours' own
//              values, so the halo's u, v, w, T are theirs bit for bit);
//              cd6 (or exterior closure) derivatives of u, v, w, T along xi
//              and eta from a double-buffered LDS plane with its xy halo, and
//              along zeta from a seven-plane register ring of the thread's own
//              column (NaiverStokes3d.cpp:86-168); chain rule with the
//              metrics, Sutherland viscosity, stresses, heat flux and the
//              contravariant fluxes over J (NaiverStokes3d.cpp:170-326).
// k_visc_div:  after the three inviscid sweeps, rhs -= dFv/dxi, dGv/deta,
//              dHv/dzeta in the reference's order (NaiverStokes3d.cpp:401-513),
//              fv_xi / fv_eta from LDS planes and fv_zeta from a register ring,
//              then the RHS (MODE 0) or the TVD-RK3 stage (MODE 1).
//
// The loads of plane k+1 (ring, halo, centre values, and J + metrics or the
// RHS) are issued before plane k is computed, so memory latency overlaps the
// arithmetic; one barrier per plane separates the LDS fill from its use.

// Waves per SIMD the viscous kernels are register-limited to.  Measured at
// 256^3: both run fastest unconstrained (187 VGPRs for the viscous fluxes on
// the uniform-mesh path, 227 with the metric fields, 235 for the divergence
// pass: 2 waves per SIMD, every load of the next plane in flight); capping
// them at 3 or 4 waves spills and is slower.
#ifndef NUC3D_VISC_FLUX_WAVES
#define NUC3D_VISC_FLUX_WAVES 2
#endif
#ifndef NUC3D_VISC_FLUX_WAVES_UNI
#define NUC3D_VISC_FLUX_WAVES_UNI 2
#endif
#ifndef NUC3D_VISC_DIV_WAVES
#define NUC3D_VISC_DIV_WAVES 2
#endif

// u, v, w, T of a cell from its conserved state Q/J (q6[0..4]) and J (q6[5])
__device__ __forceinline__ void prim4_of(const double *q6, const PhysParams &P, double *w4)
{
    double rho, u, v, w, p, cs, y = 0.0;
    con2prim(q6, q6[5], P, rho, u, v, w, p, cs, &y);
    w4[0] = u;
    w4[1] = v;
    w4[2] = w;
    // physicsModel.cpp:1029
    w4[3] = P.rdq ? div_nrs(P.gamma * P.mach * P.mach * p, rho, y) : P.gamma * P.mach * P.mach * p / rho;
}

__device__ __forceinline__ void load_q6(const nuc3d_block_dev &B, double *const *qb, long long x, double *q6)
{
#pragma unroll
    for (int c = 0; c < 5; c++) q6[c] = gld(qb[c] + x);
    q6[5] = jac_at(B, x);
}

__device__ __forceinline__ void prim4(const nuc3d_block_dev &B, double *const *qb, long long x, const PhysParams &P,
                                      double *w4)
{
    double q6[6];
    load_q6(B, qb, x, q6);
    prim4_of(q6, P, w4);
}

// k_prim: u, v, w, T of every cell into B.prim, pushed to the neighbours'
// ghost layers like Q (for the diagnostics; the stage pipeline computes them
// inside k_visc_flux).  One workgroup per (tile, plane): blockIdx.y = plane
// within the tile's k segment.
__global__ __launch_bounds__(256) void KNAME(k_prim)(const nuc3d_block_dev *__restrict__ blocks,
                                                     const nuc3d_tile *__restrict__ tiles, nuc3d_kparams K, int bin)
{
    constexpr int T = NUC3D_TILE_X;
    const nuc3d_tile tl = tiles[xcd_tile(blockIdx.x, gridDim.x)];
    if ((int)blockIdx.y >= tl.kn) return;
    const nuc3d_block_dev &B = blocks[tl.b];
    const PhysParams P = phys_of(K, B);
    const int i = tl.i0 + (threadIdx.x & (T - 1)), j = tl.j0 + threadIdx.x / T, k = tl.k0 + blockIdx.y;
    if (i >= B.n[0] || j >= B.n[1]) return;
    const long long x = cidx(B, i, j, k);
    double w4[4];
    prim4(B, B.q[bin], x, P, w4);
#pragma unroll
    for (int v = 0; v < 4; v++) B.prim[v][x] = w4[v];
#pragma unroll
    for (int f = 0; f < 6; f++) {
        int nb;
        long long y;
        if (ghost_target(blocks, B, f, i, j, k, nb, y)) {
#pragma unroll
            for (int v = 0; v < 4; v++) blocks[nb].prim[v][y] = w4[v];
        }
    }
}

// Halo cell h of a TX x TY column tile (0 <= h < 2H (TX + TY), H = 3 for
// cd6, 4 for cd8): h < 2H TY are the strips of H cells left and right of each
// row (x halo), then the rows of H cells below and above the tile (y halo,
// consecutive h along i); no corners (the derivatives are not mixed).
// Tile-local (li, lj); false for rows or columns beyond a partial tile.
template <int H>
__device__ __forceinline__ bool halo_cell(int h, int tw, int th, int &li, int &lj)
{
    constexpr int TX = NUC3D_TILE_X, TY = NUC3D_TILE_Y;
    static_assert(2 * H * (TX + TY) <= 256, "one halo cell per thread");
    if (h < TY * 2 * H) {
        const int r = h / (2 * H), s = h % (2 * H);
        if (r >= th) return false;
        lj = r;
        li = s < H ? s - H : tw + (s - H);
    } else {
        const int c = (h - TY * 2 * H) % TX, s = (h - TY * 2 * H) / TX;
        if (c >= tw) return false;
        li = c;
        lj = s < H ? s - H : th + (s - H);
    }
    return true;
}

// Viscous fluxes of one cell from its u,v,w,T gradients along xi, eta, zeta
// (dxi/deta/dzeta [4] each) and centre values (NaiverStokes3d.cpp:170-326).
__device__ __forceinline__ void visc_flux_cell(const nuc3d_kparams &K, const PhysParams &P, const double *dxi,
                                               const double *deta, const double *dzeta, const double *w4,
                                               const double *m, double jac, double *out)
{
    const double xi_x = m[0], xi_y = m[1], xi_z = m[2];
    const double eta_x = m[3], eta_y = m[4], eta_z = m[5];
    const double zeta_x = m[6], zeta_y = m[7], zeta_z = m[8];
    double g[4][3];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const double a = dxi[v], bb = deta[v], cc = dzeta[v];
        g[v][0] = a * xi_x + bb * eta_x + cc * zeta_x;
        g[v][1] = a * xi_y + bb * eta_y + cc * zeta_y;
        g[v][2] = a * xi_z + bb * eta_z + cc * zeta_z;
    }
    const double S = K.sref;
    const double u = w4[0], vv = w4[1], w = w4[2], T0 = w4[3];
    const double miu0 = (1.0 + S) / (T0 + S) * pow(T0, 1.5) / P.reynolds;  // physicsModel.cpp:226
    const double coeff0 = miu0 / ((K.gamma - 1.0) * P.mach * P.mach * K.prandtl);
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
    const double fv[4] = {txx, txy, txz, u * txx + vv * txy + w * txz + ttx};
    const double gv[4] = {txy, tyy, tyz, u * txy + vv * tyy + w * tyz + tty};
    const double hv[4] = {txz, tyz, tzz, u * txz + vv * tyz + w * tzz + ttz};
#pragma unroll
    for (int c = 0; c < 4; c++) {
        out[c] = (xi_x * fv[c] + xi_y * gv[c] + xi_z * hv[c]) / jac;
        out[4 + c] = (eta_x * fv[c] + eta_y * gv[c] + eta_z * hv[c]) / jac;
        out[8 + c] = (zeta_x * fv[c] + zeta_y * gv[c] + zeta_z * hv[c]) / jac;
    }
}

// HD: the derivatives' half-width (3 for cd6, 4 where a direction uses cd8)
// UNI: every block is on the uniform-mesh path, so J and the metric terms
// are the descriptor's constants (no per-cell loads or prefetch registers)
template <int HD, bool UNI>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(UNI ? NUC3D_VISC_FLUX_WAVES_UNI : NUC3D_VISC_FLUX_WAVES))) void KNAME(k_visc_flux)(
    const nuc3d_block_dev *__restrict__ blocks, const nuc3d_tile *__restrict__ tiles, nuc3d_kparams K, int bin)
{
    constexpr int TX = NUC3D_TILE_X, TY = NUC3D_TILE_Y, H = HD, LX = TX + 2 * H, LY = TY + 2 * H;
    __shared__ double Wp[2][4][LY][LX];  // u,v,w,T of plane k with the xy halo, double-buffered
    const nuc3d_tile tl = tiles[xcd_tile(blockIdx.x, gridDim.x)];
    const nuc3d_block_dev &B = blocks[tl.b];
    const PhysParams P = phys_of(K, B, bin);
    double *const *qb = B.q[bin];
    const int ti = threadIdx.x & (TX - 1), tj = threadIdx.x / TX;
    const int tw = min(TX, B.n[0] - tl.i0), th = min(TY, B.n[1] - tl.j0);
    const bool incell = ti < tw && tj < th;
    const int i = tl.i0 + ti, j = tl.j0 + tj;
    int hi = 0, hj = 0;
    const bool hv = threadIdx.x < 2 * H * (TX + TY) && halo_cell<H>(threadIdx.x, tw, th, hi, hj);
    const long long xc = cidx(B, i, j, 0), xh = cidx(B, tl.i0 + hi, tl.j0 + hj, 0), sz = B.sz;
    const int kb = tl.k0, ke = tl.k0 + tl.kn;

    double zr[2 * H + 1][4];  // u,v,w,T of the own column, planes p-6..p
#pragma unroll
    for (int a = 0; a < 2 * H + 1; a++)
#pragma unroll
        for (int v = 0; v < 4; v++) zr[a][v] = 0.0;
    double pc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, ph[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // Q/J and J of
                                                                                        // plane p / halo plane k
    double pm[UNI ? 1 : 10];                                             // prefetched J + metrics of cell k
#pragma unroll
    for (int e = 0; e < (UNI ? 1 : 10); e++) pm[e] = 0.0;
    // Q/J of cell x and J: J is the block's constant on the uniform-mesh path
    auto load_state = [&](long long x, double *q6) {
        if constexpr (UNI) {
#pragma unroll
            for (int c = 0; c < 5; c++) q6[c] = gld(qb[c] + x);
            q6[5] = B.jac0;
        } else {
            load_q6(B, qb, x, q6);
        }
    };
    if (incell) load_state(xc + (kb - H) * sz, pc);

    for (int p = kb - H; p < ke + H; p++) {
        const int k = p - H, buf = p & 1;
#pragma unroll
        for (int a = 0; a < 2 * H; a++)
#pragma unroll
            for (int v = 0; v < 4; v++) zr[a][v] = zr[a + 1][v];
        prim4_of(pc, P, zr[2 * H]);
        double m[10];
        if constexpr (UNI) {
            m[0] = B.jac0;
#pragma unroll
            for (int e = 0; e < 9; e++) m[1 + e] = B.met0[e];
        } else {
#pragma unroll
            for (int e = 0; e < 10; e++) m[e] = pm[e];
        }
        if (k >= kb) {
            double wh[4];
            prim4_of(ph, P, wh);
#pragma unroll
            for (int v = 0; v < 4; v++) {
                if (incell) Wp[buf][v][tj + H][ti + H] = zr[H][v];
                if (hv) Wp[buf][v][hj + H][hi + H] = wh[v];
            }
        }
        // loads for the next plane, in flight during this plane's work
        if (incell && p + 1 < ke + H) load_state(xc + (p + 1) * sz, pc);
        if (k + 1 >= kb && k + 1 < ke) {
            if (hv) load_state(xh + (k + 1) * sz, ph);
            if constexpr (!UNI) {
                if (incell) {
                    pm[0] = jac_at(B, xc + (k + 1) * sz);
#pragma unroll
                    for (int e = 0; e < 9; e++) pm[1 + e] = met_at(B, e, xc + (k + 1) * sz);
                }
            }
        }
        __syncthreads();
        if (k < kb || !incell) continue;
        double dxi[4], deta[4], dzeta[4], w4[4];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            double col[2 * H + 1];
#pragma unroll
            for (int a = 0; a < 2 * H + 1; a++) col[a] = zr[a][v];
            dxi[v] = line_diff_lds<HD == 4>(B, K, &Wp[buf][v][tj + H][ti + H], 1, 0, i);
            deta[v] = line_diff_lds<HD == 4>(B, K, &Wp[buf][v][tj + H][ti + H], LX, 1, j);
            dzeta[v] = line_diff_lds<HD == 4>(B, K, &col[H], 1, 2, k);
            w4[v] = zr[H][v];
        }
        double out[12];
        visc_flux_cell(K, P, dxi, deta, dzeta, w4, m + 1, m[0], out);
        const long long x = xc + k * sz;
#pragma unroll
        for (int e = 0; e < 12; e++) gst(B.fv[e] + x, out[e]);
#pragma unroll
        for (int f = 0; f < 6; f++) {
            int nb;
            long long y;
            if (ghost_target(blocks, B, f, i, j, k, nb, y)) {
#pragma unroll
                for (int e = 0; e < 12; e++) gst(blocks[nb].fv[e] + y, out[e]);
            }
        }
    }
}

template <int MODE, int HD>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NUC3D_VISC_DIV_WAVES))) void KNAME(k_visc_div)(
    const nuc3d_block_dev *__restrict__ blocks, const nuc3d_tile *__restrict__ tiles, nuc3d_kparams K, int stage,
    int bin, int bun, int bout, double *dtbuf, const unsigned long long *lam_cur, unsigned long long *lam_next,
    unsigned long long *counters)
{
    constexpr int TX = NUC3D_TILE_X, TY = NUC3D_TILE_Y, H = HD, LX = TX + 2 * H, LY = TY + 2 * H;
    __shared__ double Vx[2][4][TY][LX];  // fv_xi (components 1..4) of plane k, x halo, double-buffered
    __shared__ double Vy[2][4][LY][TX];  // fv_eta, y halo
    const nuc3d_tile tl = tiles[xcd_tile(blockIdx.x, gridDim.x)];
    const nuc3d_block_dev &B = blocks[tl.b];
    const PhysParams P = phys_of(K, B);
    const int ti = threadIdx.x & (TX - 1), tj = threadIdx.x / TX;
    const int tw = min(TX, B.n[0] - tl.i0), th = min(TY, B.n[1] - tl.j0);
    const bool incell = ti < tw && tj < th;
    const int i = tl.i0 + ti, j = tl.j0 + tj;
    int hi = 0, hj = 0;
    const bool hv = threadIdx.x < 2 * H * (TX + TY) && halo_cell<H>(threadIdx.x, tw, th, hi, hj);
    const bool hx = threadIdx.x < 2 * H * TY;  // x halo (fv_xi) or y halo (fv_eta)
    const long long xc = cidx(B, i, j, 0), xh = cidx(B, tl.i0 + hi, tl.j0 + hj, 0), sz = B.sz;
    const int kb = tl.k0, ke = tl.k0 + tl.kn;
    const double dt = MODE == 1 ? stage_dt(blocks, K, stage, B.group, dtbuf, lam_cur, lam_next) : 0.0;
    double lmax[3] = {0.0, 0.0, 0.0};
    StageCounts nfix;
    double r2 = 0.0;  // residual partial (stage 2)

    double zr[2 * H + 1][4];  // fv_zeta (components 1..4) of the own column, planes p-6..p
#pragma unroll
    for (int a = 0; a < 2 * H + 1; a++)
#pragma unroll
        for (int v = 0; v < 4; v++) zr[a][v] = 0.0;
    double pz[4] = {0.0, 0.0, 0.0, 0.0}, ph[4] = {0.0, 0.0, 0.0, 0.0}, pxy[8];
    double pr[5];  // RHS of cell k (prefetched; the state is loaded when the plane is computed)
#pragma unroll
    for (int e = 0; e < 8; e++) pxy[e] = 0.0;
#pragma unroll
    for (int c = 0; c < 5; c++) pr[c] = 0.0;
    if (incell)
#pragma unroll
        for (int v = 0; v < 4; v++) pz[v] = gld(B.fv[8 + v] + xc + (kb - H) * sz);

    for (int p = kb - H; p < ke + H; p++) {
        const int k = p - H, buf = p & 1;
#pragma unroll
        for (int a = 0; a < 2 * H; a++)
#pragma unroll
            for (int v = 0; v < 4; v++) zr[a][v] = zr[a + 1][v];
#pragma unroll
        for (int v = 0; v < 4; v++) zr[2 * H][v] = pz[v];
        double r[5];
#pragma unroll
        for (int c = 0; c < 5; c++) r[c] = pr[c];
        if (k >= kb) {
#pragma unroll
            for (int v = 0; v < 4; v++) {
                if (incell) {
                    Vx[buf][v][tj][ti + H] = pxy[v];
                    Vy[buf][v][tj + H][ti] = pxy[4 + v];
                }
                if (hv) {
                    if (hx)
                        Vx[buf][v][hj][hi + H] = ph[v];
                    else
                        Vy[buf][v][hj + H][hi] = ph[v];
                }
            }
        }
        // loads for the next plane, in flight during this plane's work
        if (incell && p + 1 < ke + H)
#pragma unroll
            for (int v = 0; v < 4; v++) pz[v] = gld(B.fv[8 + v] + xc + (p + 1) * sz);
        if (k + 1 >= kb && k + 1 < ke) {
            const long long x1 = xc + (k + 1) * sz;
            if (hv)
#pragma unroll
                for (int v = 0; v < 4; v++) ph[v] = gld(B.fv[(hx ? 0 : 4) + v] + xh + (k + 1) * sz);
            if (incell) {
#pragma unroll
                for (int e = 0; e < 8; e++) pxy[e] = gld(B.fv[e] + x1);
#pragma unroll
                for (int c = 0; c < 5; c++) pr[c] = gld(B.rhs[c] + x1);
            }
        }
        __syncthreads();
        if (k < kb || !incell) continue;
        const long long x = xc + k * sz;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            double col[2 * H + 1];
#pragma unroll
            for (int a = 0; a < 2 * H + 1; a++) col[a] = zr[a][v];
            const double gx = line_diff_lds<HD == 4>(B, K, &Vx[buf][v][tj][ti + H], 1, 0, i);
            const double gy = line_diff_lds<HD == 4>(B, K, &Vy[buf][v][tj + H][ti], TX, 1, j);
            const double gz = line_diff_lds<HD == 4>(B, K, &col[H], 1, 2, k);
            r[1 + v] = ((r[1 + v] - gx) - gy) - gz;
        }
        if (MODE == 0) {
#pragma unroll
            for (int c = 0; c < 5; c++) gst(B.rhs[c] + x, r[c]);
        } else {
            rk_update(blocks, B, K, P, i, j, k, x, r, stage, bin, bun, bout, dt, lmax, nfix, r2);
        }
    }
    if (MODE == 1) {
        nfix.flush(counters);
        if (stage == 2) block_max3_atomic(lmax, lam_next + 3 * tl.b);
        if (stage == 2 && K.resid) block_sum_store(r2, K.resid + xcd_tile(blockIdx.x, gridDim.x));
    }
}

// ---------------------------------------------------------------------------
// Diagnostics of postproc::solveQ (postproc.cpp:743-848), after k_prim has
// written u, v, w, T with their ghosts: velocity gradients (the derivative
// scheme with closures, chain rule in the order of postproc.cpp:1824-1951),
// vorticity (wy-vz, uz-wx, vx-uy), enstrophy density |w|^2/2, Q-criterion,
// and per workgroup the sums of |w|^2/2 and rho|u|^2/2 (partial[2 wg],
// partial[2 wg + 1]; the host adds them in order).  One workgroup per
// (tile, plane) as k_prim.  For block bf, fields (5N, compact i-fastest)
// receives the vorticity, enstrophy density and Q-criterion.
__global__ __launch_bounds__(256) void KNAME(k_diag)(const nuc3d_block_dev *__restrict__ blocks,
                                                     const nuc3d_tile *__restrict__ tiles, nuc3d_kparams K, int bin,
                                                     int bf, double *__restrict__ fields, double *__restrict__ partial)
{
    constexpr int T = NUC3D_TILE_X;
    const nuc3d_tile tl = tiles[xcd_tile(blockIdx.x, gridDim.x)];
    const nuc3d_block_dev &B = blocks[tl.b];
    const int i = tl.i0 + (threadIdx.x & (T - 1)), j = tl.j0 + threadIdx.x / T, k = tl.k0 + blockIdx.y;
    double ens = 0.0, kin = 0.0;
    if ((int)blockIdx.y < tl.kn && i < B.n[0] && j < B.n[1]) {
        const long long x = cidx(B, i, j, k);
        double g[3][3];
#pragma unroll
        for (int v = 0; v < 3; v++) {
            const double *f = B.prim[v] + x;
            const double a = line_diff(B, K, f, 0, i), bb = line_diff(B, K, f, 1, j), cc = line_diff(B, K, f, 2, k);
            g[v][0] = a * met_at(B, 0, x) + bb * met_at(B, 3, x) + cc * met_at(B, 6, x);
            g[v][1] = a * met_at(B, 1, x) + bb * met_at(B, 4, x) + cc * met_at(B, 7, x);
            g[v][2] = a * met_at(B, 2, x) + bb * met_at(B, 5, x) + cc * met_at(B, 8, x);
        }
        const double ux = g[0][0], uy = g[0][1], uz = g[0][2];
        const double vx = g[1][0], vy = g[1][1], vz = g[1][2];
        const double wx = g[2][0], wy = g[2][1], wz = g[2][2];
        const double w0 = wy - vz, w1 = uz - wx, w2 = vx - uy;
        const double q0 = ux * vy + ux * wz + vy * wz - uy * vx - uz * wx - vz * wy;
        const double rho = B.q[bin][0][x] * jac_at(B, x);  // physicsModel.cpp:313-353
        const double u = B.prim[0][x], v = B.prim[1][x], w = B.prim[2][x];
        ens = 0.5 * (w0 * w0 + w1 * w1 + w2 * w2);
        kin = 0.5 * rho * (u * u + v * v + w * w);
        if (tl.b == bf && fields) {
            const long long N = (long long)B.n[0] * B.n[1] * B.n[2];
            const long long c = i + (long long)B.n[0] * (j + (long long)B.n[1] * k);
            fields[c] = w0;
            fields[N + c] = w1;
            fields[2 * N + c] = w2;
            fields[3 * N + c] = (w0 * w0 + w1 * w1 + w2 * w2) * 0.5;
            fields[4 * N + c] = q0;
        }
    }
    __shared__ double red[2][4];
    for (int o = 32; o > 0; o >>= 1) {
        ens += __shfl_xor(ens, o);
        kin += __shfl_xor(kin, o);
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        red[0][wv] = ens;
        red[1][wv] = kin;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        const long long wg = blockIdx.y * (long long)gridDim.x + blockIdx.x;
        partial[2 * wg + threadIdx.x] =
            (red[threadIdx.x][0] + red[threadIdx.x][1]) + (red[threadIdx.x][2] + red[threadIdx.x][3]);
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
    constexpr int T = NUC3D_TILE_X;
    const nuc3d_tile tl = tiles[xcd_tile(blockIdx.x, gridDim.x)];
    const nuc3d_block_dev &B = blocks[tl.b];
    const PhysParams P = phys_of(K, B);
    const int i = tl.i0 + (threadIdx.x & (T - 1)), j = tl.j0 + threadIdx.x / T;
    double lmax[3] = {0.0, 0.0, 0.0};
    if (i < B.n[0] && j < B.n[1]) {
        for (int k = tl.k0; k < tl.k0 + tl.kn; k++) {
            const long long x = cidx(B, i, j, k);
            double Q[5];
#pragma unroll
            for (int c = 0; c < 5; c++) Q[c] = B.q[buf][c][x];
            double rho, u, v, w, p, cs;
            con2prim(Q, jac_at(B, x), P, rho, u, v, w, p, cs);
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const double mx = met_at(B, 3 * d, x), my = met_at(B, 3 * d + 1, x), mz = met_at(B, 3 * d + 2, x);
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
    constexpr int T = NUC3D_TILE_X;
    const nuc3d_tile tl = tiles[xcd_tile(blockIdx.x, gridDim.x)];
    const nuc3d_block_dev &B = blocks[tl.b];
    const PhysParams P = phys_of(K, B);
    const int i = tl.i0 + (threadIdx.x & (T - 1)), j = tl.j0 + threadIdx.x / T;
    if (i >= B.n[0] || j >= B.n[1]) return;
    StageCounts nfix;
    for (int k = tl.k0; k < tl.k0 + tl.kn; k++) {
        const long long x = cidx(B, i, j, k);
        double Q[5];
#pragma unroll
        for (int c = 0; c < 5; c++) Q[c] = B.q[buf][c][x];
        if (fixup_energy(Q, jac_at(B, x), P, nfix)) B.q[buf][4][x] = Q[4];
    }
    nfix.flush(counters);
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
// Workgroups of a direction-sweep launch: one per tile, or for a persistent
// launch (K.pq set) as many as the GPU holds at once: this kernel's
// workgroups per CU (its registers and LDS) times the CUs
#ifndef NUC3D_PERSIST_X
#define NUC3D_PERSIST_X 1
#endif
static int sweep_grid(const void *kernel, int ntiles, const nuc3d_kparams &K)
{
    if (!K.pq) return ntiles;
    int dev = 0, ncu = 0, nb = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, 256, 0) != hipSuccess || ncu < 1 || nb < 1)
        return ntiles;
    return std::min(ntiles, NUC3D_PERSIST_X * nb * ncu);
}

template <int DIR, int SCHEME, bool NS, int MODE>
static void launch_sweep(const nuc3d_block_dev *blocks, const int4 *tiles, int ntiles, nuc3d_kparams K, int stage,
                         int bin, int bun, int bout, double *dtbuf, const unsigned long long *lam_cur,
                         unsigned long long *lam_next, unsigned long long *counters, hipStream_t st, bool decide,
                         int nf)
{
    using Kernel = void (*)(const nuc3d_block_dev *, const int4 *, nuc3d_kparams, int, int, int, int, double *,
                            const unsigned long long *, unsigned long long *, unsigned long long *);
    // persistent launches exist for the WENO5-JS / WENO5-Z sweeps that keep
    // their occupancy in that form (x and y, and the Euler z sweep: 77-79
    // VGPRs, six workgroups per CU; the Navier-Stokes z sweep, which also
    // takes the viscous divergence and the RK stage, needed 148 VGPRs in that
    // form, three workgroups per CU against five); every other kernel takes
    // one workgroup per tile
    auto go = [&](Kernel kern, bool pers = false) {
        nuc3d_kparams Kl = K;
        if (!pers) Kl.pq = nullptr;
        hipLaunchKernelGGL(kern, dim3(sweep_grid(reinterpret_cast<const void *>(kern), ntiles, Kl)), dim3(256), 0, st,
                           blocks, tiles, Kl, stage, bin, bun, bout, dtbuf, lam_cur, lam_next, counters);
    };
    // the fused viscous z sweep of a case with cd8 directions takes the cd8
    // divergence (the kernels of cd6-only cases are built without it)
    if constexpr (NS && DIR == 2 && MODE == 1) {
        if (K.diff[0] || K.diff[1] || K.diff[2]) {
            if constexpr (SCHEME == SCH_HYBRID) {
                if (decide) go(KNAME(k_sweep_hybrid)<DIR, NS, MODE, true, true>);
                else go(KNAME(k_sweep_hybrid)<DIR, NS, MODE, false, true>);
            } else {
                go(KNAME(k_sweep)<DIR, SCHEME, NS, MODE, 16, true>);
            }
            return;
        }
    }
    if constexpr (SCHEME == SCH_WENO5JS || SCHEME == SCH_WENO5Z) {
        if (K.pq && !(NS && DIR == 2 && MODE == 1)) {
            if (nf == 32) go(KNAME(k_sweep)<DIR, SCHEME, NS, MODE, 32, false, true>, true);
            else go(KNAME(k_sweep)<DIR, SCHEME, NS, MODE, 16, false, true>, true);
            return;
        }
        if (nf == 32) {
            go(KNAME(k_sweep)<DIR, SCHEME, NS, MODE, 32>);
            return;
        }
    }
    if constexpr (SCHEME == SCH_HYBRID) {
        if constexpr (DIR < 2) {
            if (nf == 32) {
                if (decide) go(KNAME(k_sweep_hybrid)<DIR, NS, MODE, true, false, 32>);
                else go(KNAME(k_sweep_hybrid)<DIR, NS, MODE, false, false, 32>);
                return;
            }
        }
        if (decide) go(KNAME(k_sweep_hybrid)<DIR, NS, MODE, true>);
        else go(KNAME(k_sweep_hybrid)<DIR, NS, MODE, false>);
    } else {
        go(KNAME(k_sweep)<DIR, SCHEME, NS, MODE>);
    }
}

template <int SCHEME, bool NS>
static void sweep_dispatch(int dir, int mode, int nf, const nuc3d_block_dev *blocks, const int4 *tiles, int ntiles,
                           nuc3d_kparams K, int stage, int bin, int bun, int bout, double *dtbuf,
                           const unsigned long long *lam_cur, unsigned long long *lam_next,
                           unsigned long long *counters, hipStream_t st)
{
    // hybrid decisions: made where the host asks (a single RHS evaluation, and
    // stage 0 of a step at the case's decision interval), reused otherwise
    const bool decide = K.hdecide != 0;
    (void)mode;
#define NUC3D_A blocks, tiles, ntiles, K, stage, bin, bun, bout, dtbuf, lam_cur, lam_next, counters, st, decide, nf
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
// nf: faces per line of the tiles (16, or 32 for WENO5-JS / WENO5-Z)
extern "C" hipError_t KNAME(nuc3d_launch_sweep)(int dir, int scheme, int ns, int mode, int nf,
                                                const nuc3d_block_dev *blocks, const int4 *tiles, int ntiles,
                                                nuc3d_kparams K, int stage, int bin, int bun, int bout, double *dtbuf,
                                                const unsigned long long *lam_cur, unsigned long long *lam_next,
                                                unsigned long long *counters, hipStream_t st)
{
    if (nf != 16 && !(nf == 32 && (scheme == SCH_WENO5JS || scheme == SCH_WENO5Z || (scheme == SCH_HYBRID && dir < 2))))
        return hipErrorInvalidValue;
#define NUC3D_A dir, mode, nf, blocks, tiles, ntiles, K, stage, bin, bun, bout, dtbuf, lam_cur, lam_next, counters, st
    if (scheme == SCH_WENO5JS) {
        if (ns) sweep_dispatch<SCH_WENO5JS, true>(NUC3D_A);
        else sweep_dispatch<SCH_WENO5JS, false>(NUC3D_A);
    } else if (scheme == SCH_WENO5Z) {
        if (ns) sweep_dispatch<SCH_WENO5Z, true>(NUC3D_A);
        else sweep_dispatch<SCH_WENO5Z, false>(NUC3D_A);
    } else if (scheme == SCH_CENTRAL8) {
        if (ns) sweep_dispatch<SCH_CENTRAL8, true>(NUC3D_A);
        else sweep_dispatch<SCH_CENTRAL8, false>(NUC3D_A);
    } else if (scheme == SCH_HYBRID) {
        if (!K.sensor || !blocks) return hipErrorInvalidValue;
        if (ns) sweep_dispatch<SCH_HYBRID, true>(NUC3D_A);
        else sweep_dispatch<SCH_HYBRID, false>(NUC3D_A);
    } else {
        return hipErrorInvalidValue;
    }
#undef NUC3D_A
    return hipGetLastError();
}

template <int DIR, bool NS, int MODE>
static hipError_t launch_crweno(const nuc3d_block_dev *blocks, const int4 *lines, int nlines, int nmax,
                                nuc3d_kparams K, int stage, int bin, int bun, int bout, double *dtbuf,
                                const unsigned long long *lam_cur, unsigned long long *lam_next,
                                unsigned long long *counters, double *scratch, hipStream_t st)
{
    hipLaunchKernelGGL((KNAME(k_crweno)<DIR, NS, MODE>), dim3(nlines / NUC3D_CRW_LINES), dim3(64), 0, st, blocks,
                       lines, K, stage, bin, bun, bout, dtbuf, lam_cur, lam_next, counters, scratch, nmax);
    return hipGetLastError();
}

extern "C" hipError_t KNAME(nuc3d_launch_crweno)(int dir, int ns, int mode, const nuc3d_block_dev *blocks,
                                                 const int4 *lines, int nlines, int nmax, nuc3d_kparams K,
                                                 int stage, int bin, int bun, int bout, double *dtbuf,
                                                 const unsigned long long *lam_cur, unsigned long long *lam_next,
                                                 unsigned long long *counters, double *scratch, hipStream_t st)
{
    if (nmax > 1024 || nlines % NUC3D_CRW_LINES) return hipErrorInvalidValue;
#define NUC3D_A blocks, lines, nlines, nmax, K, stage, bin, bun, bout, dtbuf, lam_cur, lam_next, counters, scratch, st
    if (dir == 0) return ns ? launch_crweno<0, true, 0>(NUC3D_A) : launch_crweno<0, false, 0>(NUC3D_A);
    if (dir == 1) return ns ? launch_crweno<1, true, 0>(NUC3D_A) : launch_crweno<1, false, 0>(NUC3D_A);
    if (mode == 0) return ns ? launch_crweno<2, true, 0>(NUC3D_A) : launch_crweno<2, false, 0>(NUC3D_A);
    return ns ? launch_crweno<2, true, 1>(NUC3D_A) : launch_crweno<2, false, 1>(NUC3D_A);
#undef NUC3D_A
}

// kmax: the longest k segment of the tiles (one workgroup per tile and plane)
extern "C" hipError_t KNAME(nuc3d_launch_prim)(const nuc3d_block_dev *blocks, const nuc3d_tile *tiles, int ntiles,
                                               int kmax, nuc3d_kparams K, int bin, hipStream_t st)
{
    hipLaunchKernelGGL(KNAME(k_prim), dim3(ntiles, kmax), dim3(256), 0, st, blocks, tiles, K, bin);
    return hipGetLastError();
}

extern "C" hipError_t KNAME(nuc3d_launch_visc_flux)(const nuc3d_block_dev *blocks, const nuc3d_tile *tiles,
                                                    int ntiles, nuc3d_kparams K, int bin, hipStream_t st)
{
    const bool wide = K.diff[0] || K.diff[1] || K.diff[2];
    if (wide && K.alluni)
        hipLaunchKernelGGL((KNAME(k_visc_flux)<4, true>), dim3(ntiles), dim3(256), 0, st, blocks, tiles, K, bin);
    else if (wide)
        hipLaunchKernelGGL((KNAME(k_visc_flux)<4, false>), dim3(ntiles), dim3(256), 0, st, blocks, tiles, K, bin);
    else if (K.alluni)
        hipLaunchKernelGGL((KNAME(k_visc_flux)<3, true>), dim3(ntiles), dim3(256), 0, st, blocks, tiles, K, bin);
    else
        hipLaunchKernelGGL((KNAME(k_visc_flux)<3, false>), dim3(ntiles), dim3(256), 0, st, blocks, tiles, K, bin);
    return hipGetLastError();
}

extern "C" hipError_t KNAME(nuc3d_launch_visc_div)(int mode, const nuc3d_block_dev *blocks, const nuc3d_tile *tiles,
                                                   int ntiles, nuc3d_kparams K, int stage, int bin, int bun,
                                                   int bout, double *dtbuf, const unsigned long long *lam_cur,
                                                   unsigned long long *lam_next, unsigned long long *counters,
                                                   hipStream_t st)
{
    const bool wide = K.diff[0] || K.diff[1] || K.diff[2];
    if (mode && wide)
        hipLaunchKernelGGL((KNAME(k_visc_div)<1, 4>), dim3(ntiles), dim3(256), 0, st, blocks, tiles, K, stage, bin, bun,
                           bout, dtbuf, lam_cur, lam_next, counters);
    else if (mode)
        hipLaunchKernelGGL((KNAME(k_visc_div)<1, 3>), dim3(ntiles), dim3(256), 0, st, blocks, tiles, K, stage, bin, bun,
                           bout, dtbuf, lam_cur, lam_next, counters);
    else if (wide)
        hipLaunchKernelGGL((KNAME(k_visc_div)<0, 4>), dim3(ntiles), dim3(256), 0, st, blocks, tiles, K, stage, bin, bun,
                           bout, dtbuf, lam_cur, lam_next, counters);
    else
        hipLaunchKernelGGL((KNAME(k_visc_div)<0, 3>), dim3(ntiles), dim3(256), 0, st, blocks, tiles, K, stage, bin, bun,
                           bout, dtbuf, lam_cur, lam_next, counters);
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

extern "C" hipError_t KNAME(nuc3d_launch_diag)(const nuc3d_block_dev *blocks, const nuc3d_tile *tiles, int ntiles,
                                               int kmax, nuc3d_kparams K, int bin, int bf, double *fields,
                                               double *partial, hipStream_t st)
{
    hipLaunchKernelGGL(KNAME(k_diag), dim3(ntiles, kmax), dim3(256), 0, st, blocks, tiles, K, bin, bf, fields,
                       partial);
    return hipGetLastError();
}
