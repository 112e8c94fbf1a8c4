// nuc3d_types.h -- data layout shared by the HIP kernels and the host side.
//
// Every block lives on the device as SoA FP64 fields in a padded box of
// (nx+2G)(ny+2G)(nz+2G) cells, i-fastest.  Field pointers point at cell
// (0,0,0), so cell (i,j,k) with i in [-G, nx+G) is ptr[i + sy*j + sz*k].
// Only the six face slabs of the ghost frame are used (the stencils are
// one-dimensional); corner and edge ghosts are never read.
#pragma once
#include <stdint.h>

#define NUC3D_MAX_BLOCKS 256
// column tiles: NUC3D_TILE_X columns in i x NUC3D_TILE_Y in j, one thread per
// column (256 threads).  32 x 8 measures the same as 16 x 16 on 256^3.
#ifndef NUC3D_TILE_X
#define NUC3D_TILE_X 16
#endif
#define NUC3D_TILE_Y (256 / NUC3D_TILE_X)
#define NUC3D_HALO 3    // stencil half-width of WENO5 / cd6
#define NUC3D_KSEG 32   // planes per k segment of a column tile
#ifndef NUC3D_CRW_LINES
#define NUC3D_CRW_LINES 6  // CRWENO5 lines per wavefront (ten tridiagonal systems each)
#endif

struct nuc3d_block_dev {
    int n[3];            // interior cells nx, ny, nz
    int G;               // ghost layers in storage (the reference's bufferWidth)
    long long sy, sz;    // strides of j and k (stride of i is 1)
    double *q[3][5];     // three state buffers (conserved / J)
    double *rhs[5];      // RHS output (single RHS evaluations)
    double *jac;         // J = 1/det(dx/dxi)
    double *met[9];      // xi_x, xi_y, xi_z, eta_x .. zeta_z
    double *dxyz[9];     // x_xi, x_eta, x_zeta, y_xi .. z_zeta (symmetric BC only)
    double *prim[4];     // u, v, w, T (Navier-Stokes)
    double *fv[12];      // viscous flux: fv[4*d + c - 1], d = xi/eta/zeta, c = 1..4
    int ftype[6];        // -1: physical face, else neighbour block index
    int fid[6];          // BC id (0 inlet, 1 outlet, 2 wall, 3 symmetric) or neighbour face
    double fval[6][5];   // inlet primitive state per face
    unsigned char *hflag[3];  // hybrid decisions per face of each hybrid direction, else null
    // batch of independent cases: the block's case (group) takes its own dt
    // (the minimum over that group's blocks), and its own Mach and Reynolds
    // numbers (the case's by default)
    int group;
    int pad_g;
    double mach, reynolds;
    // uniform mesh: J and every metric term equal jac0 and met0 in the
    // interior and in the ghost slabs of the neighbour faces (uniform
    // Cartesian blocks); the stage kernels then take them from here instead
    // of loading the fields (the fields stay valid for every other user)
    int uniform;
    int pad_u;
    double jac0, met0[9];
    // uniform blocks: |grad xi|, |grad eta|, |grad zeta| of the constant metric
    // terms and the reciprocal half of a division by jac0 (recip_nr), both
    // computed on the device (k_block_consts); rdsplit: jac0 and every metric
    // term lie where the split's divisions by J may share that reciprocal
    double theta0[3], rjac0;
    int rdsplit;
    int pad_r;
};

// Learned sensor of the hybrid reconstruction: 30 features -> 32 -> 32 -> 1.
#define NUC3D_SENSOR_NFEAT 30
#define NUC3D_SENSOR_NH 32
// master parameters (FP32): W1[32][32] (inputs 30, 31 unused), b1[32],
// W2[32][32], b2[32], w3[32], b3
#ifndef NUC3D_SENSOR_NPARAM
#define NUC3D_SENSOR_NPARAM 2145
#endif
static_assert(NUC3D_SENSOR_NPARAM == 32 * 32 + 32 + 32 * 32 + 32 + 32 + 1, "sensor parameter count");

// Inference copy of the network: BF16 weight bits (round to nearest even),
// FP32 biases, w3 as the FP32 values of its BF16 rounding, and the logit
// threshold at and above which a face takes WENO5-JS.
struct nuc3d_sensor_dev {
    unsigned short w1[32 * 32];  // [hidden][input]
    unsigned short w2[32 * 32];  // [hidden][hidden]
    float b1[32], b2[32], w3[32];
    float b3, thr;
    float pad[2];
    // the same values arranged per lane of a wavefront as the MFMA operands
    // of the inference (nuc3d_sensor.h, load_regs_lanes): lane l of lane group
    // g = l/16 holds W1 row 16mt + l%16, inputs 8g..8g+7; W2 row 16mt + l%16
    // and w3, both at hidden inputs 4g..4g+3 and 16+4g..16+4g+3; b1, b2 of
    // hidden 16mt + 4g .. 16mt + 4g + 3
    uint4 lw1[2][64], lw2[2][64], lw3[64];
    float4 lb1[2][64], lb2[2][64];
};

struct nuc3d_tile {
    int b, i0, j0, k0, kn, pad0, pad1, pad2;
};

struct nuc3d_kparams {
    double gamma, mach, reynolds, prandtl, sref, twall;
    double cfl;
    int riemann;
    int width[3];
    int zquirk;
    int nblocks;
    int ngroups;    // independent cases among the blocks (dt per group)
    int hdecide;    // hybrid sweeps of this launch make decisions (else reuse the stored ones)
    int hrhs;       // a single RHS evaluation: its decisions go to the second half of each decision
                    // array, so that the step's own decisions (first half) stay untouched
    int ieee_div;   // 1: the compiler's IEEE division everywhere (no reduced division)
    int diff[3];    // viscous derivative per direction: 0 cd6, 1 cd8 (needs width 4)
    int alluni;     // every block takes the uniform-mesh path (J and metrics from the descriptors)
    double *resid;  // stage 2: per-work-item partial sums of (Q - Q_n)^2, or null
    const nuc3d_sensor_dev *sensor;  // hybrid reconstruction: the network, or null
    // hybrid decision counts (central, WENO5-JS by the network, by the
    // fallback), spread over NUC3D_HCOUNT_SLOTS slots of 16 words so that
    // the atomics of different workgroups do not meet on one address
    unsigned long long *hcount;
    // persistent direction sweeps: eight per-XCD queue counters and an exit
    // counter (zero when a launch starts; its last workgroup zeroes them), and
    // the number of work items (tiles) of the launch; null: one workgroup per
    // tile
    unsigned *pq;
    int pq_n;
    int pad_q;
    // range flags of the three state buffers: 1 where every cell of the buffer
    // passed momenta_in_range and gamma_p_in_range when the stage
    // that wrote it last wrote it, so that the divisions by rho of its readers
    // share one reciprocal; null: not kept (every reader divides by rho)
    unsigned *qok;
};
#define NUC3D_HCOUNT_SLOTS 256
#define NUC3D_QUEUE_WORDS 16
