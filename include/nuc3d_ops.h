/*
 * nuc3d_ops.h -- operator-level C ABI: the reference's scheme plugins
 * (fieldOperator.h:15-80) on the GPU, one call per Field operation.
 *
 * These calls implement, for whole host Fields, the private virtual methods
 * that fieldOperator3d (fieldOperator.cpp:371-633) calls on its scheme
 * objects:
 *   interoplation::interpolationInner / interpolationBoundaryL / ...BoundaryR
 *       (weno5js.cpp, weno5z.cpp, crweno5.cpp, reconstructionBoundaryScheme.cpp)
 *   differential::differentialInner / differentialBoundaryL / ...BoundaryR
 *       (centraldifference6th.cpp, centraldifference8th.cpp, boundaryScheme.cpp)
 *   integration::integrationAll (TVD-RK3.cpp)
 * The C++ classes under nuc3d_amd/shim/ wrap them with the reference's class
 * names and signatures, so the reference's own driver can run on them.
 *
 * Layout is that of field3d<double> (field.h:11-81): a Field of nx*ny*nz
 * doubles, i-fastest; the operator acts along x on the ny*nz lines (the
 * reference passes eta and zeta data as rotated Fields, euler3d.cpp:75-80).
 * Reconstruction output has nx+1 faces per line (face m between cells m-1 and
 * m).  A ghost Field holds w cells per line: the left one cells -w..-1 at
 * slots 0..w-1, the right one cells nx..nx+w-1 (weno5js.cpp:255-266).
 *
 * All pointers are host pointers.  Each call copies its inputs to the GPU,
 * runs the kernel in the reference's operation order and copies back only
 * the faces or cells the operation defines (other entries of `out` are left
 * as they were, as in the reference).  Calls are synchronous; `stream` (a
 * hipStream_t as void*, may be NULL) is the stream they run on.  Status codes
 * are those of nuc3d_hip.h.
 */
#ifndef NUC3D_OPS_H
#define NUC3D_OPS_H

#ifdef __cplusplus
extern "C" {
#endif

/* which part of a line an operation writes */
enum {
    NUC3D_PART_INNER = 0, /* faces w..nx-w (reconstruction), cells w..nx-w-1 (derivative) */
    NUC3D_PART_LEFT = 1,  /* faces 0..w-1 / cells 0..w-1, using the left ghost Field */
    NUC3D_PART_RIGHT = 2  /* faces nx-w+1..nx / cells nx-w..nx-1, using the right ghost Field */
};

/* Reconstruction schemes of nuc3d_op_interpolate: NUC3D_WENO5JS,
 * NUC3D_WENO5Z, NUC3D_CRWENO5 (nuc3d_hip.h) and the physical-face closure
 * ReconstructionboundaryScheme (reconstructionBoundaryScheme.cpp:31-474),
 * which has only the LEFT and RIGHT parts. */
#define NUC3D_RECON_BOUNDARY 3

/* Derivative schemes of nuc3d_op_differentiate: NUC3D_CD6, NUC3D_CD8
 * (nuc3d_hip.h; CD8 needs w = 4) and the physical-face closure boundaryScheme
 * (boundaryScheme.cpp:31-126), LEFT and RIGHT parts only. */
#define NUC3D_DIFF_BOUNDARY 2

/* interpolationInner(in, upwind, out, w)             part = INNER, ghost = NULL
 * interpolationBoundaryL(in, boundaryL, upwind, out, w)  part = LEFT
 * interpolationBoundaryR(in, boundaryR, upwind, out, w)  part = RIGHT
 * (fieldOperator.h:23-38).  upwind = +1 or -1.  zquirk = 1 reproduces the
 * beta2 of WENO5-Z's + inner faces as written in weno5z.cpp:80 (0: the
 * textbook beta2); it has no effect on other schemes and parts.
 * in: nx*ny*nz, ghost: w*ny*nz, out: (nx+1)*ny*nz. */
int nuc3d_op_interpolate(int scheme, int part, int upwind, int w, int zquirk, int nx, int ny, int nz,
                         const double *in, const double *ghost, double *out, void *stream);

/* differentialInner(in, out, w)                   part = INNER, ghost = NULL
 * differentialBoundaryL(in, boundaryL, out, w)    part = LEFT
 * differentialBoundaryR(in, boundaryR, out, w)    part = RIGHT
 * (fieldOperator.h:68-79), unit grid spacing.  in, out: nx*ny*nz,
 * ghost: w*ny*nz (ignored by NUC3D_DIFF_BOUNDARY). */
int nuc3d_op_differentiate(int scheme, int part, int w, int nx, int ny, int nz, const double *in,
                           const double *ghost, double *out, void *stream);

/* tvdrk3rd::integrationAll(RHS, u_n, u_i, dt, stage) (TVD-RK3.cpp:12-137)
 * over nfields fields of n doubles each:
 *   stage 0: u_i = u_i - dt RHS
 *   stage 1: u_i = 3/4 u_n + 1/4 (u_i - dt RHS)
 *   stage 2: u_i = 1/3 u_n + 2/3 (u_i - dt RHS)
 * u_n is not read at stage 0 (may be NULL there). */
int nuc3d_op_rk3(int stage, double dt, int nfields, long long n, const double *const *rhs,
                 const double *const *u_n, double *const *u_i, void *stream);

#ifdef __cplusplus
}
#endif
#endif
