/*
 * nuc3d_oracle.h -- CPU restatement of NUC3d's per-block RHS and TVD-RK3 path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; nothing in nuc3d_amd/ links or calls it.
 *
 * The restatement follows the reference's operation order (unfused FP64, libm
 * pow/sqrt; build with -ffp-contract=off -fno-builtin) so that it agrees with
 * the reference binary to the last few ulps.  Every function cites the
 * reference file:line it restates.
 *
 * Layouts (as in the reference):
 *   xi   arrays (nx,ny,nz): idx = k*nx*ny + j*nx + i        (euler3d.cpp:389)
 *   eta  arrays (ny,nz,nx): idx = i*ny*nz + k*ny + j        (euler3d.cpp:403)
 *   zeta arrays (nz,nx,ny): idx = j*nz*nx + i*nz + k        (euler3d.cpp:417)
 * All public entry points take and return xi-layout (i-fastest) fields.
 */
#ifndef NUC3D_ORACLE_H
#define NUC3D_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_WENO5JS = 0, ORC_WENO5Z = 1, ORC_CRWENO5 = 2, ORC_CENTRAL8 = 3, ORC_HYBRID = 4 };
enum { ORC_CD6 = 0, ORC_CD8 = 1 };
enum { ORC_LF = 0, ORC_AUSM = 1, ORC_AUSMP = 2 };
enum { ORC_EULER = 0, ORC_NS = 1 };
enum { ORC_BC_INLET = 0, ORC_BC_OUTLET = 1, ORC_BC_WALL = 2, ORC_BC_SYMM = 3 };

typedef struct {
    double gamma, mach, reynolds, prandtl, t_ref, t_inf, t_wall;
    int model;      /* ORC_EULER / ORC_NS */
    int riemann;    /* ORC_LF / ORC_AUSM / ORC_AUSMP */
    int scheme[3];  /* reconstruction per direction */
    int diff[3];    /* viscous difference per direction */
    int width[3];   /* ghost width per direction (fieldOperator.cpp:127-177) */
    int zquirk;     /* 1: reproduce weno5z.cpp:80 (+inner beta2 uses f[i+1]) */
} orc_params;

typedef struct orc_case orc_case;

/* Case with nblocks blocks; dims = nblocks*3 cell counts. Returns 0 or <0. */
/* How the reference's pow(x, 2) squares are restated (nuc3d_oracle.c): 0 libm
 * pow (the makefile's -O0 build), 1 literal-exponent squares as x*x (the -O2
 * build), 2 every square as x*x.  Process-wide; default 0. */
enum { ORC_SQ_LIBM = 0, ORC_SQ_O2 = 1, ORC_SQ_EXACT = 2 };
void orc_set_square_mode(int mode);
int orc_get_square_mode(void);

int orc_case_create(const orc_params *p, int nblocks, const int *dims, orc_case **out);
void orc_case_destroy(orc_case *c);

/* Face topology of block b (boundaryConditions.cpp:25-93): type[f] = -1 for a
 * physical face (id = BC id, val = 5 primitive values for inlets) or the
 * neighbour block index (id = neighbour face). */
int orc_set_faces(orc_case *c, int b, const int *type6, const int *id6, const double *val30);
/* jac N; met 9N: xi_x,xi_y,xi_z,eta_x,eta_y,eta_z,zeta_x,zeta_y,zeta_z;
 * dxyz 9N: x_xi,x_eta,x_zeta,y_xi,y_eta,y_zeta,z_xi,z_eta,z_zeta (may be NULL). */
int orc_set_metrics(orc_case *c, int b, const double *jac, const double *met, const double *dxyz);
int orc_set_q(orc_case *c, int b, const double *q5);      /* Q_work = conserved / J */
int orc_get_q(const orc_case *c, int b, double *q5);
int orc_get_rhs(const orc_case *c, int b, double *rhs5);
int orc_get_prim(const orc_case *c, int b, double *w8);    /* rho,u,v,w,p,T,e,c */
double orc_get_dt_local(const orc_case *c, int b);
double orc_get_dt_global(const orc_case *c);
int orc_get_max_eigen(const orc_case *c, int b, double *eig3);
long orc_get_fixups(const orc_case *c);                    /* con2prim e<0 fix-ups */
/* con2prim inputs with density <= 0 (or NaN) and with NaN pressure after the
 * fix-up; a negative pressure makes orc_stage return -10 (the reference exits) */
void orc_get_positivity(const orc_case *c, long *counts2);

/* Decisions of the hybrid reconstruction (ORC_HYBRID, no counterpart in the
 * reference) for direction d of block b, in the layout of the reconstructed
 * flux of that direction: (n_d+1) x n_{d+1} x n_{d+2}, face index fastest.
 * Faces flagged 1 or 2 take WENO5-JS, 0 the eighth-order central value;
 * physical-face closures ignore the flag.  NULL (the default) is all WENO5-JS.
 * The array is copied. */
int orc_set_hybrid_mask(orc_case *c, int b, int d, const unsigned char *mask);

/* One RHS evaluation of every block: EulerData3D::solve / NaiverStokesData3d::solve. */
int orc_rhs(orc_case *c);
/* One RK stage: RHS then PDEData3d::solve (PDEData3d.cpp:70-88). */
int orc_stage(orc_case *c, int stage, double cfl);
/* nsteps full TVD-RK3 steps (block.cpp:569-594); dt of each step to dts (may be NULL). */
int orc_steps(orc_case *c, int nsteps, double cfl, double *dts);
/* residual of the last step (PDEData3d::setRES): block b, or the maximum over
 * blocks for b < 0 */
double orc_get_residual(const orc_case *c, int b);

/* ---- line-level operators (one line, ghosts included) ----
 * ext has n + 2w entries: ext[w+i] = cell i, ext[w-1-g] = left ghost g, ext[w+n+g] = right ghost g.
 * typeL/typeR = -1 selects the exterior (physical) closures. */
int orc_recon_line(int scheme, int upwind, int n, int w, const double *ext,
                   int typeL, int typeR, int zquirk, double *out /* n+1 */);
int orc_diff_line(int scheme, int n, int w, const double *ext,
                  int typeL, int typeR, double *out /* n */);

/* Pointwise splitters (physicsModel.cpp:940-979 and :1323-1398). */
void orc_split_point(const orc_params *p, const double prim5[5], double jac,
                     double mx, double my, double mz, double fl[5], double fr[5]);

/* TVD-RK3 on flat arrays (TVD-RK3.cpp:37-137). */
void orc_rk3(int stage, long n, const double *rhs, const double *un, double *ui, double dt);

/* postproc::solveQ diagnostics of the current state (postproc.cpp:743-848):
 * sums over all cells of all blocks of |vorticity|^2/2 and rho|u|^2/2; for
 * block bf, fields5 (5N, may be NULL) receives the vorticity (3 fields), the
 * enstrophy density and the Q-criterion. */
int orc_diagnostics(orc_case *c, double *enstrophy, double *kinetic, int bf, double *fields5);

/* Cell centres (3N), coordinate derivatives (9N: x_xi,x_eta,x_zeta,y_xi..z_zeta),
 * J (N) and metric terms (9N) of an nx x ny x nz block from its node
 * coordinates ((nx+1)(ny+1)(nz+1) each, i-fastest): block.cpp:147-477.
 * Returns -1 unless every dimension is a multiple of 4 (the tile size),
 * -2 at a negative Jacobian (where the reference exits). */
int orc_mesh_metrics(int nx, int ny, int nz, const double *xn, const double *yn, const double *zn, double *centre,
                     double *dxyz, double *jac, double *met);

#ifdef __cplusplus
}
#endif
#endif
