/*
 * nuc3d_hip.h -- C ABI of the MI355X (gfx950) implementation of NUC3d's
 * per-block right-hand side and TVD-RK3 update.
 *
 * The ABI replaces, for all blocks of a case held on one GPU, the calls that
 * the reference's block::solve (block.cpp:569-594) makes per RK stage:
 *   myFluxes->solve(myPDE, myOP, mybuffer, myPhys, myMPI, myBC)
 *       -> EulerData3D::solve (euler3d.cpp:248-262) or
 *          NaiverStokesData3d::solve (NaiverStokes3d.cpp:36-49)
 *   myPDE.solve(myOP, cfl, stage)          -> PDEData3d::solve (PDEData3d.cpp:70-88)
 * together with the MPI halo exchange they contain
 * (MPICommunicator.cpp:36-115), which becomes on-device ghost copies.
 *
 * Plain C types only: host pointers and sizes, an opaque case handle and an
 * optional hipStream_t passed as void*.  Every call returns a status (0 = OK);
 * nuc3d_last_error() describes the last failure of the calling thread.  No
 * call exits the process (the reference's exit(-1)/MPI_Abort paths become
 * status codes).
 *
 * Field arrays are float64, i-fastest (index k*nx*ny + j*nx + i) as in
 * field3d<double> (field.h:11-81, field.cpp:95-108); a group of fields is
 * stored one after the other (q5 = five fields of nx*ny*nz).
 */
#ifndef NUC3D_HIP_H
#define NUC3D_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    NUC3D_OK = 0,
    NUC3D_ERR_BAD_ARG = -1,     /* invalid argument (reference: exit(-1)/exit(0)) */
    NUC3D_ERR_OOM = -2,         /* device allocation failed */
    NUC3D_ERR_HIP = -3,         /* a HIP runtime call failed */
    NUC3D_ERR_UNSUPPORTED = -4, /* configuration outside what this build implements */
    NUC3D_ERR_NO_DEVICE = -5,   /* no GPU visible */
    NUC3D_ERR_ORDER = -6,       /* RK stages called out of order */
    NUC3D_ERR_IO = -7,          /* a restart file could not be written or read */
    NUC3D_ERR_NEG_DENSITY = -8, /* a state with non-positive (or NaN) density: see nuc3d_get_positivity */
    NUC3D_ERR_NEG_PRESSURE = -9 /* a state with negative (or NaN) pressure after the e<0 fix-up
                                   (physicsModel.cpp:1009-1016 exits there) */
};

/* Schemes and models, named as in inp/Method.in and inp/PhysModel.in
 * (fieldOperator.cpp:127-369, physicsModel.cpp:13-112). */
enum { NUC3D_WENO5JS = 0, NUC3D_WENO5Z = 1, NUC3D_CRWENO5 = 2 };
/* Eighth-order central flux of each split part at face m from cells m-4..m+3,
 * (533(f3+f4) - 139(f2+f5) + 29(f1+f6) - 3(f0+f7)) * (1/840).  No counterpart in the
 * reference: the smooth-region branch of the hybrid reconstruction.  Needs
 * width 4 (four ghost layers, four closure faces at physical faces, which keep
 * the reference's closures). */
enum { NUC3D_CENTRAL8 = 3 };
/* Hybrid reconstruction, no counterpart in the reference: per interface the
 * NUC3D_CENTRAL8 face value in smooth flow and WENO5-JS elsewhere, as decided
 * by the learned sensor set with nuc3d_set_sensor (below); one decision per
 * face for all five components and both split parts, so each face flux is
 * computed once and differenced.  Needs width 4, like NUC3D_CENTRAL8. */
enum { NUC3D_HYBRID = 4 };
enum { NUC3D_CD6 = 0, NUC3D_CD8 = 1 };
enum { NUC3D_LF = 0, NUC3D_AUSM = 1, NUC3D_AUSMP = 2 };
enum { NUC3D_EULER3D = 0, NUC3D_NAVIERSTOKES3D = 1 };
enum { NUC3D_BC_INLET = 0, NUC3D_BC_OUTLET = 1, NUC3D_BC_WALL = 2, NUC3D_BC_SYMMETRIC = 3 };

typedef struct {
    /* physicsModel parameters (physicsModel.cpp:41-51) */
    double gamma, mach, reynolds, prandtl, t_ref, t_inf, t_wall;
    int model;     /* NUC3D_EULER3D / NUC3D_NAVIERSTOKES3D */
    int riemann;   /* flux splitting: NUC3D_LF / NUC3D_AUSM / NUC3D_AUSMP */
    int scheme[3]; /* reconstruction per direction (scheme_x_ivs ..); may differ */
    int diff[3];   /* viscous difference per direction (scheme_x_vis ..) */
    int width[3];  /* ghost width per direction (bufferSize_xi ..) */
    int zquirk;    /* 1: reproduce the +inner beta2 of weno5z.cpp:80 */
    int exact;     /* 1: reference operation order without FMA contraction */
} nuc3d_params;

typedef struct nuc3d_case nuc3d_case;

/* Allocate a case of nblocks blocks (dims = nblocks x {nx,ny,nz}) on the
 * current device.  Replaces block::initialData (block.cpp:479-520). */
int nuc3d_case_create(const nuc3d_params *p, int nblocks, const int *dims, nuc3d_case **out);
int nuc3d_case_destroy(nuc3d_case *c);

/* Face topology of block b: the six lines of bc_topo/BC_Topo_<b>.in
 * (boundaryConditions.cpp:25-93).  type6[f] = -1 for a physical face with
 * id6[f] the BC id (val30[5f..5f+4] = inlet primitives rho,u,v,w,p), else the
 * neighbour block with id6[f] its face.  Connected faces must join the same
 * direction and have equal tangential sizes (the raw slab copy of
 * MPICommunicator.cpp:55-81). */
int nuc3d_set_faces(nuc3d_case *c, int b, const int *type6, const int *id6, const double *val30);

/* Metrics of block b (block::getJacobians output, block.cpp:237-325):
 * jac (N), met (9N: xi_x,xi_y,xi_z,eta_x,..,zeta_z) and dxyz (9N:
 * x_xi,x_eta,x_zeta,y_xi,..,z_zeta; may be NULL unless a face is symmetric). */
int nuc3d_upload_metrics(nuc3d_case *c, int b, const double *jac, const double *met, const double *dxyz);

/* Metrics of block b computed on the GPU from its mesh nodes, as
 * block::initial does after reading mesh/mesh_<rank>.x (block.cpp:42-98):
 * cell centres by tiled 5-point Lagrange interpolation (getXYZ_center,
 * block.cpp:147-235), the nine coordinate derivatives, J and the metric terms
 * (getJacobians, block.cpp:237-325).  xn, yn, zn: node coordinates,
 * (nx+1)(ny+1)(nz+1) each, i-fastest.  Cells per direction must be multiples
 * of 4 (NUC3D_ERR_UNSUPPORTED otherwise; the reference leaves the remainder
 * cells unset).  A non-positive Jacobian is NUC3D_ERR_BAD_ARG (the reference
 * exits).  centre3 (3N, may be NULL) receives the cell centres, where
 * block::initialQ evaluates the initial state (block.cpp:523-567). */
int nuc3d_upload_mesh(nuc3d_case *c, int b, const double *xn, const double *yn, const double *zn, double *centre3);
/* J (N), metrics (9N) and coordinate derivatives (9N) of block b as held on
 * the GPU (any pointer may be NULL). */
int nuc3d_download_metrics(nuc3d_case *c, int b, double *jac, double *met, double *dxyz);

/* State of block b: Q_work = conserved / J (5N), PDEData3d::getQ. */
int nuc3d_upload_q(nuc3d_case *c, int b, const double *q5);
int nuc3d_download_q(nuc3d_case *c, int b, double *q5);
/* RHS of the last nuc3d_rhs call (5N), PDEData3d::getRHS. */
int nuc3d_download_rhs(nuc3d_case *c, int b, double *rhs5);

/* Restart files of the reference (block::outputQ_binary / inputQ_binary,
 * block.cpp:754-828): one float64 `time`, then the five fields of Q_work
 * (conserved / J), i-fastest, nothing else.  The reference names them
 * flowData/NUC3d_step_<step>_ID_<rank>.bin.  nuc3d_save_q writes block b's
 * current state; nuc3d_load_q reads one (its size must match the block) and
 * uploads it like nuc3d_upload_q.  Both run between full steps. */
int nuc3d_save_q(nuc3d_case *c, int b, const char *path, double time);
int nuc3d_load_q(nuc3d_case *c, int b, const char *path, double *time);

/* Tecplot ASCII files of the reference (block::outputQ_tecplot,
 * block.cpp:692-752, named flowData/NUC3d_step_<step>_id_<rank>.dat; and
 * block::outputGEO_tecplot, block.cpp:857-922, flowData/GEO__id_<rank>.dat):
 * "title=NUC3d", the variables x, y, z, Val_0..Val_{nv-1}, one zone of
 * (nx+1)(ny+1)(nz+1) nodes in BLOCK packing with the values cell-centred,
 * every value on its own line as %.12g (the reference's setprecision(12)).
 * xn, yn, zn are block b's node coordinates ((nx+1)(ny+1)(nz+1) each,
 * i-fastest: the mesh file's, block.cpp:42-98).  nuc3d_write_tecplot writes
 * the flow (nv = 8: rho, u, v, w, p, T, e, c of physicsModel::getPrim on the
 * current state, between full steps); nuc3d_write_tecplot_geo the nine
 * metric terms xi_x .. zeta_z (nv = 9). */
int nuc3d_write_tecplot(nuc3d_case *c, int b, const char *path, const double *xn, const double *yn, const double *zn);
int nuc3d_write_tecplot_geo(nuc3d_case *c, int b, const char *path, const double *xn, const double *yn,
                            const double *zn);

/* One right-hand-side evaluation of every block on the current state
 * (EulerData3D::solve / NaiverStokesData3d::solve, including the halo
 * exchange and boundary conditions). */
int nuc3d_rhs(nuc3d_case *c, void *stream);

/* RHS evaluation plus PDEData3d::solve for TVD-RK3 stage 0, 1 or 2; stages
 * must follow each other in that order.  Stage 0 takes dt = cfl * min over
 * blocks and directions of 1/max eigenvalue of the stage-0 state. */
int nuc3d_stage(nuc3d_case *c, int stage, double cfl, void *stream);

/* nsteps full TVD-RK3 steps (block::solve, block.cpp:569-594). */
int nuc3d_steps(nuc3d_case *c, int nsteps, double cfl, void *stream);

/* nsteps full TVD-RK3 steps, like nuc3d_steps (the stage loop of
 * block::solve, block.cpp:569-594), launched as a captured HIP graph of two
 * steps (all kernels of six stages, no host work between them); an odd last
 * step runs as ordinary launches.  The graph is captured on first use and
 * again when cfl changes; it runs on `stream` in order with the caller's other
 * work.  Results are bit-identical to nuc3d_steps.  With kernel timing on
 * (nuc3d_timing) the steps run as ordinary launches. */
int nuc3d_step_graph(nuc3d_case *c, int nsteps, double cfl, void *stream);

/* Diagnostics of postproc::solveQ (postproc.cpp:743-848) on the current state,
 * the values block::Post writes to enstrophy.dat: enstrophy = sum over the
 * cells of all blocks of |vorticity|^2/2 and kinetic = sum of rho|u|^2/2, with
 * velocity gradients by the case's derivative schemes and closures.  For
 * block b, fields5 (5N, may be NULL) receives the vorticity (3 fields), the
 * enstrophy density and the Q-criterion.  Sums are taken per workgroup and
 * added in a fixed order, so they are reproducible; their rounding differs
 * from the reference's cell-by-cell sum by a few ulps per cell added. */
int nuc3d_diagnostics(nuc3d_case *c, double *enstrophy, double *kinetic, int b, double *fields5, void *stream);
/* nuc3d_diagnostics without allocating: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The enstrophy and
 * kinetic-energy sums only (no fields), equal bit for bit to those of
 * nuc3d_diagnostics; the per-workgroup partials and the sums go to a fixed
 * device array of the library (16 MiB per GPU, held until the process ends)
 * instead of buffers allocated per call.  Limit: 2^20 workgroups, one per
 * 256 cells per column-tile plane (about 268 M cells at the default column
 * tiles); larger cases get NUC3D_ERR_UNSUPPORTED.  The case's device must be
 * current.  Calls on one GPU share the array: calls that overlap on one GPU
 * (other threads, other streams) race and must be serialised by the caller. */
int nuc3d_diagnostics_noalloc(nuc3d_case *c, double *enstrophy, double *kinetic, void *stream);

/* The host time loop of singleBlock::solvePDE (singleBlock.cpp:51-71) with
 * the IOController cadence (IOcontroller.cpp:181-212):
 *   while (time <= tl && step < end_step):
 *       one TVD-RK3 step (block::solve, block.cpp:569-594): step += 1,
 *       time += dt, residual as nuc3d_get_residual;
 *       verbose: the status line of block::printStatus (block.cpp:596-599);
 *       post_step > 0 and step % post_step == 0: nuc3d_diagnostics, appending
 *           "step time enstrophy kinetic" (12 significant digits) to
 *           <dir>/enstrophy.dat, as postproc::solveQ does (postproc.cpp:833-842);
 *       save_step > 0 and step % save_step == 0: a restart file per block,
 *           <dir>/flowData/NUC3d_step_<step>_ID_<block>.bin (block.cpp:754-778).
 * dir NULL writes no files.  `out` (may be NULL) receives the final step,
 * time, dt and residual. */
typedef struct {
    int start_step, end_step, save_step, post_step; /* startStep endStep saveStep postStep */
    double cfl, tl, time;                           /* cfl, tl, and the time at start_step */
    const char *dir;
    int verbose;
} nuc3d_run_config;
typedef struct {
    int step;
    double time, dt, residual;
} nuc3d_run_status;
int nuc3d_run(nuc3d_case *c, const nuc3d_run_config *cfg, nuc3d_run_status *out, void *stream);

/* Batches of independent cases (no counterpart in the reference, where each
 * case is its own MPI job).  Block b belongs to group `group` (0 by default):
 * each group is one case, with dt = cfl * min over its own blocks
 * (PDEData3d.cpp:74-80 per case), and its own Mach and Reynolds numbers
 * (mach, reynolds <= 0 keep the case's).  Blocks of different groups must not
 * be connected: the first call that launches work after faces or groups
 * changed returns NUC3D_ERR_BAD_ARG if they are (so connected blocks can be
 * moved to another group one after the other).  All groups advance together
 * in the same launches.  Set between full steps; nuc3d_run needs a single
 * group. */
int nuc3d_set_block_group(nuc3d_case *c, int b, int group, double mach, double reynolds);
/* dt of group `group` in the current step (nuc3d_get_dt gives group 0's). */
int nuc3d_get_dt_group(nuc3d_case *c, int group, double *dt);

/* Autotuner (no counterpart in the reference).  The column-tile passes
 * (viscous fluxes with u,v,w,T, viscous divergence, diagnostics) run over
 * 16 x 16 column tiles that march along k in segments of kseg planes, 32 by
 * default.  For Navier-Stokes cases this takes the distinct block shapes in
 * turn and, for kseg = 4, 8, 16, 32, 64, 128 (up to twice the block's depth) on
 * every block of that shape, the other blocks keeping their current choice,
 * times the column-tile passes of one stage (the viscous fluxes, and the viscous
 * divergence where CRWENO5 z lines leave it to its own pass) over the whole
 * case on the current state;
 * it keeps the fastest per shape.  kseg_out (nblocks, may be NULL) receives the choices.  Results do
 * not change, except the rounding of the diagnostics' partial sums; the
 * residual history restarts (the next nuc3d_get_residual after one step is 0).
 * Runs between full steps. */
int nuc3d_autotune(nuc3d_case *c, void *stream, int *kseg_out);
/* The autotuner also picks, per block shape and direction, the tile shape of
 * the WENO5-JS / WENO5-Z direction sweeps and of the learned hybrid's x and y
 * sweeps: 16 faces x 16 lines per workgroup
 * (the default) or 32 faces x 8 lines (fewer halo cells staged per face, half
 * as many lines side by side), timing each shape's candidates on the
 * direction's own launches in the context of the whole case (the x and y
 * sweeps as in an RHS evaluation, the z sweep as in stage 0; the last
 * nuc3d_rhs result is overwritten).  Results do not change.  The choice of
 * block b (faces per tile line for x, y, z) is read with
 * nuc3d_get_sweep_tiles and can be set with nuc3d_set_sweep_tiles (32 only for
 * WENO5-JS and WENO5-Z directions and the hybrid's x and y; between full
 * steps; the residual history restarts). */
int nuc3d_get_sweep_tiles(nuc3d_case *c, int b, int *nf3);
int nuc3d_set_sweep_tiles(nuc3d_case *c, int b, const int *nf3);

/* Wait for the case's work on its streams.  Like nuc3d_get_dt,
 * nuc3d_get_dt_group and nuc3d_run, it returns NUC3D_ERR_NEG_DENSITY or
 * NUC3D_ERR_NEG_PRESSURE once a state failed the positivity check (below);
 * the status holds until a new state is uploaded or loaded. */
int nuc3d_sync(nuc3d_case *c);

/* Positivity of the states advanced so far.  Every state written by an RK
 * stage (and every uploaded state, at the first call that uses it) takes the
 * e<0 fix-up of con2prim (physicsModel.cpp:322-332) and then the check that
 * the reference's con2prim makes next: pressure (E - rho|u|^2/2)(gamma - 1)
 * not negative (EoSIdealGasgetPressure, physicsModel.cpp:1009-1016, where the
 * reference prints "negative pressure!!!" and exits).  Density must also be
 * positive; a NaN fails both.  counts2 receives the cells that failed
 * [density, pressure] since the state was last uploaded or loaded.  The
 * stages go on computing (nothing exits); the synchronising calls above
 * report the failure as a status. */
int nuc3d_get_positivity(nuc3d_case *c, long long *counts2);

/* dt of the current step (PDEData3d::getDt) and per-block max eigenvalues
 * of the last stage-0 state, or of the state of the last nuc3d_rhs call
 * (EulerFlux::maxEigen, physicsModel.cpp:756-758). */
int nuc3d_get_dt(nuc3d_case *c, double *dt_global);
int nuc3d_get_max_eigen(nuc3d_case *c, int b, double *eig3);
/* Residual of the last full step, PDEData3d::setRES (PDEData3d.cpp:90-122):
 * per block res = |sqrt(sum (Q - Q_n)^2 / 5N) - the same of the step before|
 * / dt^2 over the new state Q and the step's start Q_n, and res_global = the
 * maximum over blocks (the reference's MPI_MAX over ranks, one block per
 * rank).  res_blocks (nblocks, may be NULL) receives the per-block values.
 * 0 before the first step. */
int nuc3d_get_residual(nuc3d_case *c, double *res_global, double *res_blocks);
/* Number of e<0 energy fix-ups applied so far (physicsModel.cpp:324-332). */
int nuc3d_get_fixups(nuc3d_case *c, long long *count);

/* Kernel timing (no counterpart in the reference, which times whole steps
 * with MPI_Wtime in block::solve, block.cpp:576-584).  While enabled, every
 * kernel launch is bracketed by HIP events on its stream, in NUC3D_NTIMING
 * classes: 0 x sweep, 1 y sweep, 2 z sweep (with the RK update), and for
 * Navier-Stokes 4 viscous fluxes (u,v,w,T computed in the same pass), 5
 * viscous divergence (single RHS evaluations, and the stages of cases with
 * CRWENO5 z lines; the other stages take it in the z sweep, class 2);
 * class 3 (a separate primitives pass) is no longer launched.
 * nuc3d_timing_read returns the summed
 * milliseconds and launch counts per class since the last read (arrays of
 * NUC3D_NTIMING) and resets them. */
#define NUC3D_NTIMING 6
int nuc3d_timing(nuc3d_case *c, int enable);
int nuc3d_timing_read(nuc3d_case *c, double *ms, int *launches);

/* ---- Learned sensor of the hybrid reconstruction (no counterpart in the
 * reference) ----
 * A multilayer perceptron 30 -> 32 -> 32 -> 1 (ReLU) with BF16 weights and
 * activations and FP32 accumulation on the matrix cores.  Its input at face m
 * are rho, u, v, w, p on the six cells m-3..m+2, each centred on its stencil
 * mean and divided by (stencil range + 0.01 S), S = mean rho for rho, mean p
 * for p and the sound speed of the means for the velocities.  A face takes
 * WENO5-JS when the logit reaches the threshold, when the stencil's mean rho
 * or p is not positive, or when the central face value of rho or p is not
 * positive (the positivity fallback); otherwise the central value.
 * Parameters are FP32, NUC3D_SENSOR_NPARAM of them, in the order
 *   W1[32][32] (hidden x input; inputs 30 and 31 are unused and kept zero),
 *   b1[32], W2[32][32], b2[32], w3[32], b3.
 * The training label of a face (the teacher) is 1 when, for some primitive,
 * the central face value and the WENO5-JS value of either upwind side differ
 * by more than tau S.  Training runs on the GPU: minibatch Adam with the
 * loss wt (softplus(z) - y z), wt = pos_weight for y = 1, averaged over the
 * batch.  A sensor handle owns the parameters, the optimiser state and a
 * device-resident dataset of (30 features, label) samples. */
#define NUC3D_SENSOR_NPARAM 2145
typedef struct nuc3d_sensor nuc3d_sensor;
int nuc3d_sensor_create(long long capacity, nuc3d_sensor **out);
int nuc3d_sensor_destroy(nuc3d_sensor *s);
/* parameters in and out (set_params also resets the optimiser state) */
int nuc3d_sensor_set_params(nuc3d_sensor *s, const float *params);
int nuc3d_sensor_get_params(nuc3d_sensor *s, float *params);
/* Append samples of every block and direction of the case's current state
 * (between full steps): the faces whose index L in the face array of a
 * direction ((n_d+1) x n_{d+1} x n_{d+2}, face index fastest, blocks and
 * directions in order) has L % stride == offset.  A face whose eight-cell
 * stencil m-4..m+3 meets a physical face or leaves the ghost frame is stored
 * with label -1 (not trained on).  *added (may be NULL) receives the count. */
int nuc3d_sensor_sample(nuc3d_sensor *s, nuc3d_case *c, double tau, int stride, int offset, long long *added);
int nuc3d_sensor_count(nuc3d_sensor *s, long long *count, long long *positives);
int nuc3d_sensor_clear(nuc3d_sensor *s);
/* dataset rows: 32 bf16 bit patterns (features 0..29, then two zeros) and a label */
int nuc3d_sensor_upload(nuc3d_sensor *s, const unsigned short *x32, const float *y, long long n);
int nuc3d_sensor_download(nuc3d_sensor *s, long long first, long long n, unsigned short *x32, float *y);
/* Gradient of the mean loss over samples [first, first+n) (NUC3D_SENSOR_NPARAM
 * values) and the loss, by the training kernels. */
int nuc3d_sensor_grad(nuc3d_sensor *s, long long first, long long n, float pos_weight, float *grad, double *loss);
/* logits of samples [first, first+n) by the network as the hybrid sweep runs it */
int nuc3d_sensor_infer(nuc3d_sensor *s, long long first, long long n, float *logits);
/* `steps` Adam steps (lr, beta 0.9 / 0.999) on minibatches of `batch` samples
 * drawn from the dataset by a hash of (seed, step, position); losses (steps
 * values, may be NULL) receives the loss of every step. */
int nuc3d_sensor_train(nuc3d_sensor *s, int steps, int batch, float lr, float pos_weight, unsigned long long seed,
                       float *losses);
/* The network the case's hybrid directions use (NUC3D_SENSOR_NPARAM FP32
 * parameters, rounded to BF16 as in training) and the logit threshold. */
int nuc3d_set_sensor(nuc3d_case *c, const float *params, float threshold);
/* The network's decisions are made by nuc3d_rhs and by stage 0 of a step
 * (from the stage-0 state); the other stages reuse them (see
 * nuc3d_set_sensor_interval).  The positivity fallback is evaluated again at
 * every stage on that stage's state: a reused central face whose stencil
 * means or central face values of rho or p are not positive takes WENO5-JS
 * (and is recorded as 2 from then on).  A single RHS evaluation keeps its
 * decisions apart: it never changes the decisions of the step in progress.
 * The decisions of the last evaluation (nuc3d_rhs, or a stage) of block b
 * and hybrid direction d, in the layout of the face array ((n_d+1) x n_{d+1}
 * x n_{d+2}, face index fastest): 0 central, 1 WENO5-JS by the network, 2
 * WENO5-JS by the fallback, 3 physical-face closure. */
int nuc3d_download_hybrid_flags(nuc3d_case *c, int b, int d, unsigned char *flags);
/* Overlap (no counterpart in the reference): the viscous-flux pass of a
 * Navier-Stokes stage reads only the stage's input state and writes only the
 * viscous fluxes, so it runs on a second stream of the case beside the x and
 * y sweeps, and the z sweep waits for it (also inside captured step graphs).
 * On by default; off runs every pass in order on the caller's stream.  Results
 * are identical either way; with kernel timing on, the passes run in order. */
int nuc3d_set_overlap(nuc3d_case *c, int enable);
/* Persistent direction sweeps (no counterpart in the reference, whose ranks
 * loop over their own blocks: block.cpp:576-581).  When enabled, the x and
 * y sweeps (and the z sweep of Euler cases) of WENO5-JS / WENO5-Z directions
 * are each one persistent grid of as many workgroups as the GPU holds at
 * once, whose workgroups take (block, tile) work items of every block of the
 * case from a device queue until none is left (one queue per XCD over its
 * contiguous range of the tile list, then from the others); the other
 * kernels, and every kernel when off (the default), have one workgroup per
 * tile and the hardware dispatcher assigns them.  Results are identical
 * either way. */
int nuc3d_set_persistent(nuc3d_case *c, int enable);
/* Uniform-mesh path (no counterpart in the reference, whose blocks always
 * hold J and the nine metric fields).  When enabled (the default), a block
 * whose J and metric terms are each one value, bit for bit, over its interior
 * and the ghost slabs of its neighbour faces (a uniform Cartesian block) has
 * the stage kernels take them from its descriptor instead of loading the
 * fields; results are identical, the HBM traffic per cell lower.  The check
 * runs whenever the case's metrics, faces or state were replaced.
 * nuc3d_get_uniform_mesh reports whether block b takes that path. */
int nuc3d_set_uniform_mesh(nuc3d_case *c, int enable);
int nuc3d_get_uniform_mesh(nuc3d_case *c, int b, int *uniform);
/* Steps between the network's decisions: 1 (default, stage 0 of every step)
 * or 2 (stage 0 of every other step; the stages between reuse them).  Either
 * way the first stage 0 after a new state (upload, load, faces, metrics), a
 * new sensor or a new interval decides afresh. */
int nuc3d_set_sensor_interval(nuc3d_case *c, int steps);
/* Faces decided so far (counted where decisions are made): central, WENO5-JS
 * by the network, WENO5-JS by the fallback. */
int nuc3d_get_hybrid_counts(nuc3d_case *c, long long *counts3);

/* 1 when this build implements the CRWENO5 line solves (no counterpart in the
 * reference: a capability query for callers). */
int nuc3d_has_crweno5(void);

/* Self-test of the arithmetic (no counterpart in the reference).  The exact
 * kernels divide by a shortened form of the compiler's IEEE FP64 division
 * (v_rcp_f64 and two Newton steps, quotient with one FMA correction, without
 * the scaling and special-value steps), used only where every operand is in
 * the range in which both give the same correctly rounded quotient.  This
 * compares the two on n hashed operand pairs of that range (x7 quotients:
 * one reciprocal serves three numerators and a signed zero, whose quotient
 * keeps the IEEE sign, and the sign-keeping correction serves two numerators
 * and a signed zero over |b|), and the square root of |a / b| by the same
 * shortening of the compiler's FP64 square root, on the current device, and
 * returns the number of results whose bits differ (0 expected). */
int nuc3d_selftest_division(long long n, unsigned long long seed, long long *mismatches);
/* ieee = 1: the case's kernels use the compiler's IEEE division everywhere
 * instead (results are the same bit for bit; for tests). */
int nuc3d_set_division(nuc3d_case *c, int ieee);

/* Description of the calling thread's last error. */
const char *nuc3d_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
