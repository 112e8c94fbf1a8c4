// nuc3d_capi.cpp -- host side of the C ABI declared in include/nuc3d_hip.h.
//
// Owns the device memory of a case (all blocks on the current GPU), the block
// descriptors and tile tables, the three state buffers of the TVD-RK3 cycle
// and the small control arrays (dt, eigenvalue maxima, counters), and
// enqueues the kernels of nuc3d_kernels.hip on the caller's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <array>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/nuc3d_hip.h"
#include "../../include/nuc3d_ops.h"
#include "nuc3d_types.h"

// Navier-Stokes z sweeps: i-tiles per group of the work-item order (the tiles
// of a group that share a run of j rows and a k segment are neighbours in the
// list, so the i halo of the viscous divergence is read from one XCD's L2;
// 4 measures 1.5% faster per z sweep on 256^3 than 1, and 1% than 16)
#ifndef NUC3D_ZRUN_IGROUP
#define NUC3D_ZRUN_IGROUP 4
#endif
#ifndef NUC3D_ZRUN_J
#define NUC3D_ZRUN_J 8  // j rows per run
#endif

#define NUC3D_LAUNCHERS(SFX)                                                                                         \
    extern "C" hipError_t nuc3d_launch_sweep_##SFX(int, int, int, int, int, const nuc3d_block_dev *, const int4 *,   \
                                                   int, nuc3d_kparams, int, int, int, int, double *,                 \
                                                   const unsigned long long *, unsigned long long *,                 \
                                                   unsigned long long *, hipStream_t);                               \
    extern "C" hipError_t nuc3d_launch_diag_##SFX(const nuc3d_block_dev *, const nuc3d_tile *, int, int,            \
                                                  nuc3d_kparams, int, int, double *, double *, hipStream_t);         \
    extern "C" hipError_t nuc3d_launch_prim_##SFX(const nuc3d_block_dev *, const nuc3d_tile *, int, int,            \
                                                  nuc3d_kparams, int, hipStream_t);                                  \
    extern "C" hipError_t nuc3d_launch_visc_flux_##SFX(const nuc3d_block_dev *, const nuc3d_tile *, int,           \
                                                       nuc3d_kparams, int, hipStream_t);                             \
    extern "C" hipError_t nuc3d_launch_visc_div_##SFX(int, const nuc3d_block_dev *, const nuc3d_tile *, int,        \
                                                      nuc3d_kparams, int, int, int, int, double *,                   \
                                                      const unsigned long long *, unsigned long long *,              \
                                                      unsigned long long *, hipStream_t);                            \
    extern "C" hipError_t nuc3d_launch_eigen_##SFX(const nuc3d_block_dev *, const nuc3d_tile *, int, nuc3d_kparams,  \
                                                   int, unsigned long long *, hipStream_t);                          \
    extern "C" hipError_t nuc3d_launch_fixup_##SFX(const nuc3d_block_dev *, const nuc3d_tile *, int, nuc3d_kparams,  \
                                                   int, unsigned long long *, hipStream_t);                          \
    extern "C" hipError_t nuc3d_launch_ghost_fill_##SFX(const nuc3d_block_dev *, int, int, int, hipStream_t);     \
    extern "C" hipError_t nuc3d_launch_crweno_##SFX(int, int, int, const nuc3d_block_dev *, const int4 *, int, int, \
                                                    nuc3d_kparams, int, int, int, int, double *,                     \
                                                    const unsigned long long *, unsigned long long *,                \
                                                    unsigned long long *, double *, hipStream_t);
NUC3D_LAUNCHERS(fast)
NUC3D_LAUNCHERS(exact)
extern "C" hipError_t nuc3d_launch_sensor_sample(const nuc3d_block_dev *, int, int, int, double, double, long long,
                                                 long long, long long, unsigned *, float *, unsigned long long *,
                                                 hipStream_t);
extern "C" hipError_t nuc3d_launch_sensor_fwdbwd(const unsigned *, const float *, long long, long long, long long, int,
                                                 unsigned long long, long long, const float *, float, float, int,
                                                 float *, float *, float *, hipStream_t);
extern "C" hipError_t nuc3d_launch_sensor_adam(float *, float *, float *, const float *, float, int, float *,
                                               hipStream_t);
extern "C" hipError_t nuc3d_launch_mesh_metrics(nuc3d_block_dev, const double *, const double *, const double *,
                                                double *, int *, hipStream_t);
extern "C" hipError_t nuc3d_launch_uniform_check(nuc3d_block_dev, int *, hipStream_t);
extern "C" hipError_t nuc3d_launch_block_consts(double, const double *, double *, hipStream_t);
extern "C" hipError_t nuc3d_launch_sum_pairs(const double *, long long, double *, hipStream_t);
extern "C" hipError_t nuc3d_launch_selftest_division(long long, unsigned long long, unsigned long long *,
                                                     hipStream_t);

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

int hipfail(hipError_t e, const char *what)
{
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return NUC3D_ERR_HIP;
}

#define HIPCHK(x)                                      \
    do {                                               \
        hipError_t e_ = (x);                           \
        if (e_ != hipSuccess) return hipfail(e_, #x);  \
    } while (0)

struct Launchers {
    decltype(&nuc3d_launch_sweep_fast) sweep;
    decltype(&nuc3d_launch_prim_fast) prim;
    decltype(&nuc3d_launch_diag_fast) diag;
    decltype(&nuc3d_launch_visc_flux_fast) vflux;
    decltype(&nuc3d_launch_visc_div_fast) vdiv;
    decltype(&nuc3d_launch_eigen_fast) eigen;
    decltype(&nuc3d_launch_fixup_fast) fixup;
    decltype(&nuc3d_launch_ghost_fill_fast) ghost;
    decltype(&nuc3d_launch_crweno_fast) crweno;
};

const Launchers kFast = {nuc3d_launch_sweep_fast, nuc3d_launch_prim_fast, nuc3d_launch_diag_fast, nuc3d_launch_visc_flux_fast, nuc3d_launch_visc_div_fast, nuc3d_launch_eigen_fast,
                         nuc3d_launch_fixup_fast, nuc3d_launch_ghost_fill_fast, nuc3d_launch_crweno_fast};
const Launchers kExact = {nuc3d_launch_sweep_exact, nuc3d_launch_prim_exact, nuc3d_launch_diag_exact, nuc3d_launch_visc_flux_exact, nuc3d_launch_visc_div_exact, nuc3d_launch_eigen_exact,
                          nuc3d_launch_fixup_exact, nuc3d_launch_ghost_fill_exact, nuc3d_launch_crweno_exact};

// temporary device buffers of one call, freed on return
struct OpBuffers {
    std::vector<void *> p;
    ~OpBuffers()
    {
        for (void *q : p) (void)hipFree(q);
    }
    double *get(size_t n)
    {
        void *q = nullptr;
        if (hipMalloc(&q, sizeof(double) * std::max<size_t>(n, 1)) != hipSuccess) return nullptr;
        p.push_back(q);
        return static_cast<double *>(q);
    }
};

constexpr int KSEG = NUC3D_KSEG;  // planes per k segment of a tile

}  // namespace

struct nuc3d_case {
    nuc3d_params p;
    nuc3d_kparams K;
    const Launchers *L;
    int nb = 0;
    int G = 3;
    int device = 0;
    std::vector<nuc3d_block_dev> hb;   // host copy of the descriptors
    std::vector<void *> chunks;        // one device allocation per block
    std::vector<long long> padded;     // padded cells per block
    nuc3d_block_dev *db = nullptr;     // device descriptors
    nuc3d_tile *dtiles = nullptr;      // column tiles (NUC3D_TILE_X x NUC3D_TILE_Y) with k segments (pointwise kernels)
    int ntiles = 0;
    std::vector<int> kseg;             // k-segment length of the column tiles per block (nuc3d_autotune)
    int kmax = NUC3D_KSEG;             // the longest
    // direction-sweep tiles per direction, in two groups: 16 faces x 16
    // lines, and 32 faces x 8 lines (the tile shape of each block and
    // direction, swnf, chosen by nuc3d_autotune)
    int4 *dsweep[3][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
    int nsweep[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    std::vector<std::array<int, 3>> swnf;
    int4 *dlines[3] = {nullptr, nullptr, nullptr};  // whole lines (CRWENO5 directions)
    int nlines[3] = {0, 0, 0};
    int nmax[3] = {0, 0, 0};
    double *dscratch = nullptr;                     // Thomas c', d' of the CRWENO5 lines
    double *ddt = nullptr;             // dt of the current step
    unsigned long long *dlam = nullptr;   // [2][nb][3] eigenvalue maxima (bit patterns)
    unsigned long long *dcount = nullptr; // [4] fix-ups, non-positive density, negative pressure
    unsigned long long *dhcount = nullptr; // hybrid decision counts, NUC3D_HCOUNT_SLOTS x 16
    int cur = 0;                       // state buffer holding Q_work
    int next_stage = 0;
    int lampar = 0;
    bool dirty = true;                 // descriptors / ghosts / eigenvalues need refreshing
    bool faces_set = false;
    // residual partials of the last two steps (PDEData3d::setRES): one per
    // work item of the kernel that finishes stage 2, and that item's block
    double *dresid = nullptr;
    int nwork = 0;
    std::vector<int> work_block;
    long long steps_done = 0;
    // the viscous-flux pass of a Navier-Stokes stage runs on a second stream
    // beside the x and y sweeps (it reads only the stage's input state and
    // writes only fv; the z sweep waits for it)
    bool overlap = true;
    // persistent direction sweeps over the device work queue (nuc3d_set_persistent)
    bool persistent = false;
    unsigned *dqueue = nullptr;        // NUC3D_QUEUE_WORDS counters, zero between launches
    // range flags of the three state buffers (K.qok when the case's gamma and
    // Mach numbers allow the shared reciprocal of rho), zero after a refresh
    unsigned *dqok = nullptr;
    hipStream_t vstream = nullptr;
    hipEvent_t vfork = nullptr, vjoin = nullptr;
    // captured two-step graphs, one per parity of the eigenvalue buffers
    hipStream_t gstream = nullptr;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    double gcfl[2] = {-1.0, -1.0};
    // hybrid reconstruction: the network (inference copy) and the recorded decisions
    nuc3d_sensor_dev *dsensor = nullptr;
    std::vector<void *> dflags;
    int hybrid_interval = 1;           // steps between hybrid decisions (1 or 2)
    // the next stage 0 decides whatever the interval says: the state, the
    // faces or the sensor changed since the stored decisions were made
    bool hforce = true;
    bool flags_rhs = false;            // the last evaluation was a single RHS (its decisions: second half)
    // uniform blocks take J and the metric terms from their descriptor
    bool uniform_mesh = true;
    int *duniform = nullptr;           // [nb] uniform-mesh check flags
    double *dconsts = nullptr;         // [5] constants of one uniform block (k_block_consts)
    std::vector<char> rdok;            // per block: J and metrics allow the split's shared reciprocal
    // kernel timing
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    // x sweep, y sweep, z sweep, viscous fluxes, viscous divergence (+ update)
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[NUC3D_NTIMING];
};

namespace {

// After a synchronisation: the positivity status of the states advanced
// since the last upload (nuc3d_get_positivity).
int positivity_status(nuc3d_case *c)
{
    unsigned long long v[3] = {0, 0, 0};
    HIPCHK(hipMemcpy(v, c->dcount, sizeof v, hipMemcpyDeviceToHost));
    if (v[1])
        return fail(NUC3D_ERR_NEG_DENSITY,
                    std::to_string(v[1]) + " cells with non-positive or NaN density after an RK stage");
    if (v[2])
        return fail(NUC3D_ERR_NEG_PRESSURE, std::to_string(v[2]) +
                                                " cells with negative or NaN pressure after the e<0 fix-up "
                                                "(physicsModel.cpp:1009-1016)");
    return NUC3D_OK;
}

// a new state: its positivity is checked afresh
int clear_positivity(nuc3d_case *c)
{
    HIPCHK(hipMemset(c->dcount + 1, 0, 2 * sizeof(unsigned long long)));
    return NUC3D_OK;
}

int check_block(nuc3d_case *c, int b)
{
    if (!c) return fail(NUC3D_ERR_BAD_ARG, "null case");
    if (b < 0 || b >= c->nb) return fail(NUC3D_ERR_BAD_ARG, "block index out of range");
    return NUC3D_OK;
}

// Padded field layout: a row holds XLEAD slots before cell i=0 (the last G of
// them are the ghost cells), the n0 interior cells, G ghosts and padding to a
// multiple of ROWALIGN doubles.  With 256-byte aligned allocations every
// interior row therefore starts on a 128-byte L2 line, so a 16-cell row
// segment of a tile is one line, not two.
constexpr int XLEAD = 16, ROWALIGN = 16;

long long origin(const nuc3d_block_dev &B) { return XLEAD + B.sy * B.G + B.sz * B.G; }

// copy an interior field (nz,ny,nx) into / out of the padded device layout
// through a host staging copy of the whole padded box (ghosts are rebuilt on
// the device by refresh()).
int put_field(nuc3d_case *c, int b, double *dst, const double *src)
{
    const nuc3d_block_dev &B = c->hb[b];
    const int G = B.G, nx = B.n[0], ny = B.n[1], nz = B.n[2];
    std::vector<double> st((size_t)c->padded[b], 0.0);
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            std::memcpy(&st[(size_t)(XLEAD + B.sy * (j + G) + B.sz * (k + G))], src + ((size_t)k * ny + j) * nx,
                        sizeof(double) * nx);
    HIPCHK(hipMemcpy(dst - origin(B), st.data(), sizeof(double) * st.size(), hipMemcpyHostToDevice));
    return NUC3D_OK;
}

int get_field(nuc3d_case *c, int b, double *dst, const double *src)
{
    const nuc3d_block_dev &B = c->hb[b];
    const int G = B.G, nx = B.n[0], ny = B.n[1], nz = B.n[2];
    std::vector<double> st((size_t)c->padded[b]);
    HIPCHK(hipMemcpy(st.data(), src - origin(B), sizeof(double) * st.size(), hipMemcpyDeviceToHost));
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            std::memcpy(dst + ((size_t)k * ny + j) * nx, &st[(size_t)(XLEAD + B.sy * (j + G) + B.sz * (k + G))],
                        sizeof(double) * nx);
    return NUC3D_OK;
}

hipEvent_t take_event(nuc3d_case *c)
{
    if (!c->ev_pool.empty()) {
        hipEvent_t e = c->ev_pool.back();
        c->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// brackets one launch with events when timing is enabled
struct Timed {
    nuc3d_case *c;
    hipStream_t s;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> *dst;
    hipEvent_t a = nullptr;
    Timed(nuc3d_case *c_, hipStream_t s_, std::vector<std::pair<hipEvent_t, hipEvent_t>> *d) : c(c_), s(s_), dst(d)
    {
        if (c->timing && (a = take_event(c))) (void)hipEventRecord(a, s);
    }
    ~Timed()
    {
        if (!a) return;
        hipEvent_t b = take_event(c);
        if (!b) return;
        (void)hipEventRecord(b, s);
        dst->push_back({a, b});
    }
};

int refresh(nuc3d_case *c, hipStream_t s)
{
    if (!c->dirty) return NUC3D_OK;
    // a case (group) is the reference's whole MPI job: its blocks may only be
    // connected to blocks of the same group (checked here, once the faces and
    // groups are all set, so that connected blocks can change group together)
    for (int b = 0; b < c->nb; b++)
        for (int f = 0; f < 6; f++) {
            const int t = c->hb[b].ftype[f];
            if (t >= 0 && c->hb[t].group != c->hb[b].group)
                return fail(NUC3D_ERR_BAD_ARG, "connected faces join blocks of different groups");
        }
    c->hforce = true;
    c->rdok.assign(c->nb, 0);
    for (auto &B : c->hb) B.uniform = B.rdsplit = 0;
    HIPCHK(hipMemcpy(c->db, c->hb.data(), sizeof(nuc3d_block_dev) * c->nb, hipMemcpyHostToDevice));
    // ghosts of J and metrics, fix-up and ghosts of the state, eigenvalues
    HIPCHK(c->L->ghost(c->db, c->nb, 1, c->cur, s));
    // blocks whose J and metric terms are one value each (uniform Cartesian
    // meshes): the stage kernels take them from the descriptor
    if (c->uniform_mesh) {
        HIPCHK(hipMemsetAsync(c->duniform, 0, sizeof(int) * c->nb, s));
        for (int b = 0; b < c->nb; b++) HIPCHK(nuc3d_launch_uniform_check(c->hb[b], c->duniform + b, s));
        std::vector<int> bad(c->nb);
        HIPCHK(hipStreamSynchronize(s));
        HIPCHK(hipMemcpy(bad.data(), c->duniform, sizeof(int) * c->nb, hipMemcpyDeviceToHost));
        for (int b = 0; b < c->nb; b++) {
            nuc3d_block_dev &B = c->hb[b];
            if (bad[b]) continue;
            HIPCHK(hipMemcpy(&B.jac0, B.jac, sizeof(double), hipMemcpyDeviceToHost));
            for (int m = 0; m < 9; m++) HIPCHK(hipMemcpy(&B.met0[m], B.met[m], sizeof(double), hipMemcpyDeviceToHost));
            // |grad xi_d| and the reciprocal of J as the device computes them
            double k5[5];
            HIPCHK(nuc3d_launch_block_consts(B.jac0, B.met0, c->dconsts, s));
            HIPCHK(hipMemcpyAsync(k5, c->dconsts, sizeof k5, hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            for (int d = 0; d < 3; d++) B.theta0[d] = k5[d];
            B.rjac0 = k5[3];
            c->rdok[b] = k5[4] != 0.0;
            B.rdsplit = c->rdok[b] && !c->K.ieee_div;
            B.uniform = 1;
        }
        HIPCHK(hipMemcpy(c->db, c->hb.data(), sizeof(nuc3d_block_dev) * c->nb, hipMemcpyHostToDevice));
    }
    int alluni = 1;
    for (const auto &B : c->hb) alluni = alluni && B.uniform;
    // the shared reciprocal of rho needs gamma in [1, 8] and every Mach number
    // in [2^-40, 2^40] (the temperature's numerator gamma Ma^2 p stays in the
    // reduced division's range); the flags restart at zero with every refresh
    bool qok_allowed = c->K.gamma >= 1.0 && c->K.gamma <= 8.0;
    for (const auto &B : c->hb) qok_allowed = qok_allowed && B.mach >= 0x1p-40 && B.mach <= 0x1p40;
    unsigned *qok = qok_allowed ? c->dqok : nullptr;
    HIPCHK(hipMemsetAsync(c->dqok, 0, sizeof(unsigned) * 4, s));
    if (alluni != c->K.alluni || qok != c->K.qok) {  // the captured steps hold the launch parameters by value
        c->K.alluni = alluni;
        c->K.qok = qok;
        for (auto &g : c->gexec)
            if (g) {
                (void)hipGraphExecDestroy(g);
                g = nullptr;
            }
        c->gcfl[0] = c->gcfl[1] = -1.0;
    }
    HIPCHK(c->L->fixup(c->db, c->dtiles, c->ntiles, c->K, c->cur, c->dcount, s));
    HIPCHK(c->L->ghost(c->db, c->nb, 0, c->cur, s));
    HIPCHK(hipMemsetAsync(c->dlam + 3 * c->nb * c->lampar, 0, sizeof(unsigned long long) * 3 * c->nb, s));
    HIPCHK(c->L->eigen(c->db, c->dtiles, c->ntiles, c->K, c->cur, c->dlam + 3 * c->nb * c->lampar, s));
    c->dirty = false;
    return NUC3D_OK;
}

// Navier-Stokes stages whose z sweep is a direction sweep (not CRWENO5 lines)
// finish their cells there: viscous divergence and RK stage in one pass.
bool visc_in_z_sweep(const nuc3d_case *c)
{
    return c->p.model == NUC3D_NAVIERSTOKES3D && c->p.scheme[2] != NUC3D_CRWENO5;
}

// Direction-sweep tile lists from the tile shape of every block and
// direction (c->swnf: 16 faces x 16 lines, or 32 faces x 8 lines), one list
// per shape, tiles of NF-1 cells along the sweep axis; where the z sweep
// finishes stage 2 its work items are the residual's (whose history restarts).
int build_sweeps(nuc3d_case *c)
{
    const bool zvisc = c->p.model == NUC3D_NAVIERSTOKES3D && c->p.scheme[2] != NUC3D_CRWENO5;
    std::vector<int4> sw[3][2];
    for (int b = 0; b < c->nb; b++) {
        const nuc3d_block_dev &B = c->hb[b];
        for (int d = 0; d < 3; d++) {
            const int nf = c->swnf[b][d], g = nf == 32 ? 1 : 0, nl = 256 / nf, ncell = nf - 1;
            std::vector<int4> &L = sw[d][g];
            if (d == 0) {
                for (int k = 0; k < B.n[2]; k++)
                    for (int l0 = 0; l0 < B.n[1]; l0 += nl)
                        for (int a0 = 0; a0 < B.n[0]; a0 += ncell) L.push_back(make_int4(b, a0, l0, k));
            } else if (d == 1) {
                for (int k = 0; k < B.n[2]; k++)
                    for (int l0 = 0; l0 < B.n[0]; l0 += nl)
                        for (int a0 = 0; a0 < B.n[1]; a0 += ncell) L.push_back(make_int4(b, a0, l0, k));
            } else if (zvisc) {
                // the z sweep also takes the viscous divergence, whose fv_eta
                // stencil spans j-3..j+3: runs of 8 rows in j at one k segment are
                // neighbours in the list (they run together and share one XCD's
                // L2), and so are the k segments of one run (their k halos overlap)
                // and the runs of a group of i-tiles (the fv_xi i halo)
                for (int g0 = 0; g0 < B.n[0]; g0 += NUC3D_ZRUN_IGROUP * nl)
                    for (int j0 = 0; j0 < B.n[1]; j0 += NUC3D_ZRUN_J)
                        for (int a0 = 0; a0 < B.n[2]; a0 += ncell)
                            for (int l0 = g0; l0 < std::min(g0 + NUC3D_ZRUN_IGROUP * nl, B.n[0]); l0 += nl)
                                for (int j = j0; j < std::min(j0 + NUC3D_ZRUN_J, B.n[1]); j++)
                                    L.push_back(make_int4(b, a0, l0, j));
            } else {
                for (int j = 0; j < B.n[1]; j++)
                    for (int l0 = 0; l0 < B.n[0]; l0 += nl)
                        for (int a0 = 0; a0 < B.n[2]; a0 += ncell) L.push_back(make_int4(b, a0, l0, j));
            }
        }
    }
    for (int d = 0; d < 3; d++)
        for (int g = 0; g < 2; g++) {
            (void)hipFree(c->dsweep[d][g]);
            c->dsweep[d][g] = nullptr;
            c->nsweep[d][g] = (int)sw[d][g].size();
            if (sw[d][g].empty()) continue;
            if (hipMalloc(&c->dsweep[d][g], sizeof(int4) * sw[d][g].size()) != hipSuccess)
                return fail(NUC3D_ERR_OOM, "device allocation of sweep tiles failed");
            HIPCHK(hipMemcpy(c->dsweep[d][g], sw[d][g].data(), sizeof(int4) * sw[d][g].size(), hipMemcpyHostToDevice));
        }
    if (c->p.scheme[2] != NUC3D_CRWENO5) {
        c->work_block.clear();
        for (int g = 0; g < 2; g++)
            for (const auto &t : sw[2][g]) c->work_block.push_back(t.x);
        c->nwork = (int)c->work_block.size();
        double *dr = nullptr;
        if (hipMalloc(&dr, sizeof(double) * 2 * c->nwork) != hipSuccess)
            return fail(NUC3D_ERR_OOM, "device allocation of residual partials failed");
        HIPCHK(hipMemset(dr, 0, sizeof(double) * 2 * c->nwork));
        (void)hipFree(c->dresid);
        c->dresid = dr;
    }
    for (auto &gx : c->gexec)
        if (gx) {
            (void)hipGraphExecDestroy(gx);
            gx = nullptr;
        }
    c->gcfl[0] = c->gcfl[1] = -1.0;
    return NUC3D_OK;
}

// viscous fluxes (Navier-Stokes), then the three sweeps; the z sweep
// completes the RHS (mode 0) or applies RK stage `stage` (mode 1); for
// Navier-Stokes the viscous divergence follows in its own pass, except in the
// stages of visc_in_z_sweep cases
// step: index of the step this stage belongs to (hybrid decisions are made at
// stage 0 of every hybrid_interval-th step, and by every single RHS)
int launch_rhs(nuc3d_case *c, int mode, int stage, int bin, int bun, int bout, const unsigned long long *lcur,
               unsigned long long *lnext, double *resid, hipStream_t s, long long step = 0)
{
    const bool ns = c->p.model == NUC3D_NAVIERSTOKES3D;
    nuc3d_kparams K = c->K;
    K.resid = (mode == 1 && stage == 2) ? resid : nullptr;
    K.hdecide = (mode == 0 || (stage == 0 && (step % c->hybrid_interval == 0 || c->hforce))) ? 1 : 0;
    K.hrhs = mode == 0 ? 1 : 0;
    for (int d = 0; d < 3; d++)
        if (c->p.scheme[d] == NUC3D_HYBRID && !K.sensor)
            return fail(NUC3D_ERR_BAD_ARG, "the hybrid scheme needs a sensor (nuc3d_set_sensor)");
    // u, v, w, T are computed inside the viscous-flux pass; with kernel timing
    // on, every pass runs alone on s
    const bool side = ns && c->overlap && !c->timing && c->vstream;
    if (side) {
        HIPCHK(hipEventRecord(c->vfork, s));
        HIPCHK(hipStreamWaitEvent(c->vstream, c->vfork, 0));
        HIPCHK(c->L->vflux(c->db, c->dtiles, c->ntiles, K, bin, c->vstream));
        HIPCHK(hipEventRecord(c->vjoin, c->vstream));
    } else if (ns) {
        Timed t(c, s, &c->ev[4]);
        HIPCHK(c->L->vflux(c->db, c->dtiles, c->ntiles, K, bin, s));
    }
    for (int d = 0; d < 3; d++) {
        if (d == 2 && side) HIPCHK(hipStreamWaitEvent(s, c->vjoin, 0));
        Timed t(c, s, &c->ev[d]);
        if (c->p.scheme[d] == NUC3D_CRWENO5) {
            HIPCHK(c->L->crweno(d, ns, mode, c->db, c->dlines[d], c->nlines[d], c->nmax[d], K, stage, bin, bun,
                                bout, c->ddt, lcur, lnext, c->dcount, c->dscratch, s));
            continue;
        }
        // one launch per tile shape; the residual partials of the z sweep's
        // work items follow one another in that order
        long long off = 0;
        for (int g = 0; g < 2; g++) {
            if (!c->nsweep[d][g]) continue;
            nuc3d_kparams Kg = K;
            if (Kg.resid) Kg.resid += off;
            if (c->persistent) {
                Kg.pq = c->dqueue;
                Kg.pq_n = c->nsweep[d][g];
            }
            HIPCHK(c->L->sweep(d, c->p.scheme[d], ns, mode, g ? 32 : 16, c->db, c->dsweep[d][g], c->nsweep[d][g], Kg,
                               stage, bin, bun, bout, c->ddt, lcur, lnext, c->dcount, s));
            off += c->nsweep[d][g];
        }
    }
    if (ns && !(mode == 1 && visc_in_z_sweep(c))) {
        Timed t(c, s, &c->ev[5]);
        HIPCHK(c->L->vdiv(mode, c->db, c->dtiles, c->ntiles, K, stage, bin, bun, bout, c->ddt, lcur, lnext,
                          c->dcount, s));
    }
    return NUC3D_OK;
}

}  // namespace

extern "C" {

const char *nuc3d_last_error(void) { return g_err.c_str(); }

int nuc3d_has_crweno5(void) { return 1; }

int nuc3d_selftest_division(long long n, unsigned long long seed, long long *mismatches)
{
    if (n < 0 || !mismatches) return fail(NUC3D_ERR_BAD_ARG, "bad argument");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(NUC3D_ERR_NO_DEVICE, "no HIP device");
    OpBuffers T;
    unsigned long long *d = reinterpret_cast<unsigned long long *>(T.get(1));
    if (!d) return fail(NUC3D_ERR_OOM, "device allocation failed");
    HIPCHK(hipMemset(d, 0, sizeof(unsigned long long)));
    HIPCHK(nuc3d_launch_selftest_division(n, seed, d, nullptr));
    unsigned long long v = 0;
    HIPCHK(hipMemcpy(&v, d, sizeof v, hipMemcpyDeviceToHost));
    *mismatches = (long long)v;
    return NUC3D_OK;
}

int nuc3d_case_create(const nuc3d_params *p, int nblocks, const int *dims, nuc3d_case **out)
{
    if (!p || !dims || !out) return fail(NUC3D_ERR_BAD_ARG, "null argument");
    *out = nullptr;
    if (nblocks < 1 || nblocks > NUC3D_MAX_BLOCKS) return fail(NUC3D_ERR_BAD_ARG, "nblocks out of range");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(NUC3D_ERR_NO_DEVICE, "no HIP device");
    int G = 3;
    for (int d = 0; d < 3; d++) {
        if (p->scheme[d] != NUC3D_WENO5JS && p->scheme[d] != NUC3D_WENO5Z && p->scheme[d] != NUC3D_CRWENO5 &&
            p->scheme[d] != NUC3D_CENTRAL8 && p->scheme[d] != NUC3D_HYBRID)
            return fail(NUC3D_ERR_UNSUPPORTED, "scheme not available in this build");
        // the central scheme (alone or in the hybrid) reaches four cells past the
        // face: four ghost layers and four closure faces at physical faces
        // (bufferWidth 4, as cd8)
        const bool wide = p->scheme[d] == NUC3D_CENTRAL8 || p->scheme[d] == NUC3D_HYBRID;
        const int wneed = wide ? 4 : 3;
        if (p->width[d] != wneed)
            return fail(NUC3D_ERR_UNSUPPORTED, wneed == 3 ? "ghost width must be 3 for WENO5 schemes"
                                                          : "ghost width must be 4 for the central and hybrid schemes");
        if (wneed > G) G = wneed;   // block.cpp:506-507: the buffer is the widest scheme's
        if (p->diff[d] != NUC3D_CD6 && p->diff[d] != NUC3D_CD8) return fail(NUC3D_ERR_BAD_ARG, "diff scheme");
        // cd8 reads four cells each side: a width-4 direction (the central and
        // hybrid schemes); the reference reads out of bounds with 3
        if (p->diff[d] == NUC3D_CD8 && p->width[d] != 4)
            return fail(NUC3D_ERR_UNSUPPORTED, "cd8 needs ghost width 4 (centraldifference8th.cpp:34,49): use it "
                                               "with the central or hybrid reconstruction");
    }
    if (p->model != NUC3D_EULER3D && p->model != NUC3D_NAVIERSTOKES3D) return fail(NUC3D_ERR_BAD_ARG, "model");
    if (p->riemann < 0 || p->riemann > 2) return fail(NUC3D_ERR_BAD_ARG, "riemann");
    nuc3d_case *c = new nuc3d_case();
    c->p = *p;
    c->nb = nblocks;
    c->G = G;
    c->L = p->exact ? &kExact : &kFast;
    (void)hipGetDevice(&c->device);
    nuc3d_kparams &K = c->K;
    K.gamma = p->gamma;
    K.mach = p->mach;
    K.reynolds = p->reynolds;
    K.prandtl = p->prandtl;
    K.sref = p->t_ref / p->t_inf;
    K.twall = p->t_wall;
    K.cfl = 0.5;
    K.riemann = p->riemann;
    for (int d = 0; d < 3; d++) K.width[d] = p->width[d];
    for (int d = 0; d < 3; d++) K.diff[d] = p->diff[d] == NUC3D_CD8 ? 1 : 0;
    K.zquirk = p->zquirk;
    K.nblocks = nblocks;
    K.sensor = nullptr;
    K.hcount = nullptr;
    K.pq = nullptr;
    K.pq_n = 0;
    K.qok = nullptr;
    K.ieee_div = 0;
    K.hrhs = 0;
    K.ngroups = 1;
    const bool ns = p->model == NUC3D_NAVIERSTOKES3D;
    // Q x 3, RHS, J, metrics, dxyz, u v w T (viscous terms and diagnostics),
    // and for Navier-Stokes the 12 viscous flux components
    const int nfields = 15 + 5 + 1 + 9 + 9 + 4 + (ns ? 12 : 0);
    c->hb.resize(nblocks);
    c->rdok.assign(nblocks, 0);
    std::vector<nuc3d_tile> tiles;
    for (int b = 0; b < nblocks; b++) {
        nuc3d_block_dev &B = c->hb[b];
        std::memset(&B, 0, sizeof B);
        for (int d = 0; d < 3; d++) {
            B.n[d] = dims[3 * b + d];
            if (B.n[d] < 2 * c->G + 2) {
                nuc3d_case_destroy(c);
                return fail(NUC3D_ERR_BAD_ARG, "block dimension below 2*ghost+2");
            }
        }
        B.G = c->G;
        B.group = 0;
        B.mach = p->mach;
        B.reynolds = p->reynolds;
        const long long px = (XLEAD + B.n[0] + B.G + ROWALIGN - 1) / ROWALIGN * ROWALIGN;
        const long long py = B.n[1] + 2 * B.G, pz = B.n[2] + 2 * B.G;
        B.sy = px;
        B.sz = px * py;
        const long long P = px * py * pz;
        c->padded.push_back(P);
        void *mem = nullptr;
        if (hipMalloc(&mem, sizeof(double) * P * nfields) != hipSuccess) {
            nuc3d_case_destroy(c);
            return fail(NUC3D_ERR_OOM, "device allocation of block fields failed");
        }
        c->chunks.push_back(mem);
        if (hipMemset(mem, 0, sizeof(double) * P * nfields) != hipSuccess) {
            nuc3d_case_destroy(c);
            return fail(NUC3D_ERR_HIP, "hipMemset failed");
        }
        const long long org = origin(B);  // offset of cell (0,0,0)
        double *base = static_cast<double *>(mem);
        int f = 0;
        auto next = [&]() { return base + P * (f++) + org; };
        for (int s = 0; s < 3; s++)
            for (int q = 0; q < 5; q++) B.q[s][q] = next();
        for (int q = 0; q < 5; q++) B.rhs[q] = next();
        B.jac = next();
        for (int q = 0; q < 9; q++) B.met[q] = next();
        for (int q = 0; q < 9; q++) B.dxyz[q] = next();
        for (int q = 0; q < 4; q++) B.prim[q] = next();
        if (ns) {
            for (int q = 0; q < 12; q++) B.fv[q] = next();
        }
        for (int fc = 0; fc < 6; fc++) {
            B.ftype[fc] = -1;
            B.fid[fc] = NUC3D_BC_OUTLET;
        }
        // hybrid decisions per face: made at stage 0, reused by stages 1 and 2
        for (int d = 0; d < 3; d++) {
            if (p->scheme[d] != NUC3D_HYBRID) continue;
            // the step's decisions, then those of the last single RHS evaluation
            const size_t nf = 2 * (size_t)(B.n[d] + 1) * B.n[(d + 1) % 3] * B.n[(d + 2) % 3];
            void *fm = nullptr;
            if (hipMalloc(&fm, nf) != hipSuccess || hipMemset(fm, 0xff, nf) != hipSuccess) {
                (void)hipFree(fm);
                nuc3d_case_destroy(c);
                return fail(NUC3D_ERR_OOM, "device allocation of hybrid decisions failed");
            }
            c->dflags.push_back(fm);
            B.hflag[d] = static_cast<unsigned char *>(fm);
        }
        c->kseg.push_back(KSEG);
        for (int k0 = 0; k0 < B.n[2]; k0 += KSEG)
            for (int j0 = 0; j0 < B.n[1]; j0 += NUC3D_TILE_Y)
                for (int i0 = 0; i0 < B.n[0]; i0 += NUC3D_TILE_X) {
                    nuc3d_tile t = {b, i0, j0, k0, std::min(KSEG, B.n[2] - k0), 0, 0, 0};
                    tiles.push_back(t);
                }
    }
    c->ntiles = (int)tiles.size();
    // whole lines of the CRWENO5 directions and their Thomas scratch
    size_t scratch = 0;
    for (int d = 0; d < 3; d++) {
        if (p->scheme[d] != NUC3D_CRWENO5) continue;
        std::vector<int4> ln;
        for (int b = 0; b < nblocks; b++) {
            const nuc3d_block_dev &B = c->hb[b];
            const int d2 = d == 0 ? 1 : 0, d3 = d == 2 ? 1 : 2;  // transverse axes: (j,k) / (i,k) / (i,j)
            for (int a3 = 0; a3 < B.n[d3]; a3++)
                for (int a2 = 0; a2 < B.n[d2]; a2++) ln.push_back(make_int4(b, a2, a3, 0));
            // groups of NUC3D_CRW_LINES lines of one block, the last one padded
            while (ln.size() % NUC3D_CRW_LINES) ln.push_back(make_int4(-1, 0, 0, 0));
            c->nmax[d] = std::max(c->nmax[d], B.n[d]);
        }
        if (c->nmax[d] > 1024) {
            nuc3d_case_destroy(c);
            return fail(NUC3D_ERR_UNSUPPORTED, "CRWENO5 lines longer than 1024 cells");
        }
        c->nlines[d] = (int)ln.size();
        scratch = std::max(scratch, ln.size() / NUC3D_CRW_LINES * 128 * (size_t)c->nmax[d]);
        if (hipMalloc(&c->dlines[d], sizeof(int4) * ln.size()) != hipSuccess ||
            hipMemcpy(c->dlines[d], ln.data(), sizeof(int4) * ln.size(), hipMemcpyHostToDevice) != hipSuccess) {
            nuc3d_case_destroy(c);
            return fail(NUC3D_ERR_OOM, "device allocation of CRWENO5 lines failed");
        }
    }
    if (scratch && hipMalloc(&c->dscratch, sizeof(double) * scratch) != hipSuccess) {
        nuc3d_case_destroy(c);
        return fail(NUC3D_ERR_OOM, "device allocation of CRWENO5 scratch failed");
    }
    // direction sweeps: 16 faces x 16 lines for every block until the autotuner chooses
    c->swnf.assign(nblocks, {16, 16, 16});
    if (int rc = build_sweeps(c)) {
        nuc3d_case_destroy(c);
        return rc;
    }
    // work items of the kernel that finishes stage 2, and their blocks (the
    // z sweep's are set by build_sweeps)
    if (p->scheme[2] == NUC3D_CRWENO5) {
        if (ns) {
            for (const auto &t : tiles) c->work_block.push_back(t.b);
        } else {
            const long long L = NUC3D_CRW_LINES;
            for (int b = 0; b < nblocks; b++)
                for (long long g = 0; g < ((long long)c->hb[b].n[0] * c->hb[b].n[1] + L - 1) / L; g++)
                    c->work_block.push_back(b);
        }
        c->nwork = (int)c->work_block.size();
        if (hipMalloc(&c->dresid, sizeof(double) * 2 * c->nwork) != hipSuccess ||
            hipMemset(c->dresid, 0, sizeof(double) * 2 * c->nwork) != hipSuccess) {
            nuc3d_case_destroy(c);
            return fail(NUC3D_ERR_OOM, "device allocation of residual partials failed");
        }
    }
    if (hipMalloc(&c->duniform, sizeof(int) * nblocks) != hipSuccess ||
        hipMalloc(&c->dconsts, sizeof(double) * 8) != hipSuccess ||
        hipMalloc(&c->dqueue, sizeof(unsigned) * NUC3D_QUEUE_WORDS) != hipSuccess ||
        hipMalloc(&c->dqok, sizeof(unsigned) * 4) != hipSuccess ||
        hipMalloc(&c->db, sizeof(nuc3d_block_dev) * nblocks) != hipSuccess ||
        hipMalloc(&c->dtiles, sizeof(nuc3d_tile) * tiles.size()) != hipSuccess ||
        hipMalloc(&c->ddt, sizeof(double) * NUC3D_MAX_BLOCKS) != hipSuccess ||
        hipMalloc(&c->dlam, sizeof(unsigned long long) * 6 * nblocks) != hipSuccess ||
        hipMalloc(&c->dcount, sizeof(unsigned long long) * 4) != hipSuccess ||
        hipMalloc(&c->dhcount, sizeof(unsigned long long) * 16 * NUC3D_HCOUNT_SLOTS) != hipSuccess) {
        nuc3d_case_destroy(c);
        return fail(NUC3D_ERR_OOM, "device allocation of control arrays failed");
    }
    c->K.hcount = c->dhcount;
    if (ns && (hipStreamCreateWithFlags(&c->vstream, hipStreamNonBlocking) != hipSuccess ||
               hipEventCreateWithFlags(&c->vfork, hipEventDisableTiming) != hipSuccess ||
               hipEventCreateWithFlags(&c->vjoin, hipEventDisableTiming) != hipSuccess)) {
        nuc3d_case_destroy(c);
        return fail(NUC3D_ERR_HIP, "stream for the viscous-flux pass");
    }
    if (hipMemcpy(c->dtiles, tiles.data(), sizeof(nuc3d_tile) * tiles.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(c->ddt, 0, sizeof(double) * NUC3D_MAX_BLOCKS) != hipSuccess ||
        hipMemset(c->dlam, 0, sizeof(unsigned long long) * 6 * nblocks) != hipSuccess ||
        hipMemset(c->dcount, 0, sizeof(unsigned long long) * 4) != hipSuccess ||
        hipMemset(c->dqueue, 0, sizeof(unsigned) * NUC3D_QUEUE_WORDS) != hipSuccess ||
        hipMemset(c->dqok, 0, sizeof(unsigned) * 4) != hipSuccess ||
        hipMemset(c->dhcount, 0, sizeof(unsigned long long) * 16 * NUC3D_HCOUNT_SLOTS) != hipSuccess) {
        nuc3d_case_destroy(c);
        return fail(NUC3D_ERR_HIP, "initialising control arrays failed");
    }
    *out = c;
    return NUC3D_OK;
}

int nuc3d_case_destroy(nuc3d_case *c)
{
    if (!c) return NUC3D_OK;
    (void)hipDeviceSynchronize();
    for (auto &g : c->gexec)
        if (g) (void)hipGraphExecDestroy(g);
    if (c->gstream) (void)hipStreamDestroy(c->gstream);
    if (c->vstream) (void)hipStreamDestroy(c->vstream);
    if (c->vfork) (void)hipEventDestroy(c->vfork);
    if (c->vjoin) (void)hipEventDestroy(c->vjoin);
    for (auto &l : c->ev)
        for (auto &pr : l) c->ev_pool.push_back(pr.first), c->ev_pool.push_back(pr.second);
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    for (void *m : c->chunks) (void)hipFree(m);
    (void)hipFree(c->db);
    (void)hipFree(c->duniform);
    (void)hipFree(c->dconsts);
    (void)hipFree(c->dqueue);
    (void)hipFree(c->dqok);
    (void)hipFree(c->dtiles);
    for (int d = 0; d < 3; d++)
        for (int g = 0; g < 2; g++) (void)hipFree(c->dsweep[d][g]);
    for (int d = 0; d < 3; d++) (void)hipFree(c->dlines[d]);
    (void)hipFree(c->dscratch);
    (void)hipFree(c->ddt);
    (void)hipFree(c->dlam);
    (void)hipFree(c->dcount);
    (void)hipFree(c->dhcount);
    (void)hipFree(c->dresid);
    (void)hipFree(c->dsensor);
    for (void *f : c->dflags) (void)hipFree(f);
    delete c;
    return NUC3D_OK;
}

int nuc3d_set_faces(nuc3d_case *c, int b, const int *type6, const int *id6, const double *val30)
{
    int rc = check_block(c, b);
    if (rc) return rc;
    if (!type6 || !id6) return fail(NUC3D_ERR_BAD_ARG, "null face arrays");
    nuc3d_block_dev &B = c->hb[b];
    for (int f = 0; f < 6; f++) {
        const int t = type6[f], id = id6[f];
        if (t == -1) {
            if (id < 0 || id > 3) return fail(NUC3D_ERR_BAD_ARG, "BC id must be 0..3 (boundaryConditions.hpp:56-61)");
        } else {
            if (t < 0 || t >= c->nb) return fail(NUC3D_ERR_BAD_ARG, "neighbour block out of range");
            if (id < 0 || id > 5 || id / 2 != f / 2)
                return fail(NUC3D_ERR_UNSUPPORTED, "connected faces must join the same direction");
            const nuc3d_block_dev &N = c->hb[t];
            const int d = f / 2;
            if (N.n[(d + 1) % 3] != B.n[(d + 1) % 3] || N.n[(d + 2) % 3] != B.n[(d + 2) % 3])
                return fail(NUC3D_ERR_BAD_ARG, "connected faces differ in tangential size");
        }
        B.ftype[f] = t;
        B.fid[f] = id;
        for (int q = 0; q < 5; q++) B.fval[f][q] = val30 ? val30[5 * f + q] : 0.0;
    }
    c->dirty = true;
    c->faces_set = true;
    return NUC3D_OK;
}

int nuc3d_upload_metrics(nuc3d_case *c, int b, const double *jac, const double *met, const double *dxyz)
{
    int rc = check_block(c, b);
    if (rc) return rc;
    if (!jac || !met) return fail(NUC3D_ERR_BAD_ARG, "null metrics");
    nuc3d_block_dev &B = c->hb[b];
    const long long N = (long long)B.n[0] * B.n[1] * B.n[2];
    if ((rc = put_field(c, b, B.jac, jac))) return rc;
    for (int q = 0; q < 9; q++)
        if ((rc = put_field(c, b, B.met[q], met + q * N))) return rc;
    if (dxyz)
        for (int q = 0; q < 9; q++)
            if ((rc = put_field(c, b, B.dxyz[q], dxyz + q * N))) return rc;
    c->dirty = true;
    return NUC3D_OK;
}

int nuc3d_upload_mesh(nuc3d_case *c, int b, const double *xn, const double *yn, const double *zn, double *centre3)
{
    int rc = check_block(c, b);
    if (rc) return rc;
    if (!xn || !yn || !zn) return fail(NUC3D_ERR_BAD_ARG, "null node coordinates");
    const nuc3d_block_dev &B = c->hb[b];
    for (int d = 0; d < 3; d++)
        if (B.n[d] % 4)
            return fail(NUC3D_ERR_UNSUPPORTED,
                        "mesh metrics need cells per direction divisible by 4 (tiles of block.cpp:163-165)");
    const long long N = (long long)B.n[0] * B.n[1] * B.n[2];
    const long long NN = (long long)(B.n[0] + 1) * (B.n[1] + 1) * (B.n[2] + 1);
    OpBuffers T;
    double *dn = T.get(3 * NN), *dc = centre3 ? T.get(3 * N) : nullptr;
    int *dbad = reinterpret_cast<int *>(T.get(1));
    if (!dn || !dbad || (centre3 && !dc)) return fail(NUC3D_ERR_OOM, "device allocation failed");
    HIPCHK(hipMemcpy(dn, xn, sizeof(double) * NN, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dn + NN, yn, sizeof(double) * NN, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dn + 2 * NN, zn, sizeof(double) * NN, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(dbad, 0, sizeof(int)));
    HIPCHK(nuc3d_launch_mesh_metrics(B, dn, dn + NN, dn + 2 * NN, dc, dbad, nullptr));
    int bad = 0;
    HIPCHK(hipMemcpy(&bad, dbad, sizeof(int), hipMemcpyDeviceToHost));
    if (centre3) HIPCHK(hipMemcpy(centre3, dc, sizeof(double) * 3 * N, hipMemcpyDeviceToHost));
    c->dirty = true;
    if (bad) return fail(NUC3D_ERR_BAD_ARG, "non-positive Jacobian in the mesh of this block");
    return NUC3D_OK;
}

int nuc3d_download_metrics(nuc3d_case *c, int b, double *jac, double *met, double *dxyz)
{
    int rc = check_block(c, b);
    if (rc) return rc;
    nuc3d_block_dev &B = c->hb[b];
    const long long N = (long long)B.n[0] * B.n[1] * B.n[2];
    HIPCHK(hipDeviceSynchronize());
    if (jac && (rc = get_field(c, b, jac, B.jac))) return rc;
    for (int q = 0; q < 9; q++) {
        if (met && (rc = get_field(c, b, met + q * N, B.met[q]))) return rc;
        if (dxyz && (rc = get_field(c, b, dxyz + q * N, B.dxyz[q]))) return rc;
    }
    return NUC3D_OK;
}

}  // extern "C"

namespace {

// Tecplot ASCII file of block::outputQ_tecplot / outputGEO_tecplot
// (block.cpp:692-752, 857-922): the header, the node coordinates x, y, z
// ((nx+1)(ny+1)(nz+1) each) and nv cell-centred fields (nx ny nz each), one
// value per line as block::writeField prints them (std::setprecision(12),
// block.cpp:102-117, i.e. %.12g), i fastest.
int write_tecplot_file(const char *path, const nuc3d_block_dev &B, const double *xn, const double *yn,
                       const double *zn, const std::vector<const double *> &fields)
{
    FILE *fp = std::fopen(path, "w");
    if (!fp) return fail(NUC3D_ERR_IO, std::string("cannot open ") + path + " for writing");
    const int nx = B.n[0], ny = B.n[1], nz = B.n[2], nv = (int)fields.size();
    std::fprintf(fp, "title=NUC3d\nvariables=x,y,z");
    for (int i = 0; i < nv; i++) std::fprintf(fp, ",Val_%d", i);
    std::fprintf(fp, "\n Zone I = %d, J= %d, K=%d\n DATAPACKING=BLOCK, VARLOCATION=([%d-%d]=CELLCENTERED)\n", nx + 1,
                 ny + 1, nz + 1, 4, 3 + nv);
    const size_t NN = (size_t)(nx + 1) * (ny + 1) * (nz + 1), N = (size_t)nx * ny * nz;
    for (const double *a : {xn, yn, zn})
        for (size_t e = 0; e < NN; e++) std::fprintf(fp, "%.12g\n", a[e]);
    for (const double *a : fields)
        for (size_t e = 0; e < N; e++) std::fprintf(fp, "%.12g\n", a[e]);
    const bool ok = !std::ferror(fp);
    if (std::fclose(fp) != 0 || !ok) return fail(NUC3D_ERR_IO, std::string("writing ") + path + " failed");
    return NUC3D_OK;
}

// physicsModel::getPrim on the host (con2prim with the e<0 fix-up and
// EoSIdealGasFWD, physicsModel.cpp:313-353, 981-1054), in the reference's
// operation order: rho, u, v, w, p, then T, e, c.
void host_prim(const double *q5, const double *jac, size_t N, double gamma, double mach, std::vector<double> &w8)
{
#pragma clang fp contract(off)
    w8.assign(8 * N, 0.0);
    for (size_t x = 0; x < N; x++) {
        const double J = jac[x];
        const double rho = q5[x] * J;
        const double u = q5[N + x] / rho * J;
        const double v = q5[2 * N + x] / rho * J;
        const double w = q5[3 * N + x] / rho * J;
        double E = q5[4 * N + x] * J;
        const double e0 = E - 0.5 * rho * (u * u + v * v + w * w);
        if (e0 < 0.0) E = (std::pow(rho, gamma) / ((gamma * mach * mach) * (gamma - 1.0)) + 0.5 * rho * (u * u + v * v + w * w));
        const double p = (E - 0.5 * rho * (u * u + v * v + w * w)) * (gamma - 1.0);
        w8[x] = rho;
        w8[N + x] = u;
        w8[2 * N + x] = v;
        w8[3 * N + x] = w;
        w8[4 * N + x] = p;
        w8[5 * N + x] = gamma * mach * mach * p / rho;
        w8[6 * N + x] = p / rho / (gamma - 1.0);
        w8[7 * N + x] = std::sqrt(std::fabs(gamma * p / rho));
    }
}

}  // namespace

extern "C" {

int nuc3d_write_tecplot(nuc3d_case *c, int b, const char *path, const double *xn, const double *yn, const double *zn)
{
    int rc = check_block(c, b);
    if (rc) return rc;
    if (!path || !xn || !yn || !zn) return fail(NUC3D_ERR_BAD_ARG, "null argument");
    if (c->next_stage != 0) return fail(NUC3D_ERR_ORDER, "flow files are written between full steps");
    const nuc3d_block_dev &B = c->hb[b];
    const size_t N = (size_t)B.n[0] * B.n[1] * B.n[2];
    std::vector<double> q(5 * N), jac(N), w8;
    if ((rc = nuc3d_download_q(c, b, q.data()))) return rc;
    if ((rc = get_field(c, b, jac.data(), B.jac))) return rc;
    host_prim(q.data(), jac.data(), N, c->p.gamma, B.mach, w8);
    std::vector<const double *> f;
    for (int v = 0; v < 8; v++) f.push_back(w8.data() + v * N);
    return write_tecplot_file(path, B, xn, yn, zn, f);
}

int nuc3d_write_tecplot_geo(nuc3d_case *c, int b, const char *path, const double *xn, const double *yn,
                            const double *zn)
{
    int rc = check_block(c, b);
    if (rc) return rc;
    if (!path || !xn || !yn || !zn) return fail(NUC3D_ERR_BAD_ARG, "null argument");
    const nuc3d_block_dev &B = c->hb[b];
    const size_t N = (size_t)B.n[0] * B.n[1] * B.n[2];
    std::vector<double> met(9 * N);
    HIPCHK(hipDeviceSynchronize());
    for (int m = 0; m < 9; m++)
        if ((rc = get_field(c, b, met.data() + m * N, B.met[m]))) return rc;
    std::vector<const double *> f;
    for (int m = 0; m < 9; m++) f.push_back(met.data() + m * N);
    return write_tecplot_file(path, B, xn, yn, zn, f);
}

int nuc3d_save_q(nuc3d_case *c, int b, const char *path, double time)
{
    int rc = check_block(c, b);
    if (rc) return rc;
    if (!path) return fail(NUC3D_ERR_BAD_ARG, "null path");
    if (c->next_stage != 0) return fail(NUC3D_ERR_ORDER, "restart files are written between full steps");
    const nuc3d_block_dev &B = c->hb[b];
    const size_t N = (size_t)B.n[0] * B.n[1] * B.n[2];
    std::vector<double> q(5 * N);
    if ((rc = nuc3d_download_q(c, b, q.data()))) return rc;
    FILE *fp = std::fopen(path, "wb");
    if (!fp) return fail(NUC3D_ERR_IO, std::string("cannot open ") + path + " for writing");
    const bool ok = std::fwrite(&time, sizeof(double), 1, fp) == 1 && std::fwrite(q.data(), sizeof(double), 5 * N, fp) == 5 * N;
    if (std::fclose(fp) != 0 || !ok) return fail(NUC3D_ERR_IO, std::string("writing ") + path + " failed");
    return NUC3D_OK;
}

int nuc3d_load_q(nuc3d_case *c, int b, const char *path, double *time)
{
    int rc = check_block(c, b);
    if (rc) return rc;
    if (!path) return fail(NUC3D_ERR_BAD_ARG, "null path");
    const nuc3d_block_dev &B = c->hb[b];
    const size_t N = (size_t)B.n[0] * B.n[1] * B.n[2];
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return fail(NUC3D_ERR_IO, std::string("cannot open ") + path);
    std::fseek(fp, 0, SEEK_END);
    const long size = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    if (size != (long)(sizeof(double) * (1 + 5 * N))) {
        std::fclose(fp);
        return fail(NUC3D_ERR_BAD_ARG, std::string(path) + " does not hold time + 5 fields of this block");
    }
    double t = 0.0;
    std::vector<double> q(5 * N);
    const bool ok = std::fread(&t, sizeof(double), 1, fp) == 1 && std::fread(q.data(), sizeof(double), 5 * N, fp) == 5 * N;
    std::fclose(fp);
    if (!ok) return fail(NUC3D_ERR_IO, std::string("reading ") + path + " failed");
    if (time) *time = t;
    return nuc3d_upload_q(c, b, q.data());
}

int nuc3d_upload_q(nuc3d_case *c, int b, const double *q5)
{
    int rc = check_block(c, b);
    if (rc) return rc;
    if (!q5) return fail(NUC3D_ERR_BAD_ARG, "null state");
    if (c->next_stage != 0) return fail(NUC3D_ERR_ORDER, "upload in the middle of an RK step");
    nuc3d_block_dev &B = c->hb[b];
    const long long N = (long long)B.n[0] * B.n[1] * B.n[2];
    for (int q = 0; q < 5; q++)
        if ((rc = put_field(c, b, B.q[c->cur][q], q5 + q * N))) return rc;
    if ((rc = clear_positivity(c))) return rc;
    c->dirty = true;
    return NUC3D_OK;
}

int nuc3d_download_q(nuc3d_case *c, int b, double *q5)
{
    int rc = check_block(c, b);
    if (rc) return rc;
    if (!q5) return fail(NUC3D_ERR_BAD_ARG, "null output");
    HIPCHK(hipDeviceSynchronize());
    nuc3d_block_dev &B = c->hb[b];
    const long long N = (long long)B.n[0] * B.n[1] * B.n[2];
    for (int q = 0; q < 5; q++)
        if ((rc = get_field(c, b, q5 + q * N, B.q[c->cur][q]))) return rc;
    return NUC3D_OK;
}

int nuc3d_download_rhs(nuc3d_case *c, int b, double *rhs5)
{
    int rc = check_block(c, b);
    if (rc) return rc;
    if (!rhs5) return fail(NUC3D_ERR_BAD_ARG, "null output");
    HIPCHK(hipDeviceSynchronize());
    nuc3d_block_dev &B = c->hb[b];
    const long long N = (long long)B.n[0] * B.n[1] * B.n[2];
    for (int q = 0; q < 5; q++)
        if ((rc = get_field(c, b, rhs5 + q * N, B.rhs[q]))) return rc;
    return NUC3D_OK;
}

int nuc3d_rhs(nuc3d_case *c, void *stream)
{
    if (!c) return fail(NUC3D_ERR_BAD_ARG, "null case");
    hipStream_t s = static_cast<hipStream_t>(stream);
    int rc = refresh(c, s);
    if (rc) return rc;
    // eigenvalues of the evaluated state (the stage-0 maxima stay untouched)
    unsigned long long *lam = c->dlam + 3 * c->nb * (1 - c->lampar);
    if (c->next_stage != 0) {
        HIPCHK(hipMemsetAsync(lam, 0, sizeof(unsigned long long) * 3 * c->nb, s));
        HIPCHK(c->L->eigen(c->db, c->dtiles, c->ntiles, c->K, c->cur, lam, s));
    }
    if ((rc = launch_rhs(c, 0, 0, c->cur, 0, 0, c->dlam, lam, nullptr, s))) return rc;
    c->flags_rhs = true;
    return NUC3D_OK;
}

int nuc3d_stage(nuc3d_case *c, int stage, double cfl, void *stream)
{
    if (!c) return fail(NUC3D_ERR_BAD_ARG, "null case");
    if (stage < 0 || stage > 2) return fail(NUC3D_ERR_BAD_ARG, "stage must be 0, 1 or 2");
    if (stage != c->next_stage) return fail(NUC3D_ERR_ORDER, "RK stages must run 0, 1, 2 in order");
    hipStream_t s = static_cast<hipStream_t>(stream);
    int rc = refresh(c, s);
    if (rc) return rc;
    c->K.cfl = cfl;
    const bool ns = c->p.model == NUC3D_NAVIERSTOKES3D;
    // buffers: stage 0 A->B, stage 1 (B, u_n=A)->C, stage 2 (C, u_n=A)->A
    const int bin = stage, bun = 0, bout = stage == 2 ? 0 : stage + 1;
    (void)ns;
    unsigned long long *lcur = c->dlam + 3 * c->nb * c->lampar;
    unsigned long long *lnext = c->dlam + 3 * c->nb * (1 - c->lampar);
    if ((rc = launch_rhs(c, 1, stage, bin, bun, bout, lcur, lnext, c->dresid + (long long)c->lampar * c->nwork, s,
                         c->steps_done)))
        return rc;
    if (stage == 0) c->hforce = false;
    c->flags_rhs = false;
    c->cur = bout;
    c->next_stage = (stage + 1) % 3;
    if (stage == 2) {
        c->lampar ^= 1;
        c->steps_done++;
    }
    return NUC3D_OK;
}

int nuc3d_steps(nuc3d_case *c, int nsteps, double cfl, void *stream)
{
    if (!c) return fail(NUC3D_ERR_BAD_ARG, "null case");
    if (nsteps < 0) return fail(NUC3D_ERR_BAD_ARG, "negative step count");
    for (int n = 0; n < nsteps; n++)
        for (int st = 0; st < 3; st++) {
            int rc = nuc3d_stage(c, st, cfl, stream);
            if (rc) return rc;
        }
    return NUC3D_OK;
}

int nuc3d_step_graph(nuc3d_case *c, int nsteps, double cfl, void *stream)
{
    if (!c) return fail(NUC3D_ERR_BAD_ARG, "null case");
    if (nsteps < 0) return fail(NUC3D_ERR_BAD_ARG, "negative step count");
    if (c->next_stage != 0) return fail(NUC3D_ERR_ORDER, "graph steps start at stage 0");
    // with kernel timing on, every launch is bracketed by events: run eagerly
    if (c->timing) return nuc3d_steps(c, nsteps, cfl, stream);
    hipStream_t s = static_cast<hipStream_t>(stream);
    int rc = refresh(c, s);
    if (rc) return rc;
    int n = 0;
    // the captured steps decide on the interval's pattern only: a decision
    // forced by a new state or sensor is made by an eager step
    if (c->hforce && nsteps > 0) {
        if ((rc = nuc3d_steps(c, 1, cfl, stream))) return rc;
        n = 1;
    }
    for (; n + 2 <= nsteps; n += 2) {
        const int par = c->lampar;
        if (!c->gexec[par] || c->gcfl[par] != cfl) {
            if (c->gexec[par]) {
                HIPCHK(hipGraphExecDestroy(c->gexec[par]));
                c->gexec[par] = nullptr;
            }
            if (!c->gstream) HIPCHK(hipStreamCreateWithFlags(&c->gstream, hipStreamNonBlocking));
            // capture two steps: the eigenvalue buffers swap every step, so two
            // steps return every buffer index to where it started
            HIPCHK(hipStreamBeginCapture(c->gstream, hipStreamCaptureModeThreadLocal));
            const double cfl0 = c->K.cfl;
            c->K.cfl = cfl;
            int lp = par;
            for (int st = 0; st < 6 && rc == 0; st++) {
                const int stage = st % 3;
                const int bin = stage, bun = 0, bout = stage == 2 ? 0 : stage + 1;
                // steps_done and the eigenvalue parity advance together, so a
                // graph (kept per parity) sees the same step pattern at every replay
                rc = launch_rhs(c, 1, stage, bin, bun, bout, c->dlam + 3 * c->nb * lp,
                                c->dlam + 3 * c->nb * (1 - lp), c->dresid + (long long)lp * c->nwork, c->gstream,
                                c->steps_done + st / 3);
                if (stage == 2) lp ^= 1;
            }
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(c->gstream, &g);
            c->K.cfl = cfl0;
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return hipfail(e, "hipStreamEndCapture");
            const hipError_t ei = hipGraphInstantiate(&c->gexec[par], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ei != hipSuccess) {
                c->gexec[par] = nullptr;
                return hipfail(ei, "hipGraphInstantiate");
            }
            c->gcfl[par] = cfl;
        }
        HIPCHK(hipGraphLaunch(c->gexec[par], s));
        c->K.cfl = cfl;
        c->cur = 0;
        c->steps_done += 2;
    }
    if (n < nsteps) return nuc3d_steps(c, nsteps - n, cfl, stream);
    return NUC3D_OK;
}

int nuc3d_diagnostics(nuc3d_case *c, double *enstrophy, double *kinetic, int b, double *fields5, void *stream)
{
    if (!c || !enstrophy || !kinetic) return fail(NUC3D_ERR_BAD_ARG, "null argument");
    if (fields5 && (b < 0 || b >= c->nb)) return fail(NUC3D_ERR_BAD_ARG, "block index out of range");
    hipStream_t s = static_cast<hipStream_t>(stream);
    int rc = refresh(c, s);
    if (rc) return rc;
    const long long nwg = (long long)c->ntiles * c->kmax;
    long long N = 0;
    if (fields5) N = (long long)c->hb[b].n[0] * c->hb[b].n[1] * c->hb[b].n[2];
    OpBuffers T;
    double *dpart = T.get(2 * nwg), *dfields = fields5 ? T.get(5 * N) : nullptr;
    if (!dpart || (fields5 && !dfields)) return fail(NUC3D_ERR_OOM, "device allocation failed");
    HIPCHK(c->L->prim(c->db, c->dtiles, c->ntiles, c->kmax, c->K, c->cur, s));
    HIPCHK(c->L->diag(c->db, c->dtiles, c->ntiles, c->kmax, c->K, c->cur, fields5 ? b : -1, dfields, dpart, s));
    std::vector<double> part(2 * nwg);
    HIPCHK(hipMemcpyAsync(part.data(), dpart, sizeof(double) * 2 * nwg, hipMemcpyDeviceToHost, s));
    if (fields5) HIPCHK(hipMemcpyAsync(fields5, dfields, sizeof(double) * 5 * N, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    double e = 0.0, k = 0.0;
    for (long long w = 0; w < nwg; w++) {
        e += part[2 * w];
        k += part[2 * w + 1];
    }
    *enstrophy = e;
    *kinetic = k;
    return NUC3D_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Scratch of the entry point that allocates nothing (nuc3d_diagnostics_noalloc):
// a file-scope device array of NUC3D_NOALLOC_WORK partial pairs plus the two
// sums.  Every GPU the library runs on holds its own copy until the process
// ends; its device address is looked up once per device (hipGetSymbolAddress
// with that device current) and only that address is handed to kernels and
// copies -- the array's name in host code is not a device address.
#define NUC3D_NOALLOC_WORK (1 << 20)
static __device__ __align__(256) double scratch[2 * NUC3D_NOALLOC_WORK + 2];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// the current device's copy of `scratch` (looked up on the first call there)
int scratch_of_device(int dev, double **out)
{
    if (dev < 0 || dev >= kMaxDevices) return fail(NUC3D_ERR_UNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hipfail(e, "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return NUC3D_OK;
}

}  // namespace

extern "C" {

int nuc3d_diagnostics_noalloc(nuc3d_case *c, double *enstrophy, double *kinetic, void *stream)
{
    if (!c || !enstrophy || !kinetic) return fail(NUC3D_ERR_BAD_ARG, "null argument");
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev != c->device) return fail(NUC3D_ERR_BAD_ARG, "the case's device is not the current device");
    const long long nwg = (long long)c->ntiles * c->kmax;
    if (nwg > NUC3D_NOALLOC_WORK)
        return fail(NUC3D_ERR_UNSUPPORTED, "case too large for the fixed scratch of nuc3d_diagnostics_noalloc (" +
                                               std::to_string(nwg) + " work items > " +
                                               std::to_string(NUC3D_NOALLOC_WORK) + "); use nuc3d_diagnostics");
    double *sp = nullptr;
    int rc = scratch_of_device(dev, &sp);
    if (rc) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if ((rc = refresh(c, s))) return rc;
    // every partial pair is written by its workgroup before the sum reads it
    HIPCHK(c->L->prim(c->db, c->dtiles, c->ntiles, c->kmax, c->K, c->cur, s));
    HIPCHK(c->L->diag(c->db, c->dtiles, c->ntiles, c->kmax, c->K, c->cur, -1, nullptr, sp, s));
    double *sums = sp + 2 * (long long)NUC3D_NOALLOC_WORK;
    HIPCHK(nuc3d_launch_sum_pairs(sp, nwg, sums, s));
    double h[2];
    HIPCHK(hipMemcpyAsync(h, sums, sizeof h, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    *enstrophy = h[0];
    *kinetic = h[1];
    return NUC3D_OK;
}

int nuc3d_run(nuc3d_case *c, const nuc3d_run_config *cfg, nuc3d_run_status *out, void *stream)
{
    if (!c || !cfg) return fail(NUC3D_ERR_BAD_ARG, "null argument");
    if (cfg->save_step < 0 || cfg->post_step < 0) return fail(NUC3D_ERR_BAD_ARG, "negative save or post interval");
    if (c->K.ngroups > 1) return fail(NUC3D_ERR_UNSUPPORTED, "the time loop runs one case (group) at a time");
    int step = cfg->start_step;
    double time = cfg->time, dt = 0.0, res = 0.0;
    int rc;
    const std::string dir = cfg->dir ? cfg->dir : "";
    while (time <= cfg->tl && step < cfg->end_step) {  // IOController::ifsolve
        if ((rc = nuc3d_steps(c, 1, cfg->cfl, stream))) return rc;
        if ((rc = nuc3d_get_dt(c, &dt))) return rc;
        if ((rc = nuc3d_get_residual(c, &res, nullptr))) return rc;
        step++;
        time += dt;
        if (cfg->verbose)
            std::printf("========step = %d\n time = %.6g, dt = %.6g, residual = %.6g\n", step, time, dt, res);
        if (cfg->post_step > 0 && step % cfg->post_step == 0) {  // ifpost
            double ens = 0.0, kin = 0.0;
            if ((rc = nuc3d_diagnostics(c, &ens, &kin, 0, nullptr, stream))) return rc;
            if (cfg->dir) {
                FILE *fp = std::fopen((dir + "/enstrophy.dat").c_str(), "a");
                if (!fp) return fail(NUC3D_ERR_IO, "cannot append to " + dir + "/enstrophy.dat");
                std::fprintf(fp, "%d %.12g %.12g %.12g\n", step, time, ens, kin);
                std::fclose(fp);
            }
        }
        if (cfg->save_step > 0 && step % cfg->save_step == 0 && cfg->dir) {  // ifsave
            for (int b = 0; b < c->nb; b++) {
                char name[96];
                std::snprintf(name, sizeof name, "/flowData/NUC3d_step_%d_ID_%d.bin", step, b);
                if ((rc = nuc3d_save_q(c, b, (dir + name).c_str(), time))) return rc;
            }
        }
    }
    if (out) {
        out->step = step;
        out->time = time;
        out->dt = dt;
        out->residual = res;
    }
    return NUC3D_OK;
}

int nuc3d_sync(nuc3d_case *c)
{
    if (!c) return fail(NUC3D_ERR_BAD_ARG, "null case");
    HIPCHK(hipDeviceSynchronize());
    return positivity_status(c);
}

int nuc3d_get_positivity(nuc3d_case *c, long long *counts2)
{
    if (!c || !counts2) return fail(NUC3D_ERR_BAD_ARG, "null argument");
    HIPCHK(hipDeviceSynchronize());
    unsigned long long v[3] = {0, 0, 0};
    HIPCHK(hipMemcpy(v, c->dcount, sizeof v, hipMemcpyDeviceToHost));
    counts2[0] = (long long)v[1];
    counts2[1] = (long long)v[2];
    return NUC3D_OK;
}

int nuc3d_get_dt(nuc3d_case *c, double *dt)
{
    if (!c || !dt) return fail(NUC3D_ERR_BAD_ARG, "null argument");
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(dt, c->ddt, sizeof(double), hipMemcpyDeviceToHost));
    return positivity_status(c);
}

int nuc3d_get_dt_group(nuc3d_case *c, int group, double *dt)
{
    if (!c || !dt) return fail(NUC3D_ERR_BAD_ARG, "null argument");
    if (group < 0 || group >= c->K.ngroups) return fail(NUC3D_ERR_BAD_ARG, "group out of range");
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(dt, c->ddt + group, sizeof(double), hipMemcpyDeviceToHost));
    return positivity_status(c);
}

int nuc3d_set_block_group(nuc3d_case *c, int b, int group, double mach, double reynolds)
{
    int rc = check_block(c, b);
    if (rc) return rc;
    if (group < 0 || group >= c->nb) return fail(NUC3D_ERR_BAD_ARG, "group must be in [0, nblocks)");
    if (c->next_stage != 0) return fail(NUC3D_ERR_ORDER, "groups change between full steps");
    nuc3d_block_dev &B = c->hb[b];
    B.group = group;
    B.mach = mach > 0.0 ? mach : c->p.mach;
    B.reynolds = reynolds > 0.0 ? reynolds : c->p.reynolds;
    int ng = 0;
    for (const auto &x : c->hb) ng = std::max(ng, x.group + 1);
    c->K.ngroups = ng;
    // the captured steps hold the launch parameters by value
    for (auto &g : c->gexec)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    c->gcfl[0] = c->gcfl[1] = -1.0;
    c->dirty = true;
    return NUC3D_OK;
}

int nuc3d_get_max_eigen(nuc3d_case *c, int b, double *eig3)
{
    int rc = check_block(c, b);
    if (rc) return rc;
    if (!eig3) return fail(NUC3D_ERR_BAD_ARG, "null output");
    HIPCHK(hipDeviceSynchronize());
    // after a step the maxima of the next stage-0 state sit in lampar;
    // after a lone RHS evaluation in the middle of a step, in the other half.
    const int half = c->next_stage == 0 ? c->lampar : 1 - c->lampar;
    unsigned long long v[3];
    HIPCHK(hipMemcpy(v, c->dlam + 3 * c->nb * half + 3 * b, sizeof v, hipMemcpyDeviceToHost));
    for (int d = 0; d < 3; d++) std::memcpy(&eig3[d], &v[d], sizeof(double));
    return NUC3D_OK;
}

int nuc3d_get_residual(nuc3d_case *c, double *res_global, double *res_blocks)
{
    if (!c || !res_global) return fail(NUC3D_ERR_BAD_ARG, "null argument");
    HIPCHK(hipDeviceSynchronize());
    std::vector<double> res(c->nb, 0.0);
    if (c->steps_done > 0) {
        std::vector<double> part(2 * (size_t)c->nwork);
        HIPCHK(hipMemcpy(part.data(), c->dresid, sizeof(double) * part.size(), hipMemcpyDeviceToHost));
        double dt = 0.0;
        HIPCHK(hipMemcpy(&dt, c->ddt, sizeof(double), hipMemcpyDeviceToHost));
        // the last step wrote the half of the eigenvalue parity it started with
        const int last = 1 - c->lampar, prev = c->lampar;
        std::vector<double> s_last(c->nb, 0.0), s_prev(c->nb, 0.0);
        for (int w = 0; w < c->nwork; w++) {
            s_last[c->work_block[w]] += part[(size_t)last * c->nwork + w];
            s_prev[c->work_block[w]] += part[(size_t)prev * c->nwork + w];
        }
        for (int b = 0; b < c->nb; b++) {
            const double n5 = 5.0 * c->hb[b].n[0] * c->hb[b].n[1] * c->hb[b].n[2];
            const double rms = std::sqrt(s_last[b] / n5);
            const double rms_last = c->steps_done > 1 ? std::sqrt(s_prev[b] / n5) : 0.0;
            res[b] = std::fabs(rms - rms_last) / (dt * dt);
        }
    }
    double g = 0.0;
    for (int b = 0; b < c->nb; b++) g = res[b] > g ? res[b] : g;  // MPI_MAX over ranks
    *res_global = g;
    if (res_blocks)
        for (int b = 0; b < c->nb; b++) res_blocks[b] = res[b];
    return NUC3D_OK;
}

int nuc3d_get_fixups(nuc3d_case *c, long long *count)
{
    if (!c || !count) return fail(NUC3D_ERR_BAD_ARG, "null argument");
    HIPCHK(hipDeviceSynchronize());
    unsigned long long v = 0;
    HIPCHK(hipMemcpy(&v, c->dcount, sizeof v, hipMemcpyDeviceToHost));
    *count = (long long)v;
    return NUC3D_OK;
}

int nuc3d_timing(nuc3d_case *c, int enable)
{
    if (!c) return fail(NUC3D_ERR_BAD_ARG, "null case");
    c->timing = enable != 0;
    return NUC3D_OK;
}

int nuc3d_timing_read(nuc3d_case *c, double *ms, int *launches)
{
    if (!c) return fail(NUC3D_ERR_BAD_ARG, "null case");
    HIPCHK(hipDeviceSynchronize());
    for (int l = 0; l < NUC3D_NTIMING; l++) {
        double acc = 0.0;
        int cnt = 0;
        for (auto &pr : c->ev[l]) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, pr.first, pr.second));
            acc += ms;
            cnt++;
            c->ev_pool.push_back(pr.first);
            c->ev_pool.push_back(pr.second);
        }
        c->ev[l].clear();
        if (ms) ms[l] = acc;
        if (launches) launches[l] = cnt;
    }
    return NUC3D_OK;
}


// ---------------------------------------------------------------------------
// Autotuner of the column tiles (no counterpart in the reference)

}  // extern "C"

namespace {

// column tiles of the given blocks with their k-segment lengths
std::vector<nuc3d_tile> column_tiles(const nuc3d_case *c, const std::vector<int> &blocks, const std::vector<int> &kseg)
{
    std::vector<nuc3d_tile> tiles;
    for (int b : blocks) {
        const nuc3d_block_dev &B = c->hb[b];
        const int ks = kseg[b];
        for (int k0 = 0; k0 < B.n[2]; k0 += ks)
            for (int j0 = 0; j0 < B.n[1]; j0 += NUC3D_TILE_Y)
                for (int i0 = 0; i0 < B.n[0]; i0 += NUC3D_TILE_X) {
                    nuc3d_tile t = {b, i0, j0, k0, std::min(ks, B.n[2] - k0), 0, 0, 0};
                    tiles.push_back(t);
                }
    }
    return tiles;
}

// Replace the case's column tiles (and, where k_visc_div finishes the stages,
// the residual work items that follow them; the residual history restarts).
int set_column_tiles(nuc3d_case *c, const std::vector<int> &kseg)
{
    std::vector<int> all(c->nb);
    for (int b = 0; b < c->nb; b++) all[b] = b;
    const std::vector<nuc3d_tile> tiles = column_tiles(c, all, kseg);
    nuc3d_tile *dt = nullptr;
    if (hipMalloc(&dt, sizeof(nuc3d_tile) * tiles.size()) != hipSuccess) return fail(NUC3D_ERR_OOM, "tile table");
    HIPCHK(hipMemcpy(dt, tiles.data(), sizeof(nuc3d_tile) * tiles.size(), hipMemcpyHostToDevice));
    const bool ns = c->p.model == NUC3D_NAVIERSTOKES3D;
    if (ns && !visc_in_z_sweep(c)) {  // the residual partials come from k_visc_div
        double *dr = nullptr;
        if (hipMalloc(&dr, sizeof(double) * 2 * tiles.size()) != hipSuccess) {
            (void)hipFree(dt);
            return fail(NUC3D_ERR_OOM, "residual partials");
        }
        HIPCHK(hipMemset(dr, 0, sizeof(double) * 2 * tiles.size()));
        (void)hipFree(c->dresid);
        c->dresid = dr;
        c->work_block.clear();
        for (const auto &t : tiles) c->work_block.push_back(t.b);
        c->nwork = (int)tiles.size();
    }
    (void)hipFree(c->dtiles);
    c->dtiles = dt;
    c->ntiles = (int)tiles.size();
    c->kseg = kseg;
    c->kmax = *std::max_element(kseg.begin(), kseg.end());
    for (auto &g : c->gexec)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    c->gcfl[0] = c->gcfl[1] = -1.0;
    return NUC3D_OK;
}

// whether direction d may take 32-face tiles
bool wide_tiles_ok(const nuc3d_case *c, int d)
{
    if (c->p.scheme[d] == NUC3D_HYBRID) return d < 2;  // the hybrid's x and y sweeps
    if (c->p.scheme[d] != NUC3D_WENO5JS && c->p.scheme[d] != NUC3D_WENO5Z) return false;
    const bool cd8 = c->K.diff[0] || c->K.diff[1] || c->K.diff[2];
    return !(d == 2 && cd8 && c->p.model == NUC3D_NAVIERSTOKES3D);
}

// Tile shape of the WENO5-JS / WENO5-Z direction sweeps (and of the hybrid's
// x and y sweeps) per block shape and direction: 16 faces x 16 lines or 32
// faces x 8 lines (fewer halo cells staged per face, half-width rows in the
// transverse directions).  Each shape's candidates are timed in the context
// of the whole case (the other blocks keep their choice) on the direction's
// own launches: the x and y sweeps as in an RHS evaluation (the hybrid's with
// its stored decisions, whose positivity fallback on this state is the one
// the next stage makes), the z sweep as in stage 0 (into the stage's output
// buffer, which the next stage 0 writes again; the fix-up and positivity
// counts are restored).  The last nuc3d_rhs result is overwritten.
int tune_sweeps(nuc3d_case *c, hipStream_t s)
{
    const bool ns = c->p.model == NUC3D_NAVIERSTOKES3D;
    // the hybrid's sweeps are timed once its sensor is set
    auto tunable = [c](int d) { return wide_tiles_ok(c, d) && (c->p.scheme[d] != NUC3D_HYBRID || c->K.sensor); };
    bool any = false;
    for (int d = 0; d < 3; d++) any = any || tunable(d);
    if (!any) return NUC3D_OK;
    unsigned long long saved[4];
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipMemcpy(saved, c->dcount, sizeof saved, hipMemcpyDeviceToHost));
    if (ns) HIPCHK(c->L->vflux(c->db, c->dtiles, c->ntiles, c->K, c->cur, s));  // the z sweep reads fv
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    unsigned long long *lcur = c->dlam + 3 * c->nb * c->lampar, *lnext = c->dlam + 3 * c->nb * (1 - c->lampar);
    int rc = NUC3D_OK;
    auto run = [&](int d) {
        const int mode = d == 2 ? 1 : 0;
        hipError_t e = hipSuccess;
        for (int g = 0; g < 2 && e == hipSuccess; g++)
            if (c->nsweep[d][g])
                e = c->L->sweep(d, c->p.scheme[d], ns, mode, g ? 32 : 16, c->db, c->dsweep[d][g], c->nsweep[d][g],
                                c->K, 0, c->cur, 0, 1, c->ddt, lcur, lnext, c->dcount, s);
        return e;
    };
    for (int d = 0; d < 3 && rc == 0; d++) {
        if (!tunable(d)) continue;
        std::vector<int> done(c->nb, 0);
        for (int b = 0; b < c->nb && rc == 0; b++) {
            if (done[b]) continue;
            std::vector<int> same;
            for (int b2 = b; b2 < c->nb; b2++)
                if (c->hb[b2].n[0] == c->hb[b].n[0] && c->hb[b2].n[1] == c->hb[b].n[1] &&
                    c->hb[b2].n[2] == c->hb[b].n[2])
                    same.push_back(b2);
            float best = 1e30f;
            int bestnf = c->swnf[b][d];
            for (int nf : {16, 32}) {
                for (int b2 : same) c->swnf[b2][d] = nf;
                if ((rc = build_sweeps(c))) break;
                hipError_t e = run(d);
                if (e == hipSuccess) e = hipEventRecord(e0, s);
                for (int r = 0; r < 3 && e == hipSuccess; r++) e = run(d);
                if (e == hipSuccess) e = hipEventRecord(e1, s);
                if (e == hipSuccess) e = hipEventSynchronize(e1);
                float ms = 0.0f;
                if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
                if (e != hipSuccess) {
                    rc = hipfail(e, "sweep tile tuning");
                    break;
                }
                if (ms < best) {
                    best = ms;
                    bestnf = nf;
                }
            }
            for (int b2 : same) {
                c->swnf[b2][d] = bestnf;
                done[b2] = 1;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if ((rc = build_sweeps(c))) return rc;
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipMemcpy(c->dcount, saved, sizeof saved, hipMemcpyHostToDevice));
    return NUC3D_OK;
}

}  // namespace

extern "C" {

int nuc3d_autotune(nuc3d_case *c, void *stream, int *kseg_out)
{
    if (!c) return fail(NUC3D_ERR_BAD_ARG, "null case");
    if (c->next_stage != 0) return fail(NUC3D_ERR_ORDER, "autotune between full steps");
    hipStream_t s = static_cast<hipStream_t>(stream);
    int rc = refresh(c, s);
    if (rc) return rc;
    std::vector<int> kseg = c->kseg;
    if (c->p.model == NUC3D_NAVIERSTOKES3D) {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        std::vector<int> done(c->nb, 0);
        for (int b = 0; b < c->nb && rc == 0; b++) {
            if (done[b]) continue;
            const nuc3d_block_dev &B = c->hb[b];
            // every block of this shape takes the candidate, the others keep
            // their current choice, and the whole case's passes are timed: the
            // choice is made in the context of the blocks it runs with
            std::vector<int> same;
            for (int b2 = b; b2 < c->nb; b2++) {
                const nuc3d_block_dev &B2 = c->hb[b2];
                if (B2.n[0] == B.n[0] && B2.n[1] == B.n[1] && B2.n[2] == B.n[2]) same.push_back(b2);
            }
            std::vector<int> all(c->nb);
            for (int b2 = 0; b2 < c->nb; b2++) all[b2] = b2;
            float best = 1e30f;
            int bestk = kseg[b];
            for (int ks : {4, 8, 16, 32, 64, 128}) {
                if (ks > 2 * B.n[2] && ks > 8) break;
                std::vector<int> trial = kseg;
                for (int b2 : same) trial[b2] = ks;
                const std::vector<nuc3d_tile> tiles = column_tiles(c, all, trial);
                nuc3d_tile *dt = nullptr;
                if (hipMalloc(&dt, sizeof(nuc3d_tile) * tiles.size()) != hipSuccess) {
                    rc = fail(NUC3D_ERR_OOM, "tile table");
                    break;
                }
                hipError_t e = hipMemcpy(dt, tiles.data(), sizeof(nuc3d_tile) * tiles.size(), hipMemcpyHostToDevice);
                const int nt = (int)tiles.size();
                // the column-tile passes of one stage over the block: viscous
                // fluxes, and where the z sweep does not take it the viscous
                // divergence (into the RHS: no state change)
                const bool div_pass = !visc_in_z_sweep(c);
                auto pass = [&]() {
                    hipError_t r = c->L->vflux(c->db, dt, nt, c->K, c->cur, s);
                    if (r == hipSuccess && div_pass)
                        r = c->L->vdiv(0, c->db, dt, nt, c->K, 0, c->cur, 0, 0, c->ddt, c->dlam, c->dlam, c->dcount, s);
                    return r;
                };
                if (e == hipSuccess) e = pass();
                if (e == hipSuccess) e = hipEventRecord(e0, s);
                for (int r = 0; r < 3 && e == hipSuccess; r++) e = pass();
                if (e == hipSuccess) e = hipEventRecord(e1, s);
                if (e == hipSuccess) e = hipEventSynchronize(e1);
                float ms = 0.0f;
                if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
                (void)hipFree(dt);
                if (e != hipSuccess) {
                    rc = hipfail(e, "autotune");
                    break;
                }
                if (ms < best) {
                    best = ms;
                    bestk = ks;
                }
            }
            for (int b2 : same) {
                kseg[b2] = bestk;
                done[b2] = 1;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (rc) return rc;
        if ((rc = set_column_tiles(c, kseg))) return rc;
    }
    if ((rc = tune_sweeps(c, s))) return rc;
    if (kseg_out)
        for (int b = 0; b < c->nb; b++) kseg_out[b] = c->kseg[b];
    return NUC3D_OK;
}

int nuc3d_set_sweep_tiles(nuc3d_case *c, int b, const int *nf3)
{
    int rc = check_block(c, b);
    if (rc) return rc;
    if (!nf3) return fail(NUC3D_ERR_BAD_ARG, "null argument");
    if (c->next_stage != 0) return fail(NUC3D_ERR_ORDER, "tile shapes change between full steps");
    for (int d = 0; d < 3; d++) {
        if (nf3[d] != 16 && nf3[d] != 32) return fail(NUC3D_ERR_BAD_ARG, "faces per tile line must be 16 or 32");
        if (nf3[d] == 32 && !wide_tiles_ok(c, d))
            return fail(NUC3D_ERR_UNSUPPORTED, "32-face tiles are built for the WENO5-JS and WENO5-Z sweeps and the "
                                               "hybrid's x and y sweeps only (and not for the viscous z sweep of a "
                                               "case with cd8 directions)");
    }
    for (int d = 0; d < 3; d++) c->swnf[b][d] = nf3[d];
    return build_sweeps(c);
}

int nuc3d_get_sweep_tiles(nuc3d_case *c, int b, int *nf3)
{
    int rc = check_block(c, b);
    if (rc) return rc;
    if (!nf3) return fail(NUC3D_ERR_BAD_ARG, "null output");
    for (int d = 0; d < 3; d++) nf3[d] = c->swnf[b][d];
    return NUC3D_OK;
}

// ---------------------------------------------------------------------------
// Learned sensor of the hybrid reconstruction (no counterpart in the reference)

}  // extern "C"

struct nuc3d_sensor {
    long long cap = 0, count = 0;
    unsigned *dX = nullptr;        // [cap][16] bf16 feature pairs
    float *dY = nullptr;           // [cap] labels 1 / 0 / -1
    float *dparams = nullptr, *dmom = nullptr, *dvel = nullptr;  // [NPARAM]
    float *dgrad = nullptr;        // [NPARAM + 1], the loss last
    float *dpartial = nullptr;     // per-workgroup partial gradients
    long long npartial = 0;
    unsigned long long *dpos = nullptr;  // positive labels in the dataset
    int t = 0;                     // Adam steps taken
};

namespace {

constexpr int NPAR = NUC3D_SENSOR_NPARAM;

unsigned short host_bf16(float x)
{
    unsigned u;
    std::memcpy(&u, &x, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

float host_bf16_float(unsigned short b)
{
    const unsigned u = (unsigned)b << 16;
    float x;
    std::memcpy(&x, &u, 4);
    return x;
}

int sensor_partials(nuc3d_sensor *s, long long nsamp)
{
    const long long nwg = (nsamp + 127) / 128;
    if (nwg <= s->npartial) return NUC3D_OK;
    (void)hipFree(s->dpartial);
    s->dpartial = nullptr;
    s->npartial = 0;
    if (hipMalloc(&s->dpartial, sizeof(float) * (NPAR + 1) * nwg) != hipSuccess)
        return fail(NUC3D_ERR_OOM, "device allocation of sensor partials failed");
    s->npartial = nwg;
    return NUC3D_OK;
}

int check_range(nuc3d_sensor *s, long long first, long long n)
{
    if (!s) return fail(NUC3D_ERR_BAD_ARG, "null sensor");
    if (first < 0 || n < 1 || first + n > s->count) return fail(NUC3D_ERR_BAD_ARG, "sample range outside the dataset");
    return NUC3D_OK;
}

}  // namespace

extern "C" {

int nuc3d_sensor_create(long long capacity, nuc3d_sensor **out)
{
    if (!out) return fail(NUC3D_ERR_BAD_ARG, "null argument");
    *out = nullptr;
    if (capacity < 1) return fail(NUC3D_ERR_BAD_ARG, "capacity must be positive");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(NUC3D_ERR_NO_DEVICE, "no HIP device");
    nuc3d_sensor *s = new nuc3d_sensor();
    s->cap = capacity;
    if (hipMalloc(&s->dX, sizeof(unsigned) * 16 * capacity) != hipSuccess ||
        hipMalloc(&s->dY, sizeof(float) * capacity) != hipSuccess ||
        hipMalloc(&s->dparams, sizeof(float) * NPAR) != hipSuccess ||
        hipMalloc(&s->dmom, sizeof(float) * NPAR) != hipSuccess ||
        hipMalloc(&s->dvel, sizeof(float) * NPAR) != hipSuccess ||
        hipMalloc(&s->dgrad, sizeof(float) * (NPAR + 1)) != hipSuccess ||
        hipMalloc(&s->dpos, sizeof(unsigned long long)) != hipSuccess) {
        nuc3d_sensor_destroy(s);
        return fail(NUC3D_ERR_OOM, "device allocation of the sensor dataset failed");
    }
    if (hipMemset(s->dparams, 0, sizeof(float) * NPAR) != hipSuccess ||
        hipMemset(s->dmom, 0, sizeof(float) * NPAR) != hipSuccess ||
        hipMemset(s->dvel, 0, sizeof(float) * NPAR) != hipSuccess ||
        hipMemset(s->dpos, 0, sizeof(unsigned long long)) != hipSuccess) {
        nuc3d_sensor_destroy(s);
        return fail(NUC3D_ERR_HIP, "initialising the sensor failed");
    }
    *out = s;
    return NUC3D_OK;
}

int nuc3d_sensor_destroy(nuc3d_sensor *s)
{
    if (!s) return NUC3D_OK;
    (void)hipDeviceSynchronize();
    (void)hipFree(s->dX);
    (void)hipFree(s->dY);
    (void)hipFree(s->dparams);
    (void)hipFree(s->dmom);
    (void)hipFree(s->dvel);
    (void)hipFree(s->dgrad);
    (void)hipFree(s->dpartial);
    (void)hipFree(s->dpos);
    delete s;
    return NUC3D_OK;
}

int nuc3d_sensor_set_params(nuc3d_sensor *s, const float *params)
{
    if (!s || !params) return fail(NUC3D_ERR_BAD_ARG, "null argument");
    std::vector<float> p(params, params + NPAR);
    for (int o = 0; o < 32; o++)
        for (int i = NUC3D_SENSOR_NFEAT; i < 32; i++) p[o * 32 + i] = 0.0f;  // padding inputs of W1
    HIPCHK(hipMemcpy(s->dparams, p.data(), sizeof(float) * NPAR, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(s->dmom, 0, sizeof(float) * NPAR));
    HIPCHK(hipMemset(s->dvel, 0, sizeof(float) * NPAR));
    s->t = 0;
    return NUC3D_OK;
}

int nuc3d_sensor_get_params(nuc3d_sensor *s, float *params)
{
    if (!s || !params) return fail(NUC3D_ERR_BAD_ARG, "null argument");
    HIPCHK(hipMemcpy(params, s->dparams, sizeof(float) * NPAR, hipMemcpyDeviceToHost));
    return NUC3D_OK;
}

int nuc3d_sensor_clear(nuc3d_sensor *s)
{
    if (!s) return fail(NUC3D_ERR_BAD_ARG, "null sensor");
    s->count = 0;
    HIPCHK(hipMemset(s->dpos, 0, sizeof(unsigned long long)));
    return NUC3D_OK;
}

int nuc3d_sensor_sample(nuc3d_sensor *s, nuc3d_case *c, double tau, int stride, int offset, long long *added)
{
    if (!s || !c) return fail(NUC3D_ERR_BAD_ARG, "null argument");
    if (stride < 1 || offset < 0 || offset >= stride) return fail(NUC3D_ERR_BAD_ARG, "need stride >= 1, 0 <= offset < stride");
    if (!(tau > 0.0)) return fail(NUC3D_ERR_BAD_ARG, "tau must be positive");
    if (c->next_stage != 0) return fail(NUC3D_ERR_ORDER, "sample between full steps");
    HIPCHK(hipDeviceSynchronize());
    int rc = refresh(c, nullptr);
    if (rc) return rc;
    long long total = 0;
    std::vector<long long> ns;
    for (int b = 0; b < c->nb; b++)
        for (int d = 0; d < 3; d++) {
            const nuc3d_block_dev &B = c->hb[b];
            const long long nf = (long long)(B.n[d] + 1) * B.n[(d + 1) % 3] * B.n[(d + 2) % 3];
            const long long n = offset < nf ? (nf - offset + stride - 1) / stride : 0;
            ns.push_back(n);
            total += n;
        }
    if (s->count + total > s->cap) return fail(NUC3D_ERR_BAD_ARG, "sensor dataset capacity exceeded");
    long long pos = s->count;
    int e = 0;
    for (int b = 0; b < c->nb; b++)
        for (int d = 0; d < 3; d++, e++) {
            HIPCHK(nuc3d_launch_sensor_sample(c->db, b, d, c->cur, c->p.gamma, tau, stride, offset, ns[e],
                                              s->dX + 16 * pos, s->dY + pos, s->dpos, nullptr));
            pos += ns[e];
        }
    HIPCHK(hipDeviceSynchronize());
    s->count = pos;
    if (added) *added = total;
    return NUC3D_OK;
}

int nuc3d_sensor_count(nuc3d_sensor *s, long long *count, long long *positives)
{
    if (!s) return fail(NUC3D_ERR_BAD_ARG, "null sensor");
    if (count) *count = s->count;
    if (positives) {
        unsigned long long v = 0;
        HIPCHK(hipMemcpy(&v, s->dpos, sizeof v, hipMemcpyDeviceToHost));
        *positives = (long long)v;
    }
    return NUC3D_OK;
}

int nuc3d_sensor_upload(nuc3d_sensor *s, const unsigned short *x32, const float *y, long long n)
{
    if (!s || !x32 || !y) return fail(NUC3D_ERR_BAD_ARG, "null argument");
    if (n < 0 || s->count + n > s->cap) return fail(NUC3D_ERR_BAD_ARG, "sensor dataset capacity exceeded");
    HIPCHK(hipMemcpy(s->dX + 16 * s->count, x32, sizeof(unsigned) * 16 * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(s->dY + s->count, y, sizeof(float) * n, hipMemcpyHostToDevice));
    unsigned long long pos = 0;
    HIPCHK(hipMemcpy(&pos, s->dpos, sizeof pos, hipMemcpyDeviceToHost));
    for (long long i = 0; i < n; i++) pos += y[i] > 0.5f;
    HIPCHK(hipMemcpy(s->dpos, &pos, sizeof pos, hipMemcpyHostToDevice));
    s->count += n;
    return NUC3D_OK;
}

int nuc3d_sensor_download(nuc3d_sensor *s, long long first, long long n, unsigned short *x32, float *y)
{
    int rc = check_range(s, first, n);
    if (rc) return rc;
    if (x32) HIPCHK(hipMemcpy(x32, s->dX + 16 * first, sizeof(unsigned) * 16 * n, hipMemcpyDeviceToHost));
    if (y) HIPCHK(hipMemcpy(y, s->dY + first, sizeof(float) * n, hipMemcpyDeviceToHost));
    return NUC3D_OK;
}

int nuc3d_sensor_grad(nuc3d_sensor *s, long long first, long long n, float pos_weight, float *grad, double *loss)
{
    int rc = check_range(s, first, n);
    if (rc) return rc;
    if ((rc = sensor_partials(s, n))) return rc;
    HIPCHK(nuc3d_launch_sensor_fwdbwd(s->dX, s->dY, s->count, first, n, 0, 0ull, 0, s->dparams, pos_weight,
                                      1.0f / (float)n, 1, s->dpartial, s->dgrad, nullptr, nullptr));
    std::vector<float> g(NPAR + 1);
    HIPCHK(hipMemcpy(g.data(), s->dgrad, sizeof(float) * (NPAR + 1), hipMemcpyDeviceToHost));
    if (grad) std::memcpy(grad, g.data(), sizeof(float) * NPAR);
    if (loss) *loss = g[NPAR];
    return NUC3D_OK;
}

int nuc3d_sensor_infer(nuc3d_sensor *s, long long first, long long n, float *logits)
{
    int rc = check_range(s, first, n);
    if (rc) return rc;
    if (!logits) return fail(NUC3D_ERR_BAD_ARG, "null logits");
    OpBuffers ob;
    float *dz = reinterpret_cast<float *>(ob.get((size_t)(n + 1) / 2));
    if (!dz) return fail(NUC3D_ERR_OOM, "device allocation failed");
    HIPCHK(nuc3d_launch_sensor_fwdbwd(s->dX, s->dY, s->count, first, n, 0, 0ull, 0, s->dparams, 1.0f, 1.0f, 0,
                                      nullptr, nullptr, dz, nullptr));
    HIPCHK(hipMemcpy(logits, dz, sizeof(float) * n, hipMemcpyDeviceToHost));
    return NUC3D_OK;
}

int nuc3d_sensor_train(nuc3d_sensor *s, int steps, int batch, float lr, float pos_weight, unsigned long long seed,
                       float *losses)
{
    if (!s) return fail(NUC3D_ERR_BAD_ARG, "null sensor");
    if (steps < 0 || batch < 1 || !(lr > 0.0f)) return fail(NUC3D_ERR_BAD_ARG, "need steps >= 0, batch >= 1, lr > 0");
    if (s->count < 1) return fail(NUC3D_ERR_BAD_ARG, "empty dataset");
    int rc = sensor_partials(s, batch);
    if (rc) return rc;
    OpBuffers ob;
    float *dl = losses ? reinterpret_cast<float *>(ob.get((size_t)(steps + 2) / 2)) : nullptr;
    if (losses && !dl) return fail(NUC3D_ERR_OOM, "device allocation failed");
    for (int i = 0; i < steps; i++) {
        HIPCHK(nuc3d_launch_sensor_fwdbwd(s->dX, s->dY, s->count, 0, batch, 1, seed, s->t, s->dparams, pos_weight,
                                          1.0f / (float)batch, 1, s->dpartial, s->dgrad, nullptr, nullptr));
        s->t++;
        HIPCHK(nuc3d_launch_sensor_adam(s->dparams, s->dmom, s->dvel, s->dgrad, lr, s->t, dl ? dl + i : nullptr,
                                        nullptr));
    }
    HIPCHK(hipDeviceSynchronize());
    if (losses && steps) HIPCHK(hipMemcpy(losses, dl, sizeof(float) * steps, hipMemcpyDeviceToHost));
    return NUC3D_OK;
}

int nuc3d_set_sensor(nuc3d_case *c, const float *params, float threshold)
{
    if (!c || !params) return fail(NUC3D_ERR_BAD_ARG, "null argument");
    nuc3d_sensor_dev h;
    std::memset(&h, 0, sizeof h);
    for (int o = 0; o < 32; o++)
        for (int i = 0; i < 32; i++) {
            h.w1[o * 32 + i] = i < NUC3D_SENSOR_NFEAT ? host_bf16(params[o * 32 + i]) : 0;
            h.w2[o * 32 + i] = host_bf16(params[1056 + o * 32 + i]);
        }
    for (int u = 0; u < 32; u++) {
        h.b1[u] = params[1024 + u];
        h.b2[u] = params[2080 + u];
        h.w3[u] = host_bf16_float(host_bf16(params[2112 + u]));
    }
    h.b3 = params[2144];
    h.thr = threshold;
    // lane-arranged copies (the layout of load_regs_lanes)
    for (int l = 0; l < 64; l++) {
        const int g = l / 16, i = l % 16;
        unsigned short c[8];
        for (int e = 0; e < 8; e++) c[e] = host_bf16(h.w3[e < 4 ? 4 * g + e : 16 + 4 * g + (e - 4)]);
        std::memcpy(&h.lw3[l], c, 16);
        for (int mt = 0; mt < 2; mt++) {
            unsigned short a[8], b[8];
            for (int e = 0; e < 8; e++) {
                a[e] = h.w1[(16 * mt + i) * 32 + 8 * g + e];
                b[e] = h.w2[(16 * mt + i) * 32 + (e < 4 ? 4 * g + e : 16 + 4 * g + (e - 4))];
            }
            std::memcpy(&h.lw1[mt][l], a, 16);
            std::memcpy(&h.lw2[mt][l], b, 16);
            const int u = 16 * mt + 4 * g;
            h.lb1[mt][l] = make_float4(h.b1[u], h.b1[u + 1], h.b1[u + 2], h.b1[u + 3]);
            h.lb2[mt][l] = make_float4(h.b2[u], h.b2[u + 1], h.b2[u + 2], h.b2[u + 3]);
        }
    }
    if (!c->dsensor) {
        void *m = nullptr;
        if (hipMalloc(&m, sizeof h) != hipSuccess) return fail(NUC3D_ERR_OOM, "device allocation of the sensor failed");
        c->dsensor = static_cast<nuc3d_sensor_dev *>(m);
    }
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(c->dsensor, &h, sizeof h, hipMemcpyHostToDevice));
    c->K.sensor = c->dsensor;
    c->hforce = true;
    return NUC3D_OK;
}

int nuc3d_set_division(nuc3d_case *c, int ieee)
{
    if (!c) return fail(NUC3D_ERR_BAD_ARG, "null case");
    c->K.ieee_div = ieee != 0;
    // the split's divisions by J follow the same switch
    for (int b = 0; b < c->nb; b++) c->hb[b].rdsplit = c->hb[b].uniform && c->rdok[b] && !c->K.ieee_div;
    HIPCHK(hipMemcpy(c->db, c->hb.data(), sizeof(nuc3d_block_dev) * c->nb, hipMemcpyHostToDevice));
    for (auto &g : c->gexec)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    c->gcfl[0] = c->gcfl[1] = -1.0;
    return NUC3D_OK;
}

int nuc3d_set_overlap(nuc3d_case *c, int enable)
{
    if (!c) return fail(NUC3D_ERR_BAD_ARG, "null case");
    if (c->next_stage != 0) return fail(NUC3D_ERR_ORDER, "switch the overlap between full steps");
    c->overlap = enable != 0;
    for (auto &g : c->gexec)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    c->gcfl[0] = c->gcfl[1] = -1.0;
    return NUC3D_OK;
}

int nuc3d_set_persistent(nuc3d_case *c, int enable)
{
    if (!c) return fail(NUC3D_ERR_BAD_ARG, "null case");
    if (c->next_stage != 0) return fail(NUC3D_ERR_ORDER, "switch persistent sweeps between full steps");
    c->persistent = enable != 0;
    for (auto &g : c->gexec)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    c->gcfl[0] = c->gcfl[1] = -1.0;
    return NUC3D_OK;
}

int nuc3d_set_uniform_mesh(nuc3d_case *c, int enable)
{
    if (!c) return fail(NUC3D_ERR_BAD_ARG, "null case");
    if (c->next_stage != 0) return fail(NUC3D_ERR_ORDER, "switch the uniform-mesh path between full steps");
    c->uniform_mesh = enable != 0;
    c->dirty = true;
    return NUC3D_OK;
}

int nuc3d_get_uniform_mesh(nuc3d_case *c, int b, int *uniform)
{
    if (!c || !uniform) return fail(NUC3D_ERR_BAD_ARG, "null argument");
    if (b < 0 || b >= c->nb) return fail(NUC3D_ERR_BAD_ARG, "block index out of range");
    int rc = refresh(c, nullptr);
    if (rc) return rc;
    *uniform = c->hb[b].uniform;
    return NUC3D_OK;
}

int nuc3d_set_sensor_interval(nuc3d_case *c, int steps)
{
    if (!c) return fail(NUC3D_ERR_BAD_ARG, "null case");
    if (steps != 1 && steps != 2) return fail(NUC3D_ERR_BAD_ARG, "the decision interval is 1 or 2 steps");
    if (c->next_stage != 0) return fail(NUC3D_ERR_ORDER, "set the interval between full steps");
    c->hybrid_interval = steps;
    c->hforce = true;
    for (auto &g : c->gexec)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    c->gcfl[0] = c->gcfl[1] = -1.0;
    return NUC3D_OK;
}

int nuc3d_download_hybrid_flags(nuc3d_case *c, int b, int d, unsigned char *flags)
{
    int rc = check_block(c, b);
    if (rc) return rc;
    if (d < 0 || d > 2 || !flags) return fail(NUC3D_ERR_BAD_ARG, "bad direction or null flags");
    const nuc3d_block_dev &B = c->hb[b];
    if (!B.hflag[d]) return fail(NUC3D_ERR_BAD_ARG, "direction d does not use the hybrid scheme");
    const size_t nf = (size_t)(B.n[d] + 1) * B.n[(d + 1) % 3] * B.n[(d + 2) % 3];
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(flags, B.hflag[d] + (c->flags_rhs ? nf : 0), nf, hipMemcpyDeviceToHost));
    return NUC3D_OK;
}

int nuc3d_get_hybrid_counts(nuc3d_case *c, long long *counts3)
{
    if (!c || !counts3) return fail(NUC3D_ERR_BAD_ARG, "null argument");
    HIPCHK(hipDeviceSynchronize());
    std::vector<unsigned long long> v(16 * NUC3D_HCOUNT_SLOTS);
    HIPCHK(hipMemcpy(v.data(), c->dhcount, sizeof(unsigned long long) * v.size(), hipMemcpyDeviceToHost));
    for (int i = 0; i < 3; i++) {
        unsigned long long t = 0;
        for (int sl = 0; sl < NUC3D_HCOUNT_SLOTS; sl++) t += v[16 * sl + i];
        counts3[i] = (long long)t;
    }
    return NUC3D_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Operator-level ABI (include/nuc3d_ops.h): one Field operation per call, on
// device copies of the caller's host Fields.

extern "C" hipError_t nuc3d_launch_op_interp(int, int, int, int, int, int, int, int, int, const double *,
                                             const double *, const double *, double *, hipStream_t);
extern "C" hipError_t nuc3d_launch_op_crweno(int, int, int, int, const double *, double *, double *, hipStream_t);
extern "C" hipError_t nuc3d_launch_op_diff(int, int, int, int, int, int, int, const double *, const double *,
                                           const double *, double *, hipStream_t);
extern "C" hipError_t nuc3d_launch_op_rk3(int, double, long long, const double *, const double *, double *,
                                          hipStream_t);

namespace {


int op_device_ready()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return fail(NUC3D_ERR_NO_DEVICE, "no GPU visible");
    return NUC3D_OK;
}

}  // namespace

extern "C" {

int nuc3d_op_interpolate(int scheme, int part, int upwind, int w, int zquirk, int nx, int ny, int nz,
                         const double *in, const double *ghost, double *out, void *stream)
{
    if (!in || !out) return fail(NUC3D_ERR_BAD_ARG, "null Field");
    if (scheme < NUC3D_WENO5JS || scheme > NUC3D_RECON_BOUNDARY) return fail(NUC3D_ERR_BAD_ARG, "unknown scheme");
    if (part < NUC3D_PART_INNER || part > NUC3D_PART_RIGHT) return fail(NUC3D_ERR_BAD_ARG, "unknown part");
    if (upwind != 1 && upwind != -1) return fail(NUC3D_ERR_BAD_ARG, "upwind must be +1 or -1");
    if (w < 2 || w > 4) return fail(NUC3D_ERR_UNSUPPORTED, "ghost width must be 2..4");
    if (nx < 2 * w + 1 || ny < 1 || nz < 1) return fail(NUC3D_ERR_BAD_ARG, "line shorter than 2w+1 or empty Field");
    if (scheme == NUC3D_RECON_BOUNDARY && part == NUC3D_PART_INNER)
        return fail(NUC3D_ERR_BAD_ARG, "the physical-face closure has no inner part");
    // the five-point stencils of faces w.. read cells w-3..: the reference runs
    // them with a ghost width of 3 (fieldOperator.cpp:150-160)
    if (scheme != NUC3D_RECON_BOUNDARY && w < 3)
        return fail(NUC3D_ERR_UNSUPPORTED, "five-point reconstructions need a ghost width of 3 or more");
    if (part != NUC3D_PART_INNER && !ghost) return fail(NUC3D_ERR_BAD_ARG, "null ghost Field");
    if (int r = op_device_ready()) return r;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const long long nl = (long long)ny * nz;
    OpBuffers B;
    double *din = B.get(nl * nx), *dout = B.get(nl * (nx + 1));
    double *dg = part == NUC3D_PART_INNER ? nullptr : B.get(nl * w);
    if (!din || !dout || (part != NUC3D_PART_INNER && !dg)) return fail(NUC3D_ERR_OOM, "device allocation failed");
    HIPCHK(hipMemcpyAsync(din, in, sizeof(double) * nl * nx, hipMemcpyHostToDevice, s));
    if (dg) HIPCHK(hipMemcpyAsync(dg, ghost, sizeof(double) * nl * w, hipMemcpyHostToDevice, s));
    const double *gl = part == NUC3D_PART_LEFT ? dg : nullptr, *gr = part == NUC3D_PART_RIGHT ? dg : nullptr;
    const int plus = upwind > 0;
    int m0, nf;
    if (part == NUC3D_PART_INNER) {
        m0 = w;
        nf = nx - 2 * w + 1;
        if (scheme == NUC3D_CRWENO5) {
            double *scr = B.get(nl * 2 * nf);
            if (!scr) return fail(NUC3D_ERR_OOM, "device allocation failed");
            HIPCHK(nuc3d_launch_op_crweno(plus, nx, w, (int)nl, din, dout, scr, s));
        } else {
            HIPCHK(nuc3d_launch_op_interp(scheme, 0, plus, zquirk != 0 && plus && scheme == NUC3D_WENO5Z, nx, w,
                                          (int)nl, m0, nf, din, gl, gr, dout, s));
        }
    } else {
        m0 = part == NUC3D_PART_LEFT ? 0 : nx - w + 1;
        nf = w;
        // halo faces of every scheme use the WENO5 face formula of that scheme
        // (crweno5.cpp:424-836 is WENO5-JS; weno5z.cpp:195-640 the textbook
        // beta2); kind 1 / 2 is the physical-face closure
        const int kind = scheme == NUC3D_RECON_BOUNDARY ? part : 0;
        const int sch = scheme == NUC3D_WENO5Z ? NUC3D_WENO5Z : NUC3D_WENO5JS;
        HIPCHK(nuc3d_launch_op_interp(sch, kind, plus, 0, nx, w, (int)nl, m0, nf, din, gl, gr, dout, s));
    }
    HIPCHK(hipMemcpy2DAsync(out + m0, sizeof(double) * (nx + 1), dout + m0, sizeof(double) * (nx + 1),
                            sizeof(double) * nf, nl, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return NUC3D_OK;
}

int nuc3d_op_differentiate(int scheme, int part, int w, int nx, int ny, int nz, const double *in,
                           const double *ghost, double *out, void *stream)
{
    if (!in || !out) return fail(NUC3D_ERR_BAD_ARG, "null Field");
    if (scheme < NUC3D_CD6 || scheme > NUC3D_DIFF_BOUNDARY) return fail(NUC3D_ERR_BAD_ARG, "unknown scheme");
    if (part < NUC3D_PART_INNER || part > NUC3D_PART_RIGHT) return fail(NUC3D_ERR_BAD_ARG, "unknown part");
    if (w < 2 || w > 4) return fail(NUC3D_ERR_UNSUPPORTED, "ghost width must be 2..4");
    if (scheme == NUC3D_CD6 && w < 3) return fail(NUC3D_ERR_UNSUPPORTED, "cd6 needs a ghost width of 3 or more");
    // the reference's cd8 reads a fourth ghost layer (centraldifference8th.cpp:49-52)
    if (scheme == NUC3D_CD8 && w < 4) return fail(NUC3D_ERR_UNSUPPORTED, "cd8 needs a ghost width of 4");
    if (nx < 2 * w + 1 || ny < 1 || nz < 1) return fail(NUC3D_ERR_BAD_ARG, "line shorter than 2w+1 or empty Field");
    if (scheme == NUC3D_DIFF_BOUNDARY && part == NUC3D_PART_INNER)
        return fail(NUC3D_ERR_BAD_ARG, "the physical-face closure has no inner part");
    const bool needs_ghost = part != NUC3D_PART_INNER && scheme != NUC3D_DIFF_BOUNDARY;
    if (needs_ghost && !ghost) return fail(NUC3D_ERR_BAD_ARG, "null ghost Field");
    if (int r = op_device_ready()) return r;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const long long nl = (long long)ny * nz;
    OpBuffers B;
    double *din = B.get(nl * nx), *dout = B.get(nl * nx);
    double *dg = needs_ghost ? B.get(nl * w) : nullptr;
    if (!din || !dout || (needs_ghost && !dg)) return fail(NUC3D_ERR_OOM, "device allocation failed");
    HIPCHK(hipMemcpyAsync(din, in, sizeof(double) * nl * nx, hipMemcpyHostToDevice, s));
    if (dg) HIPCHK(hipMemcpyAsync(dg, ghost, sizeof(double) * nl * w, hipMemcpyHostToDevice, s));
    const double *gl = part == NUC3D_PART_LEFT ? dg : nullptr, *gr = part == NUC3D_PART_RIGHT ? dg : nullptr;
    int t0, nt;
    if (part == NUC3D_PART_INNER) {
        t0 = w;
        nt = nx - 2 * w;
    } else {
        t0 = part == NUC3D_PART_LEFT ? 0 : nx - w;
        nt = w;
    }
    const int kind = scheme == NUC3D_DIFF_BOUNDARY ? part : 0;
    if (nt > 0) {
        HIPCHK(nuc3d_launch_op_diff(scheme == NUC3D_CD8, kind, nx, w, (int)nl, t0, nt, din, gl, gr, dout, s));
        HIPCHK(hipMemcpy2DAsync(out + t0, sizeof(double) * nx, dout + t0, sizeof(double) * nx, sizeof(double) * nt,
                                nl, hipMemcpyDeviceToHost, s));
    }
    HIPCHK(hipStreamSynchronize(s));
    return NUC3D_OK;
}

int nuc3d_op_rk3(int stage, double dt, int nfields, long long n, const double *const *rhs,
                 const double *const *u_n, double *const *u_i, void *stream)
{
    if (stage < 0 || stage > 2) return fail(NUC3D_ERR_BAD_ARG, "stage must be 0, 1 or 2");
    if (nfields < 1 || n < 1 || !rhs || !u_i || (stage > 0 && !u_n)) return fail(NUC3D_ERR_BAD_ARG, "null field");
    for (int f = 0; f < nfields; f++)
        if (!rhs[f] || !u_i[f] || (stage > 0 && !u_n[f])) return fail(NUC3D_ERR_BAD_ARG, "null field");
    if (int r = op_device_ready()) return r;
    hipStream_t s = static_cast<hipStream_t>(stream);
    OpBuffers B;
    double *dr = B.get(n), *du = B.get(n), *dn = stage > 0 ? B.get(n) : nullptr;
    if (!dr || !du || (stage > 0 && !dn)) return fail(NUC3D_ERR_OOM, "device allocation failed");
    for (int f = 0; f < nfields; f++) {
        HIPCHK(hipMemcpyAsync(dr, rhs[f], sizeof(double) * n, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(du, u_i[f], sizeof(double) * n, hipMemcpyHostToDevice, s));
        if (dn) HIPCHK(hipMemcpyAsync(dn, u_n[f], sizeof(double) * n, hipMemcpyHostToDevice, s));
        HIPCHK(nuc3d_launch_op_rk3(stage, dt, n, dr, dn, du, s));
        HIPCHK(hipMemcpyAsync(u_i[f], du, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    return NUC3D_OK;
}

}  // extern "C"
