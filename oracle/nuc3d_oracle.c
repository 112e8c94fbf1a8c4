/*
 * nuc3d_oracle.c -- CPU restatement of NUC3d's per-block RHS + TVD-RK3 path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Nothing in the product (nuc3d_amd/)
 * links, loads or calls this code.  It is compiled with -ffp-contract=off
 * -fno-builtin so that, like the reference's -O0 build (src/makefile:34-35),
 * every operation is an unfused IEEE double operation in the reference's
 * order and pow()/sqrt() come from libm.
 *
 * Each function cites the reference file:line it restates (paths relative to
 * the reference's src/ directory).
 */
#include "nuc3d_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- constants
 * schemes.hpp:17-44 */
static const double C_W5A[3][3] = {
    {1.0 / 3.0, -7.0 / 6.0, 11.0 / 6.0},
    {-1.0 / 6.0, 5.0 / 6.0, 1.0 / 3.0},
    {1.0 / 3.0, 5.0 / 6.0, -1.0 / 6.0}};
static const double C_W5C[3] = {0.1, 0.6, 0.3};
static const double C_CRC[3] = {0.2, 0.5, 0.3};
static const double C_G0 = 13.0 / 12.0;
static const double C_G1 = 1.0 / 4.0;
static const double C_CD8[4] = {4.0 / 5.0, -1.0 / 5.0, 4.0 / 105.0, -1.0 / 280.0};
static const double C_CD6[3] = {3.0 / 4.0, -3.0 / 20.0, 1.0 / 60.0};
static const double C_CD4[2] = {2.0 / 3.0, -1.0 / 12.0};
static const double C_CD2 = 0.5;
static const double W_SS = 1.0e-6; /* weno5js.cpp:7, weno5z.cpp:6, crweno5.cpp:13 */
static const double W_P = 2.0;     /* weno5js.cpp:8 (runtime member p) */

/* ------------------------------------------------------------- squares
 * The reference squares with pow(): with a literal exponent 2 / 2.0
 * (smoothness indicators, closures, AUSM polynomials, residual) and with the
 * runtime member p = 2 (the WENO weight denominators, weno5js.cpp:97-99,
 * weno5z.cpp:96-98, crweno5.cpp:113-115).  Which calls reach libm depends on
 * how the reference is built, so the restatement has three conventions:
 *   ORC_SQ_LIBM  (0, default): every square from libm pow() -- the
 *                makefile's -O0 build (src/makefile:34-35);
 *   ORC_SQ_O2    (1): literal-exponent squares as x*x, runtime-p squares
 *                from libm pow() -- the same sources built with -O2, where
 *                the compiler folds pow(x, 2.0) into x*x;
 *   ORC_SQ_EXACT (2): every square as x*x, the correctly rounded square.
 * glibc's pow(x, 2.0) differs from the correctly rounded x*x by one ulp for
 * about 0.1% of arguments. */
static int g_sqmode = ORC_SQ_LIBM;

void orc_set_square_mode(int mode) { g_sqmode = mode; }
int orc_get_square_mode(void) { return g_sqmode; }

static double sql(double x) { return g_sqmode == ORC_SQ_LIBM ? pow(x, 2.0) : x * x; }
static double sqp(double x) { return g_sqmode == ORC_SQ_EXACT ? x * x : pow(x, W_P); }

/* ------------------------------------------------------------- WENO pieces */

/* Smoothness indicators of weno5js.cpp:75-82 on f[0..4] (already mirrored for
 * the minus side).  zq=1 reproduces the +inner beta2 of weno5z.cpp:80. */
static void smooth5(const double *f, int zq, double *is0, double *is1, double *is2)
{
    *is0 = C_G0 * sql((f[0] - 2.0 * f[1] + f[2])) + C_G1 * sql((f[0] - 4.0 * f[1] + 3.0 * f[2]));
    *is1 = C_G0 * sql((f[1] - 2.0 * f[2] + f[3])) + C_G1 * sql((f[1] - f[3]));
    if (zq)
        *is2 = C_G0 * sql((f[2] - 2.0 * f[3] + f[3])) + C_G1 * sql((3.0 * f[2] - 4.0 * f[3] + f[4]));
    else
        *is2 = C_G0 * sql((f[2] - 2.0 * f[3] + f[4])) + C_G1 * sql((3.0 * f[2] - 4.0 * f[3] + f[4]));
}

static void cand5(const double *f, double *q0, double *q1, double *q2)
{
    *q0 = C_W5A[0][0] * f[0] + C_W5A[0][1] * f[1] + C_W5A[0][2] * f[2];
    *q1 = C_W5A[1][0] * f[1] + C_W5A[1][1] * f[2] + C_W5A[1][2] * f[3];
    *q2 = C_W5A[2][0] * f[2] + C_W5A[2][1] * f[3] + C_W5A[2][2] * f[4];
}

/* WENO5-JS at one face: weno5js.cpp:75-108 */
static double weno5js_face(const double *f)
{
    double is0, is1, is2, q0, q1, q2;
    smooth5(f, 0, &is0, &is1, &is2);
    cand5(f, &q0, &q1, &q2);
    double a0 = C_W5C[0] / sqp((W_SS + is0));
    double a1 = C_W5C[1] / sqp((W_SS + is1));
    double a2 = C_W5C[2] / sqp((W_SS + is2));
    double as = a0 + a1 + a2;
    double o0 = a0 / as, o1 = a1 / as, o2 = a2 / as;
    return o0 * q0 + o1 * q1 + o2 * q2;
}

/* WENO5-Z at one face: weno5z.cpp:74-109 (zq: quirk of :80) */
static double weno5z_face(const double *f, int zq)
{
    double is0, is1, is2, q0, q1, q2;
    smooth5(f, zq, &is0, &is1, &is2);
    cand5(f, &q0, &q1, &q2);
    double tau = fabs(is2 - is0);
    double a0 = C_W5C[0] * (1.0 + sqp(tau / (is0 + W_SS)));
    double a1 = C_W5C[1] * (1.0 + sqp(tau / (is1 + W_SS)));
    double a2 = C_W5C[2] * (1.0 + sqp(tau / (is2 + W_SS)));
    double as = a0 + a1 + a2;
    double o0 = a0 / as, o1 = a1 / as, o2 = a2 / as;
    return o0 * q0 + o1 * q1 + o2 * q2;
}

/* Gather the 5-cell stencil of face m (between cells m-1 and m) from ext
 * (ext[W+i] = cell i).  upwind +1: cells m-3..m+1 (weno5js.cpp:72-73,255-267);
 * upwind -1: cells m+2..m-2 mirrored (weno5js.cpp:149-172,353-366). */
static void stencil5(const double *ext, int W, int m, int upwind, double *f)
{
    int z;
    if (upwind > 0)
        for (z = 0; z < 5; z++) f[z] = ext[W + m - 3 + z];
    else
        for (z = 0; z < 5; z++) f[z] = ext[W + m + 2 - z];
}

/* Eighth-order central face value from cells m-4..m+3 (c[0..7]): the face
 * point value of the degree-7 polynomial with the eight cell averages, the
 * division by 840 taken as a product with its FP64 reciprocal.  Not in the
 * reference; it is the smooth-region scheme of the learned hybrid. */
static double central8_face(const double *c)
{
    return ((533.0 * (c[3] + c[4]) - 139.0 * (c[2] + c[5])) + (29.0 * (c[1] + c[6]) - 3.0 * (c[0] + c[7]))) *
           (1.0 / 840.0);
}

/* ReconstructionboundaryScheme stencils: reconstructionBoundaryScheme.cpp:322-343 */
static double third_order(double f0, double f1, double f2)
{
    double is0 = sql((f0 - f1));
    double is1 = sql((f1 - f2));
    double q20 = -0.5 * f0 + 1.5 * f1;
    double q21 = 0.5 * f1 + 0.5 * f2;
    double aa0 = 1.0 / 3.0 / sql(1.0e-6 + is0);
    double aa1 = 2.0 / 3.0 / sql(1.0e-6 + is1);
    double w0 = aa0 / (aa0 + aa1);
    double w1 = aa1 / (aa0 + aa1);
    return w0 * q20 + w1 * q21;
}

/* fifthOrderP/N (reconstructionBoundaryScheme.cpp:345-388,425-474): the
 * WENO5-JS formula with the same constants, its weight denominators squared
 * with a literal power of 2.0 (weno5js.cpp:97-99 raise them to the runtime
 * member p instead). */
static double fifth_order(const double *f)
{
    double is0, is1, is2, q0, q1, q2;
    smooth5(f, 0, &is0, &is1, &is2);
    cand5(f, &q0, &q1, &q2);
    double a0 = C_W5C[0] / sql((1.0e-6 + is0));
    double a1 = C_W5C[1] / sql((1.0e-6 + is1));
    double a2 = C_W5C[2] / sql((1.0e-6 + is2));
    double as = a0 + a1 + a2;
    double o0 = a0 / as, o1 = a1 / as, o2 = a2 / as;
    return o0 * q0 + o1 * q1 + o2 * q2;
}

/* Exterior (physical-face) closure, reconstructionBoundaryScheme.cpp:52-312.
 * Writes faces 0..w-1 (left) or n-w+1..n (right). */
static void recon_exterior(int left, int upwind, int n, int w, int W, const double *ext, double *out)
{
    const double *c = ext + W; /* c[i] = cell i */
    double f[5];
    int i, z;
    for (i = 0; i < w; i++) {
        if (left && upwind > 0) { /* boundary_pL: pf = {ghost[-1], &c[0], &c[0], &c[0]} */
            switch (i) {
            case 0: out[0] = ext[W - 1]; break;
            case 1: out[1] = 0.5 * (c[0] + c[1]); break;
            case 2: out[2] = third_order(c[0], c[1], c[2]); break;
            default:
                for (z = 0; z < 5; z++) f[z] = c[z];
                out[3] = fifth_order(f);
            }
        } else if (left) { /* boundary_nL: pf[i] = &c[i] */
            switch (i) {
            case 0: out[0] = c[0]; break;
            case 1: out[1] = 0.5 * (c[1] + c[0]); break;
            case 2: out[2] = third_order(c[3], c[2], c[1]); break;
            default:
                for (z = 0; z < 5; z++) f[z] = c[5 - z];
                out[3] = fifth_order(f);
            }
        } else if (upwind > 0) { /* boundary_pR: pf = {&c[n-1], &c[n-2], &c[n-4], &c[n-6]} */
            switch (i) {
            case 0: out[n] = c[n - 1]; break;
            case 1: out[n - 1] = 0.5 * (c[n - 2] + c[n - 1]); break;
            case 2: out[n - 2] = third_order(c[n - 4], c[n - 3], c[n - 2]); break;
            default:
                for (z = 0; z < 5; z++) f[z] = c[n - 6 + z];
                out[n - 3] = fifth_order(f);
            }
        } else { /* boundary_nR: pf = {ghost[n], &c[n-1], &c[n-2], &c[n-3]} */
            switch (i) {
            case 0: out[n] = ext[W + n]; break;
            case 1: out[n - 1] = 0.5 * (c[n - 1] + c[n - 2]); break;
            case 2: out[n - 2] = third_order(c[n - 1], c[n - 2], c[n - 3]); break;
            default:
                for (z = 0; z < 5; z++) f[z] = c[n - 1 - z];
                out[n - 3] = fifth_order(f);
            }
        }
    }
}

/* CRWENO5 interior system and Thomas solve: crweno5.cpp:44-222 (+) and
 * :224-401 (-).  Rows r=0..M-1 are faces w..n-w; rows 0 and M-1 are explicit
 * WENO5-JS. */
static void crweno5_inner(int upwind, int n, int w, int W, const double *ext, double *out,
                          double *a, double *b, double *c, double *d)
{
    const double *p = ext + W;
    int M = n - 2 * w + 1, r;
    double f[5];
    for (r = 0; r < M; r++) {
        int i = w - 1 + r; /* reference loop index; face i+1 */
        if (r == 0 || r == M - 1) {
            stencil5(ext, W, i + 1, upwind, f);
            a[r] = 0.0;
            b[r] = 1.0;
            c[r] = 0.0;
            d[r] = weno5js_face(f);
            continue;
        }
        double is0, is1, is2;
        stencil5(ext, W, i + 1, upwind, f);
        smooth5(f, 0, &is0, &is1, &is2);
        double a0 = C_CRC[0] / sqp((W_SS + is0));
        double a1 = C_CRC[1] / sqp((W_SS + is1));
        double a2 = C_CRC[2] / sqp((W_SS + is2));
        double as = a0 + a1 + a2;
        double o0 = a0 / as, o1 = a1 / as, o2 = a2 / as;
        if (upwind > 0) { /* crweno5.cpp:155-160 */
            a[r] = 2.0 / 3.0 * o0 + 1.0 / 3.0 * o1;
            b[r] = 1.0 / 3.0 * o0 + 2.0 / 3.0 * (o1 + o2);
            c[r] = 1.0 / 3.0 * o2;
            d[r] = 1.0 / 6.0 * o0 * p[i - 1] + (5.0 * (o0 + o1) + o2) * p[i] / 6.0 + (o1 + 5.0 * o2) * p[i + 1] / 6.0;
        } else { /* crweno5.cpp:334-339 */
            c[r] = 2.0 / 3.0 * o0 + 1.0 / 3.0 * o1;
            b[r] = 1.0 / 3.0 * o0 + 2.0 / 3.0 * (o1 + o2);
            a[r] = 1.0 / 3.0 * o2;
            d[r] = 1.0 / 6.0 * o0 * p[i + 2] + (5.0 * (o0 + o1) + o2) * p[i + 1] / 6.0 + (o1 + 5.0 * o2) * p[i] / 6.0;
        }
    }
    /* Thomas: crweno5.cpp:207-218 */
    for (r = 1; r < M; r++) {
        c[r] = c[r] / (b[r] - c[r - 1] * a[r]);
        d[r] = (d[r] - d[r - 1] * a[r]) / (b[r] - c[r - 1] * a[r]);
    }
    out[n - w] = d[M - 1];
    for (r = M - 2; r >= 0; r--) out[r + w] = d[r] - c[r] * out[r + w + 1];
}

/* One line of fieldOperator3d::reconstructionInner + reconstructionBoundary
 * (fieldOperator.cpp:371-413): inner faces w..n-w, then the left and right
 * face groups from the halo (scheme face variants) or the exterior closure. */
static int recon_line_W(int scheme, int upwind, int n, int w, int W, const double *ext,
                        int typeL, int typeR, int zq, double *out, double *work, const unsigned char *mask)
{
    double f[5];
    int m;
    if (n < 2 * w + 1 || w < 2 || w > 4 || W < w) return -1;
    if (scheme == ORC_HYBRID) {
        /* no counterpart in the reference: per face WENO5-JS (mask 1, 2 or no
         * mask) or the eighth-order central value (mask 0); physical faces keep
         * the reference's closure (width 4) */
        if (w != 4) return -1;
        for (m = 0; m <= n; m++) {
            if ((m < w && typeL == -1) || (m > n - w && typeR == -1)) continue;
            if (!mask || mask[m] == 1 || mask[m] == 2) {
                stencil5(ext, W, m, upwind, f);
                out[m] = weno5js_face(f);
            } else {
                out[m] = central8_face(ext + W + m - 4);
            }
        }
        if (typeL == -1) recon_exterior(1, upwind, n, w, W, ext, out);
        if (typeR == -1) recon_exterior(0, upwind, n, w, W, ext, out);
        return 0;
    }
    if (scheme == ORC_CENTRAL8) {
        /* no counterpart in the reference: the symmetric eight-cell face value,
         * the same for both split parts; interior and halo faces alike,
         * physical faces keep the reference's closure (width 4) */
        if (w != 4) return -1;
        for (m = 0; m <= n; m++) {
            if ((m < w && typeL == -1) || (m > n - w && typeR == -1)) continue;
            out[m] = central8_face(ext + W + m - 4);
        }
        if (typeL == -1) recon_exterior(1, upwind, n, w, W, ext, out);
        if (typeR == -1) recon_exterior(0, upwind, n, w, W, ext, out);
        return 0;
    }
    if (scheme == ORC_CRWENO5) {
        crweno5_inner(upwind, n, w, W, ext, out, work, work + n, work + 2 * n, work + 3 * n);
    } else {
        for (m = w; m <= n - w; m++) {
            stencil5(ext, W, m, upwind, f);
            out[m] = scheme == ORC_WENO5Z ? weno5z_face(f, zq && upwind > 0) : weno5js_face(f);
        }
    }
    /* halo face variants are always the textbook formula (weno5js.cpp:212-624,
     * weno5z.cpp:195-640 with correct beta2, crweno5.cpp:424-836 = WENO5-JS) */
    if (typeL == -1) {
        recon_exterior(1, upwind, n, w, W, ext, out);
    } else {
        for (m = 0; m < w; m++) {
            stencil5(ext, W, m, upwind, f);
            out[m] = scheme == ORC_WENO5Z ? weno5z_face(f, 0) : weno5js_face(f);
        }
    }
    if (typeR == -1) {
        recon_exterior(0, upwind, n, w, W, ext, out);
    } else {
        for (m = n - w + 1; m <= n; m++) {
            stencil5(ext, W, m, upwind, f);
            out[m] = scheme == ORC_WENO5Z ? weno5z_face(f, 0) : weno5js_face(f);
        }
    }
    return 0;
}

/* ------------------------------------------------------------- differences */

static double cd_formula(int scheme, const double *c, int i)
{
    if (scheme == ORC_CD8) /* centraldifference8th.cpp:49-52 */
        return C_CD8[3] * (c[i + 4] - c[i - 4]) + C_CD8[2] * (c[i + 3] - c[i - 3]) +
               C_CD8[1] * (c[i + 2] - c[i - 2]) + C_CD8[0] * (c[i + 1] - c[i - 1]);
    /* centraldifference6th.cpp:43-45 */
    return C_CD6[2] * (c[i + 3] - c[i - 3]) + C_CD6[1] * (c[i + 2] - c[i - 2]) + C_CD6[0] * (c[i + 1] - c[i - 1]);
}

/* boundaryScheme.cpp:31-126: exterior derivative closure; k = distance from
 * the physical face (0 = boundary cell). */
static double bnd_diff(const double *c, int i, int k, int left)
{
    switch (k) {
    case 0: return left ? c[i + 1] - c[i] : c[i] - c[i - 1];
    case 1: return C_CD2 * (c[i + 1] - c[i - 1]);
    case 2: return C_CD4[1] * (c[i + 2] - c[i - 2]) + C_CD4[0] * (c[i + 1] - c[i - 1]);
    default: /* sixthOrder with the cd4 leading coefficient, boundaryScheme.cpp:121-126 */
        return C_CD6[2] * (c[i + 3] - c[i - 3]) + C_CD6[1] * (c[i + 2] - c[i - 2]) + C_CD4[0] * (c[i + 1] - c[i - 1]);
    }
}

/* fieldOperator3d::differenceInner + differenceBoundary (fieldOperator.cpp:505-542) */
static int diff_line_W(int scheme, int n, int w, int W, const double *ext, int typeL, int typeR, double *out)
{
    const double *c = ext + W;
    int i;
    if (n < 2 * w || (scheme == ORC_CD8 && w < 4) || w < 3 || W < w) return -1;
    for (i = w; i < n - w; i++) out[i] = cd_formula(scheme, c, i);
    for (i = 0; i < w; i++) out[i] = typeL == -1 ? bnd_diff(c, i, i, 1) : cd_formula(scheme, c, i);
    for (i = n - w; i < n; i++) out[i] = typeR == -1 ? bnd_diff(c, i, n - 1 - i, 0) : cd_formula(scheme, c, i);
    return 0;
}

int orc_recon_line(int scheme, int upwind, int n, int w, const double *ext, int typeL, int typeR, int zquirk,
                   double *out)
{
    double *work = (double *)malloc(sizeof(double) * 4 * (size_t)(n + 1));
    int rc = recon_line_W(scheme, upwind, n, w, w, ext, typeL, typeR, zquirk, out, work, NULL);
    free(work);
    return rc;
}

int orc_diff_line(int scheme, int n, int w, const double *ext, int typeL, int typeR, double *out)
{
    return diff_line_W(scheme, n, w, w, ext, typeL, typeR, out);
}

/* ------------------------------------------------------------- physics */

/* AUSM Mach/pressure splittings: physicsModel.cpp:803-850 (AUSM), :891-938 (AUSM+) */
static double machL(double m) { return fabs(m) < 1.0 ? 0.25 * sql((m + 1.0)) + 0.125 * sql((m * m - 1)) : 0.50 * (m + fabs(m)); }
static double machR(double m) { return fabs(m) < 1.0 ? -0.25 * sql((m - 1.0)) - 0.125 * sql((m * m - 1)) : 0.50 * (m - fabs(m)); }
static double presL(double m, double p) { return fabs(m) < 1.0 ? p * (m + 1.0) / 2.0 : 0.50 * p * (m + fabs(m)) / m; }
static double presR(double m, double p) { return fabs(m) < 1.0 ? -p * (m - 1.0) / 2.0 : 0.5 * p * (m - fabs(m)) / m; }
static double presLp(double m, double p)
{
    return fabs(m) < 1.0 ? p * (0.25 * sql((m + 1.0)) * (2.0 - m) + 0.1875 * m * sql((m * m - 1.0)))
                         : 0.50 * p * (m + fabs(m)) / m;
}
static double presRp(double m, double p)
{
    return fabs(m) < 1.0 ? p * (0.25 * sql(m - 1.0) * (2.0 + m) - 0.1875 * m * sql((m * m - 1.0)))
                         : 0.5 * p * (m - fabs(m)) / m;
}

/* Cell splitter: RiemannLF physicsModel.cpp:940-979, RiemannAUSM :762-801,
 * RiemannAUSMp :852-889.  Q = conserved/J. */
static void split_cell(int riemann, double U0, double al, const double *Q, double p, double mx, double my,
                       double mz, double jac, double *fp, double *fn)
{
    if (riemann == ORC_LF) {
        double flux[5];
        double lam = fabs(U0) + al;
        flux[0] = U0 * Q[0];
        flux[1] = U0 * Q[1] + mx * p / jac;
        flux[2] = U0 * Q[2] + my * p / jac;
        flux[3] = U0 * Q[3] + mz * p / jac;
        flux[4] = U0 * (Q[4] + p / jac);
        for (int c = 0; c < 5; c++) {
            fp[c] = 0.5 * (flux[c] + lam * Q[c]);
            fn[c] = 0.5 * (flux[c] - lam * Q[c]);
        }
        return;
    }
    double mach = U0 / al;
    double mp = machL(mach), mn = machR(mach);
    double pp = riemann == ORC_AUSM ? presL(mach, p) : presLp(mach, p);
    double pn = riemann == ORC_AUSM ? presR(mach, p) : presRp(mach, p);
    fp[0] = mp * al * Q[0];
    fp[1] = mp * al * Q[1] + mx * pp / jac;
    fp[2] = mp * al * Q[2] + my * pp / jac;
    fp[3] = mp * al * Q[3] + mz * pp / jac;
    fp[4] = mp * al * Q[4] + mp * al * p / jac;
    fn[0] = mn * al * Q[0];
    fn[1] = mn * al * Q[1] + mx * pn / jac;
    fn[2] = mn * al * Q[2] + my * pn / jac;
    fn[3] = mn * al * Q[3] + mz * pn / jac;
    fn[4] = mn * al * Q[4] + mn * al * p / jac;
}

/* Point splitters for BC ghosts: physicsModel.cpp:1159-1398 (with EoS BWD
 * :1056-1077 and FWD :981-1054). */
static int split_point(const orc_params *P, const double *prim, double jac, double mx, double my, double mz,
                       double *fl, double *fr)
{
    double rho = prim[0], u = prim[1], v = prim[2], w = prim[3], p = prim[4];
    double g = P->gamma;
    double _rho = rho, _rhou = rho * u, _rhov = rho * v, _rhow = rho * w;
    double _E = p / (g - 1.0) + 0.5 * rho * (u * u + v * v + w * w);
    double _p = (_E - 0.5 * rho * (u * u + v * v + w * w)) * (g - 1.0);
    if (_p < 0.0) return -1; /* physicsModel.cpp:1009-1016 exits */
    double _alpha = sqrt(fabs(g * _p / rho));
    double theta = sqrt(mx * mx + my * my + mz * mz);
    double alpha = _alpha * theta;
    double U0 = (mx * u + my * v + mz * w);
    if (P->riemann == ORC_LF) {
        double flux[5];
        double lam = fabs(U0) + alpha;
        flux[0] = U0 * _rho / jac;
        flux[1] = U0 * _rhou / jac + mx * p / jac;
        flux[2] = U0 * _rhov / jac + my * p / jac;
        flux[3] = U0 * _rhow / jac + mz * p / jac;
        flux[4] = U0 * (_E / jac + p / jac);
        fl[0] = 0.5 * (flux[0] + lam * _rho / jac);
        fl[1] = 0.5 * (flux[1] + lam * _rhou / jac);
        fl[2] = 0.5 * (flux[2] + lam * _rhov / jac);
        fl[3] = 0.5 * (flux[3] + lam * _rhow / jac);
        fl[4] = 0.5 * (flux[4] + lam * _E / jac);
        fr[0] = 0.5 * (flux[0] - lam * _rho / jac);
        fr[1] = 0.5 * (flux[1] - lam * _rhou / jac);
        fr[2] = 0.5 * (flux[2] - lam * _rhov / jac);
        fr[3] = 0.5 * (flux[3] - lam * _rhow / jac);
        fr[4] = 0.5 * (flux[4] - lam * _E / jac);
        return 0;
    }
    double mach = U0 / alpha;
    double mp = machL(mach), mn = machR(mach);
    double pp = P->riemann == ORC_AUSM ? presL(mach, p) : presLp(mach, p);
    double pn = P->riemann == ORC_AUSM ? presR(mach, p) : presRp(mach, p);
    fl[0] = mp * alpha * _rho / jac;
    fl[1] = mp * alpha * _rhou / jac + mx * pp / jac;
    fl[2] = mp * alpha * _rhov / jac + my * pp / jac;
    fl[3] = mp * alpha * _rhow / jac + mz * pp / jac;
    fl[4] = mp * alpha * _E / jac + mp * alpha * p / jac;
    fr[0] = mn * alpha * _rho / jac;
    fr[1] = mn * alpha * _rhou / jac + mx * pn / jac;
    fr[2] = mn * alpha * _rhov / jac + my * pn / jac;
    fr[3] = mn * alpha * _rhow / jac + mz * pn / jac;
    fr[4] = mn * alpha * _E / jac + mn * alpha * p / jac;
    return 0;
}

void orc_split_point(const orc_params *P, const double *prim, double jac, double mx, double my, double mz,
                     double *fl, double *fr)
{
    (void)split_point(P, prim, jac, mx, my, mz, fl, fr);
}

/* TVD-RK3: TVD-RK3.h:9-12, TVD-RK3.cpp:37-137 (rhs enters as -RHS) */
void orc_rk3(int stage, long n, const double *R, const double *un, double *ui, double dt)
{
    static const double A[3][2] = {{1.0, 1.0}, {3.0 / 4.0, 1.0 / 4.0}, {1.0 / 3.0, 2.0 / 3.0}};
    long i;
    for (i = 0; i < n; i++) {
        double rhs = -R[i];
        if (stage == 0)
            ui[i] = A[0][0] * ui[i] + A[0][1] * rhs * dt;
        else
            ui[i] = A[stage][0] * un[i] + A[stage][1] * (rhs * dt + ui[i]);
    }
}

/* ------------------------------------------------------------- blocks */

typedef struct {
    int n[3];
    long N;
    int ftype[6], fid[6];
    double fval[6][5];
    double *jac, *met[9], *dxyz[9];
    double *q[5], *qn[5], *rhs[5];
    double *W[5], *W0[3];
    double *fl[3][5], *fr[3][5];   /* split fluxes, rotated layout of direction d */
    double *rf[3][5];              /* combined reconstructed flux, (n1+1,n2,n3) */
    double *rtmp;                  /* one reconstructed field */
    double *df[3][5];              /* flux differences, rotated */
    double maxeig[3];
    double dt_local;
    double res_last, res_local; /* PDEData3d::setRES: previous rms and residual */
    /* NS */
    double *grad[4][3];            /* d(u,v,w,T)/d(xi,eta,zeta), rotated of that direction */
    double *rot[3];                /* setGrad transposes of one scalar */
    double *miu, *coef;
    double *fv[3][5], *dfv[3][5];
    /* halo slabs per component x face: (bfsize, n2, n3) of the face's direction */
    double *bsend[5][6], *brecv[5][6];
    unsigned char *hmask[3];       /* hybrid decisions per face of direction d, or NULL */
} oblock;

struct orc_case {
    orc_params P;
    int nb, bf;
    oblock *B;
    double dt_global;
    double res_global;
    long fixups;
    long negd, negp; /* con2prim inputs with density <= 0 / NaN pressure */
    double *work; /* line scratch */
    double *ext, *lout;
};

/* rotated dims/index of direction d (euler3d.cpp:75-80,403,417) */
static void rdims(const oblock *b, int d, int *n1, int *n2, int *n3)
{
    *n1 = b->n[d];
    *n2 = b->n[(d + 1) % 3];
    *n3 = b->n[(d + 2) % 3];
}
static long ridx(const oblock *b, int d, int i, int j, int k)
{
    int c[3] = {i, j, k};
    int a1 = c[d], a2 = c[(d + 1) % 3], a3 = c[(d + 2) % 3];
    return a1 + (long)b->n[d] * (a2 + (long)b->n[(d + 1) % 3] * a3);
}
static long xidx(const oblock *b, int i, int j, int k) { return i + (long)b->n[0] * (j + (long)b->n[1] * k); }

static double *dalloc(long n)
{
    double *p = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    return p;
}

int orc_case_create(const orc_params *p, int nblocks, const int *dims, orc_case **out)
{
    int b, d, c, f, maxn = 0;
    orc_case *C;
    if (!p || nblocks < 1 || !dims || !out) return -1;
    for (d = 0; d < 3; d++) {
        if (p->width[d] < 3 || p->width[d] > 4) return -2;
        if (p->diff[d] == ORC_CD8 && p->width[d] < 4) return -3; /* centraldifference8th.cpp:34,49,91 */
    }
    C = (orc_case *)calloc(1, sizeof(orc_case));
    C->P = *p;
    C->nb = nblocks;
    C->bf = p->width[0];
    for (d = 1; d < 3; d++)
        if (p->width[d] > C->bf) C->bf = p->width[d]; /* block.cpp:506-507 */
    C->B = (oblock *)calloc((size_t)nblocks, sizeof(oblock));
    for (b = 0; b < nblocks; b++) {
        oblock *B = &C->B[b];
        for (d = 0; d < 3; d++) {
            B->n[d] = dims[3 * b + d];
            if (B->n[d] < 2 * C->bf + 2) { orc_case_destroy(C); return -4; }
            if (B->n[d] > maxn) maxn = B->n[d];
        }
        B->N = (long)B->n[0] * B->n[1] * B->n[2];
        for (f = 0; f < 6; f++) { B->ftype[f] = -1; B->fid[f] = 1; }
        B->jac = dalloc(B->N);
        for (c = 0; c < 9; c++) { B->met[c] = dalloc(B->N); B->dxyz[c] = dalloc(B->N); }
        for (c = 0; c < 5; c++) {
            B->q[c] = dalloc(B->N); B->qn[c] = dalloc(B->N); B->rhs[c] = dalloc(B->N);
            B->W[c] = dalloc(B->N);
        }
        for (c = 0; c < 3; c++) B->W0[c] = dalloc(B->N);
        for (d = 0; d < 3; d++) {
            int n1, n2, n3;
            rdims(B, d, &n1, &n2, &n3);
            for (c = 0; c < 5; c++) {
                B->fl[d][c] = dalloc(B->N); B->fr[d][c] = dalloc(B->N);
                B->rf[d][c] = dalloc((long)(n1 + 1) * n2 * n3);
                B->df[d][c] = dalloc(B->N);
                if (p->model == ORC_NS) { B->fv[d][c] = dalloc(B->N); B->dfv[d][c] = dalloc(B->N); }
            }
            if (p->model == ORC_NS) {
                for (c = 0; c < 4; c++) B->grad[c][d] = dalloc(B->N);
                B->rot[d] = dalloc(B->N);
            }
        }
        if (p->model == ORC_NS) { B->miu = dalloc(B->N); B->coef = dalloc(B->N); }
        B->rtmp = dalloc((long)(maxn + 1) * maxn * maxn);
        for (c = 0; c < 5; c++)
            for (f = 0; f < 6; f++) {
                int n1, n2, n3;
                rdims(B, f / 2, &n1, &n2, &n3);
                B->bsend[c][f] = dalloc((long)C->bf * n2 * n3);
                B->brecv[c][f] = dalloc((long)C->bf * n2 * n3);
            }
        /* Field's constructor fills 1.0 (field.cpp:45-53): metrics default to identity-ish */
        for (long i = 0; i < B->N; i++) B->jac[i] = 1.0;
    }
    C->work = dalloc(4L * (maxn + 2));
    C->ext = dalloc(maxn + 2L * C->bf);
    C->lout = dalloc(maxn + 1L);
    *out = C;
    return 0;
}

void orc_case_destroy(orc_case *C)
{
    int b, c, d, f;
    if (!C) return;
    for (b = 0; b < C->nb; b++) {
        oblock *B = &C->B[b];
        free(B->jac);
        for (c = 0; c < 9; c++) { free(B->met[c]); free(B->dxyz[c]); }
        for (c = 0; c < 5; c++) { free(B->q[c]); free(B->qn[c]); free(B->rhs[c]); free(B->W[c]); }
        for (c = 0; c < 3; c++) free(B->W0[c]);
        for (d = 0; d < 3; d++) {
            for (c = 0; c < 5; c++) {
                free(B->fl[d][c]); free(B->fr[d][c]); free(B->rf[d][c]); free(B->df[d][c]);
                free(B->fv[d][c]); free(B->dfv[d][c]);
            }
            for (c = 0; c < 4; c++) free(B->grad[c][d]);
            free(B->rot[d]);
        }
        free(B->miu); free(B->coef); free(B->rtmp);
        for (d = 0; d < 3; d++) free(B->hmask[d]);
        for (c = 0; c < 5; c++)
            for (f = 0; f < 6; f++) { free(B->bsend[c][f]); free(B->brecv[c][f]); }
    }
    free(C->B); free(C->work); free(C->ext); free(C->lout);
    free(C);
}

int orc_set_hybrid_mask(orc_case *C, int b, int d, const unsigned char *mask)
{
    if (!C || b < 0 || b >= C->nb || d < 0 || d > 2) return -1;
    oblock *B = &C->B[b];
    int n1, n2, n3;
    rdims(B, d, &n1, &n2, &n3);
    free(B->hmask[d]);
    B->hmask[d] = NULL;
    if (!mask) return 0;
    size_t nf = (size_t)(n1 + 1) * n2 * n3;
    B->hmask[d] = (unsigned char *)malloc(nf);
    if (!B->hmask[d]) return -2;
    memcpy(B->hmask[d], mask, nf);
    return 0;
}

int orc_set_faces(orc_case *C, int b, const int *type6, const int *id6, const double *val30)
{
    int f, c;
    if (!C || b < 0 || b >= C->nb) return -1;
    for (f = 0; f < 6; f++) {
        oblock *B = &C->B[b];
        B->ftype[f] = type6[f];
        B->fid[f] = id6[f];
        if (type6[f] == -1 && (id6[f] < 0 || id6[f] > 3)) return -2;
        if (type6[f] >= C->nb) return -3;
        for (c = 0; c < 5; c++) B->fval[f][c] = val30 ? val30[5 * f + c] : 0.0;
    }
    return 0;
}

int orc_set_metrics(orc_case *C, int b, const double *jac, const double *met, const double *dxyz)
{
    int c;
    if (!C || b < 0 || b >= C->nb) return -1;
    oblock *B = &C->B[b];
    memcpy(B->jac, jac, sizeof(double) * B->N);
    for (c = 0; c < 9; c++) {
        memcpy(B->met[c], met + c * B->N, sizeof(double) * B->N);
        if (dxyz) memcpy(B->dxyz[c], dxyz + c * B->N, sizeof(double) * B->N);
    }
    return 0;
}

int orc_set_q(orc_case *C, int b, const double *q5)
{
    if (!C || b < 0 || b >= C->nb) return -1;
    for (int c = 0; c < 5; c++) memcpy(C->B[b].q[c], q5 + c * C->B[b].N, sizeof(double) * C->B[b].N);
    return 0;
}

int orc_get_q(const orc_case *C, int b, double *q5)
{
    if (!C || b < 0 || b >= C->nb) return -1;
    for (int c = 0; c < 5; c++) memcpy(q5 + c * C->B[b].N, C->B[b].q[c], sizeof(double) * C->B[b].N);
    return 0;
}

int orc_get_rhs(const orc_case *C, int b, double *r5)
{
    if (!C || b < 0 || b >= C->nb) return -1;
    for (int c = 0; c < 5; c++) memcpy(r5 + c * C->B[b].N, C->B[b].rhs[c], sizeof(double) * C->B[b].N);
    return 0;
}

int orc_get_prim(const orc_case *C, int b, double *w8)
{
    if (!C || b < 0 || b >= C->nb) return -1;
    const oblock *B = &C->B[b];
    for (int c = 0; c < 5; c++) memcpy(w8 + c * B->N, B->W[c], sizeof(double) * B->N);
    for (int c = 0; c < 3; c++) memcpy(w8 + (5 + c) * B->N, B->W0[c], sizeof(double) * B->N);
    return 0;
}

double orc_get_dt_local(const orc_case *C, int b) { return C->B[b].dt_local; }
double orc_get_dt_global(const orc_case *C) { return C->dt_global; }
long orc_get_fixups(const orc_case *C) { return C->fixups; }
void orc_get_positivity(const orc_case *C, long *counts2)
{
    counts2[0] = C->negd;
    counts2[1] = C->negp;
}
int orc_get_max_eigen(const orc_case *C, int b, double *e3)
{
    for (int d = 0; d < 3; d++) e3[d] = C->B[b].maxeig[d];
    return 0;
}

/* physicsModel::con2prim, physicsModel.cpp:258-357 with EoSIdealGasFWD :981-1054 */
static int con2prim(orc_case *C, oblock *B)
{
    double g = C->P.gamma, mach = C->P.mach;
    for (long i = 0; i < B->N; i++) {
        double jac = B->jac[i];
        double rho = B->q[0][i] * jac;
        double u = B->q[1][i] / rho * jac;
        double v = B->q[2][i] / rho * jac;
        double w = B->q[3][i] / rho * jac;
        double E = B->q[4][i] * jac;
        double e0 = E - 0.5 * rho * (u * u + v * v + w * w);
        if (e0 < 0.0) {
            E = (pow(rho, g) / ((g * mach * mach) * (g - 1.0)) + 0.5 * rho * (u * u + v * v + w * w));
            B->q[4][i] = E / jac;
            C->fixups++;
        }
        double tmp = (E - 0.5 * rho * (u * u + v * v + w * w)) * (g - 1.0);
        if (tmp < 0.0) return -10; /* physicsModel.cpp:1009-1016 exits */
        /* states the reference carries on with but the GPU path reports
         * (nuc3d_get_positivity): density not positive, pressure NaN */
        if (!(rho > 0.0)) C->negd++;
        if (!(tmp >= 0.0)) C->negp++;
        double p = tmp;
        B->W[0][i] = rho; B->W[1][i] = u; B->W[2][i] = v; B->W[3][i] = w; B->W[4][i] = p;
        B->W0[0][i] = g * mach * mach * p / rho;
        B->W0[1][i] = p / rho / (g - 1.0);
        B->W0[2][i] = sqrt(fabs(g * p / rho));
    }
    return 0;
}

/* physicsModel::RiemannSolver, physicsModel.cpp:420-760 */
static void riemann(orc_case *C, oblock *B)
{
    int d, i, j, k, c;
    for (d = 0; d < 3; d++) {
        double me = 0.0;
        const double *mx = B->met[3 * d], *my = B->met[3 * d + 1], *mz = B->met[3 * d + 2];
        for (k = 0; k < B->n[2]; k++)
            for (j = 0; j < B->n[1]; j++)
                for (i = 0; i < B->n[0]; i++) {
                    long x = xidx(B, i, j, k), r = ridx(B, d, i, j, k);
                    double Q[5], fp[5], fn[5];
                    for (c = 0; c < 5; c++) Q[c] = B->q[c][x];
                    double theta = sqrt(mx[x] * mx[x] + my[x] * my[x] + mz[x] * mz[x]);
                    double U0 = (mx[x] * B->W[1][x] + my[x] * B->W[2][x] + mz[x] * B->W[3][x]);
                    double al = B->W0[2][x] * theta;
                    split_cell(C->P.riemann, U0, al, Q, B->W[4][x], mx[x], my[x], mz[x], B->jac[x], fp, fn);
                    for (c = 0; c < 5; c++) { B->fl[d][c][r] = fp[c]; B->fr[d][c][r] = fn[c]; }
                    double lam = fabs(U0) + al;
                    me = (me < lam) ? lam : me; /* std::max(lam, me) */
                }
        B->maxeig[d] = me;
    }
}

/* boundaryCondition::setBC (boundaryConditions.cpp:95-1086): ghost split
 * fluxes of physical faces into BufferSend (F+) / BufferRecv (F-). */
static int set_bc(orc_case *C, oblock *B)
{
    int f, c;
    const orc_params *P = &C->P;
    int w = C->bf;
    for (f = 0; f < 6; f++) {
        if (B->ftype[f] != -1) continue;
        int d = f / 2, lr = f % 2, id = B->fid[f];
        int n1, n2, n3, a2, a3, g;
        rdims(B, d, &n1, &n2, &n3);
        for (a3 = 0; a3 < n3; a3++)
            for (a2 = 0; a2 < n2; a2++) {
                int layers = (id == ORC_BC_INLET || id == ORC_BC_OUTLET) ? 1 : w;
                for (g = 0; g < layers; g++) {
                    int cellq[3];
                    int t = (id == ORC_BC_INLET || id == ORC_BC_OUTLET) ? lr * (n1 - 1) : lr * (n1 - 1) + (1 - 2 * lr) * g;
                    cellq[d] = t; cellq[(d + 1) % 3] = a2; cellq[(d + 2) % 3] = a3;
                    long x = xidx(B, cellq[0], cellq[1], cellq[2]);
                    double jac = B->jac[x];
                    double mx = B->met[3 * d][x], my = B->met[3 * d + 1][x], mz = B->met[3 * d + 2][x];
                    double q[5], fl[5], fr[5];
                    if (id == ORC_BC_INLET) {
                        for (c = 0; c < 5; c++) q[c] = B->fval[f][c];
                    } else {
                        for (c = 0; c < 5; c++) q[c] = B->W[c][x];
                    }
                    if (id == ORC_BC_WALL) { /* boundaryConditions.cpp:564-567 */
                        q[0] = q[4] * P->gamma * P->mach * P->mach / P->t_wall;
                        q[1] = 0.0; q[2] = 0.0; q[3] = 0.0;
                    } else if (id == ORC_BC_SYMM) { /* :871-888, :971-977, :1061-1067 */
                        double U = B->met[0][x] * q[1] + B->met[1][x] * q[2] + B->met[2][x] * q[3];
                        double V = B->met[3][x] * q[1] + B->met[4][x] * q[2] + B->met[5][x] * q[3];
                        double Wc = B->met[6][x] * q[1] + B->met[7][x] * q[2] + B->met[8][x] * q[3];
                        int m;
                        for (m = 0; m < 3; m++) {
                            double xx = B->dxyz[3 * m][x], xe = B->dxyz[3 * m + 1][x], xz = B->dxyz[3 * m + 2][x];
                            if (d == 0) q[1 + m] = -xx * U + xe * V + xz * Wc;
                            else if (d == 1) q[1 + m] = xx * U - xe * V + xz * Wc;
                            else q[1 + m] = xx * U + xe * V - xz * Wc;
                        }
                    }
                    if (split_point(P, q, jac, mx, my, mz, fl, fr)) return -12;
                    for (c = 0; c < 5; c++) {
                        double *ps = B->bsend[c][f], *pr = B->brecv[c][f];
                        if (id == ORC_BC_INLET || id == ORC_BC_OUTLET) {
                            for (g = 0; g < w; g++) { /* every layer gets the same value, :173-179 */
                                long bi = g + (long)w * (a2 + (long)n2 * a3);
                                ps[bi] = fl[c]; pr[bi] = fr[c];
                            }
                        } else {
                            int gb = (w - 1) * (1 - lr) - (1 - 2 * lr) * g; /* ibuff, :526 */
                            long bi = gb + (long)w * (a2 + (long)n2 * a3);
                            ps[bi] = fl[c]; pr[bi] = fr[c];
                        }
                    }
                }
            }
    }
    return 0;
}

/* bufferData::setBufferSend (bufferData.cpp:56-192) for all non-physical faces
 * of direction d, then the point-to-point copy that MPI performs
 * (MPICommunicator.cpp:55-115): recv[f] of block b <- send[nf] of block nb. */
static int exchange(orc_case *C, int d, int comp, double *(*field)(oblock *, int))
{
    int b, s, w = C->bf;
    for (b = 0; b < C->nb; b++) {
        oblock *B = &C->B[b];
        int n1, n2, n3;
        rdims(B, d, &n1, &n2, &n3);
        const double *F = field(B, d);
        for (s = 0; s < 2; s++) {
            int f = 2 * d + s;
            if (B->ftype[f] == -1) continue;
            double *bf = B->bsend[comp][f];
            for (long a = 0; a < (long)n2 * n3; a++)
                for (int g = 0; g < w; g++) bf[g + w * a] = F[(s ? n1 - w : 0) + g + (long)n1 * a];
        }
    }
    for (b = 0; b < C->nb; b++) {
        oblock *B = &C->B[b];
        int n1, n2, n3;
        rdims(B, d, &n1, &n2, &n3);
        for (s = 0; s < 2; s++) {
            int f = 2 * d + s;
            if (B->ftype[f] == -1) continue;
            oblock *NB = &C->B[B->ftype[f]];
            int nf = B->fid[f];
            int m1, m2, m3;
            if (nf / 2 != d) return -20; /* only same-direction connections */
            rdims(NB, d, &m1, &m2, &m3);
            if (m2 != n2 || m3 != n3) return -21;
            memcpy(B->brecv[comp][f], NB->bsend[comp][nf], sizeof(double) * (size_t)w * n2 * n3);
        }
    }
    return 0;
}

static int cur_comp, cur_sign;
static double *field_flux(oblock *B, int d) { return cur_sign > 0 ? B->fl[d][cur_comp] : B->fr[d][cur_comp]; }
static double *field_rot(oblock *B, int d) { return B->rot[d]; }
static double *field_visc(oblock *B, int d) { return B->fv[d][cur_comp]; }

/* Build ext for line (a2,a3) of a rotated field with the halo slabs of the two faces. */
static void gather_line(const double *F, const double *bl, const double *br, int n1, int n2, int w, int a2, int a3,
                        double *ext)
{
    long base = (long)n1 * (a2 + (long)n2 * a3), bb = (long)w * (a2 + (long)n2 * a3);
    for (int g = 0; g < w; g++) { ext[g] = bl[bb + g]; ext[w + n1 + g] = br[bb + g]; }
    for (int i = 0; i < n1; i++) ext[w + i] = F[base + i];
}

/* EulerData3D::solveInv + setDerivativesInv (euler3d.cpp:132-241,285-366) */
static int solve_inv(orc_case *C)
{
    int d, c, s, b, rc;
    for (s = 0; s < 2; s++)
        for (d = 0; d < 3; d++)
            for (c = 0; c < 5; c++) {
                cur_comp = c;
                cur_sign = s == 0 ? 1 : -1;
                if ((rc = exchange(C, d, c, field_flux))) return rc;
                for (b = 0; b < C->nb; b++) {
                    oblock *B = &C->B[b];
                    int n1, n2, n3, a2, a3, w = C->P.width[d];
                    rdims(B, d, &n1, &n2, &n3);
                    int tL = B->ftype[2 * d] == -1 ? -1 : 0, tR = B->ftype[2 * d + 1] == -1 ? -1 : 0;
                    /* euler3d.cpp:327-328, 359-360: F+ takes BufferSend at physical faces */
                    const double *bl = (s == 0 && tL == -1) ? B->bsend[c][2 * d] : B->brecv[c][2 * d];
                    const double *br = (s == 0 && tR == -1) ? B->bsend[c][2 * d + 1] : B->brecv[c][2 * d + 1];
                    const double *F = s == 0 ? B->fl[d][c] : B->fr[d][c];
                    double *R = s == 0 ? B->rf[d][c] : B->rtmp;
                    for (a3 = 0; a3 < n3; a3++)
                        for (a2 = 0; a2 < n2; a2++) {
                            gather_line(F, bl, br, n1, n2, C->bf, a2, a3, C->ext);
                            const unsigned char *mk =
                                B->hmask[d] ? B->hmask[d] + (long)(n1 + 1) * (a2 + (long)n2 * a3) : NULL;
                            rc = recon_line_W(C->P.scheme[d], cur_sign, n1, w, C->bf, C->ext, tL, tR, C->P.zquirk,
                                              C->lout, C->work, mk);
                            if (rc) return rc;
                            long ob = (long)(n1 + 1) * (a2 + (long)n2 * a3);
                            if (s == 0)
                                for (int m = 0; m <= n1; m++) R[ob + m] = C->lout[m];
                            else /* combineFluxLR: rf = rfp + rfn (euler3d.cpp:45-49) */
                                for (int m = 0; m <= n1; m++) B->rf[d][c][ob + m] = B->rf[d][c][ob + m] + C->lout[m];
                        }
                }
            }
    /* flux differences: euler3d.cpp:163-169 */
    for (b = 0; b < C->nb; b++) {
        oblock *B = &C->B[b];
        for (d = 0; d < 3; d++) {
            int n1, n2, n3;
            rdims(B, d, &n1, &n2, &n3);
            for (c = 0; c < 5; c++)
                for (long a = 0; a < (long)n2 * n3; a++)
                    for (int i = 1; i < n1 + 1; i++) {
                        const double *f = B->rf[d][c] + (n1 + 1) * a;
                        B->df[d][c][n1 * a + i - 1] = f[i] - f[i - 1];
                    }
        }
    }
    return 0;
}

/* difference of a rotated field along its contiguous axis with halos from
 * brecv[comp] (NaiverStokes3d.cpp:122-168, 339-398) */
static int diff_rot(orc_case *C, oblock *B, int d, int comp, const double *F, double *out)
{
    int n1, n2, n3, a2, a3, rc, w = C->P.width[d];
    rdims(B, d, &n1, &n2, &n3);
    int tL = B->ftype[2 * d] == -1 ? -1 : 0, tR = B->ftype[2 * d + 1] == -1 ? -1 : 0;
    for (a3 = 0; a3 < n3; a3++)
        for (a2 = 0; a2 < n2; a2++) {
            gather_line(F, B->brecv[comp][2 * d], B->brecv[comp][2 * d + 1], n1, n2, C->bf, a2, a3, C->ext);
            rc = diff_line_W(C->P.diff[d], n1, w, C->bf, C->ext, tL, tR, C->lout);
            if (rc) return rc;
            long base = (long)n1 * (a2 + (long)n2 * a3);
            for (int i = 0; i < n1; i++) out[base + i] = C->lout[i];
        }
    return 0;
}

/* NaiverStokesData3d::solveVis (NaiverStokes3d.cpp:51-398) */
static int solve_vis(orc_case *C)
{
    int b, d, c, v, rc, i, j, k;
    const orc_params *P = &C->P;
    /* solveGrads: u, v, w, T with buffers 0..3 (NaiverStokes3d.cpp:63-120) */
    for (v = 0; v < 4; v++) {
        for (b = 0; b < C->nb; b++) { /* gradvector::setGrad, gradvector.cpp:22-73 */
            oblock *B = &C->B[b];
            const double *src = v < 3 ? B->W[1 + v] : B->W0[0];
            for (k = 0; k < B->n[2]; k++)
                for (j = 0; j < B->n[1]; j++)
                    for (i = 0; i < B->n[0]; i++)
                        for (d = 0; d < 3; d++) B->rot[d][ridx(B, d, i, j, k)] = src[xidx(B, i, j, k)];
        }
        for (d = 0; d < 3; d++) {
            if ((rc = exchange(C, d, v, field_rot))) return rc;
            for (b = 0; b < C->nb; b++)
                if ((rc = diff_rot(C, &C->B[b], d, v, C->B[b].rot[d], C->B[b].grad[v][d]))) return rc;
        }
    }
    /* solveViscousFlux + sutherland (NaiverStokes3d.cpp:170-326, physicsModel.cpp:199-234) */
    for (b = 0; b < C->nb; b++) {
        oblock *B = &C->B[b];
        for (long x = 0; x < B->N; x++) {
            double T = B->W0[0][x];
            double S = P->t_ref / P->t_inf;
            double miu0 = (1.0 + S) / (T + S) * pow(T, 1.5) / P->reynolds;
            B->miu[x] = miu0;
            B->coef[x] = miu0 / ((P->gamma - 1.0) * P->mach * P->mach * P->prandtl);
        }
        for (k = 0; k < B->n[2]; k++)
            for (j = 0; j < B->n[1]; j++)
                for (i = 0; i < B->n[0]; i++) {
                    long x = xidx(B, i, j, k), e = ridx(B, 1, i, j, k), z = ridx(B, 2, i, j, k);
                    double jac = B->jac[x];
                    double xi_x = B->met[0][x], xi_y = B->met[1][x], xi_z = B->met[2][x];
                    double eta_x = B->met[3][x], eta_y = B->met[4][x], eta_z = B->met[5][x];
                    double zeta_x = B->met[6][x], zeta_y = B->met[7][x], zeta_z = B->met[8][x];
                    double g[4][3];
                    for (v = 0; v < 4; v++) {
                        double a = B->grad[v][0][x], bb = B->grad[v][1][e], cc = B->grad[v][2][z];
                        g[v][0] = a * xi_x + bb * eta_x + cc * zeta_x;
                        g[v][1] = a * xi_y + bb * eta_y + cc * zeta_y;
                        g[v][2] = a * xi_z + bb * eta_z + cc * zeta_z;
                    }
                    double ux = g[0][0], uy = g[0][1], uz = g[0][2];
                    double vx = g[1][0], vy = g[1][1], vz = g[1][2];
                    double wx = g[2][0], wy = g[2][1], wz = g[2][2];
                    double tx = g[3][0], ty = g[3][1], tz = g[3][2];
                    double grad = ux + vy + wz;
                    double miu0 = B->miu[x];
                    double txx = miu0 * (2.0 * ux - 2.0 / 3.0 * grad);
                    double txy = miu0 * (uy + vx - 2.0 / 3.0 * grad);
                    double txz = miu0 * (uz + wx - 2.0 / 3.0 * grad);
                    double tyy = miu0 * (2.0 * vy - 2.0 / 3.0 * grad);
                    double tyz = miu0 * (vz + wy - 2.0 / 3.0 * grad);
                    double tzz = miu0 * (2.0 * wz - 2.0 / 3.0 * grad);
                    double coeff0 = B->coef[x];
                    double ttx = coeff0 * tx, tty = coeff0 * ty, ttz = coeff0 * tz;
                    double u = B->W[1][x], vv = B->W[2][x], w = B->W[3][x];
                    double fv[5] = {0.0, txx, txy, txz, u * txx + vv * txy + w * txz + ttx};
                    double gv[5] = {0.0, txy, tyy, tyz, u * txy + vv * tyy + w * tyz + tty};
                    double hv[5] = {0.0, txz, tyz, tzz, u * txz + vv * tyz + w * tzz + ttz};
                    for (c = 0; c < 5; c++) {
                        B->fv[0][c][x] = (xi_x * fv[c] + xi_y * gv[c] + xi_z * hv[c]) / jac;
                        B->fv[1][c][e] = (eta_x * fv[c] + eta_y * gv[c] + eta_z * hv[c]) / jac;
                        B->fv[2][c][z] = (zeta_x * fv[c] + zeta_y * gv[c] + zeta_z * hv[c]) / jac;
                    }
                }
    }
    /* setDerivativesVis (NaiverStokes3d.cpp:328-398) */
    for (d = 0; d < 3; d++)
        for (c = 0; c < 5; c++) {
            cur_comp = c;
            if ((rc = exchange(C, d, c, field_visc))) return rc;
            for (b = 0; b < C->nb; b++)
                if ((rc = diff_rot(C, &C->B[b], d, c, C->B[b].fv[d][c], C->B[b].dfv[d][c]))) return rc;
        }
    return 0;
}

/* solveRHS (euler3d.cpp:368-435, NaiverStokes3d.cpp:401-513) */
static void solve_rhs(orc_case *C, oblock *B)
{
    int c, i, j, k;
    for (c = 0; c < 5; c++) {
        double *r = B->rhs[c];
        for (long x = 0; x < B->N; x++) r[x] = B->df[0][c][x];
        for (k = 0; k < B->n[2]; k++)
            for (j = 0; j < B->n[1]; j++)
                for (i = 0; i < B->n[0]; i++) r[xidx(B, i, j, k)] += B->df[1][c][ridx(B, 1, i, j, k)];
        for (k = 0; k < B->n[2]; k++)
            for (j = 0; j < B->n[1]; j++)
                for (i = 0; i < B->n[0]; i++) r[xidx(B, i, j, k)] += B->df[2][c][ridx(B, 2, i, j, k)];
        if (C->P.model == ORC_NS) {
            for (long x = 0; x < B->N; x++) r[x] -= B->dfv[0][c][x];
            for (k = 0; k < B->n[2]; k++)
                for (j = 0; j < B->n[1]; j++)
                    for (i = 0; i < B->n[0]; i++) r[xidx(B, i, j, k)] -= B->dfv[1][c][ridx(B, 1, i, j, k)];
            for (k = 0; k < B->n[2]; k++)
                for (j = 0; j < B->n[1]; j++)
                    for (i = 0; i < B->n[0]; i++) r[xidx(B, i, j, k)] -= B->dfv[2][c][ridx(B, 2, i, j, k)];
        }
    }
    /* EulerData3D::getDt, euler3d.cpp:429-435: std::min(std::min(a,b),c) */
    double a = 1.0 / B->maxeig[0], bb = 1.0 / B->maxeig[1], cc = 1.0 / B->maxeig[2];
    double m = (bb < a) ? bb : a;
    B->dt_local = (cc < m) ? cc : m;
}

int orc_rhs(orc_case *C)
{
    int b, rc;
    for (b = 0; b < C->nb; b++)
        if ((rc = con2prim(C, &C->B[b]))) return rc;
    for (b = 0; b < C->nb; b++) riemann(C, &C->B[b]);
    for (b = 0; b < C->nb; b++)
        if ((rc = set_bc(C, &C->B[b]))) return rc;
    if ((rc = solve_inv(C))) return rc;
    if (C->P.model == ORC_NS && (rc = solve_vis(C))) return rc;
    for (b = 0; b < C->nb; b++) solve_rhs(C, &C->B[b]);
    return 0;
}

/* PDEData3d::solve, PDEData3d.cpp:70-88 */
int orc_stage(orc_case *C, int stage, double cfl)
{
    int b, c, rc;
    if ((rc = orc_rhs(C))) return rc;
    if (stage == 0) {
        double m = C->B[0].dt_local;
        for (b = 0; b < C->nb; b++) {
            for (c = 0; c < 5; c++) memcpy(C->B[b].qn[c], C->B[b].q[c], sizeof(double) * C->B[b].N);
            if (C->B[b].dt_local < m) m = C->B[b].dt_local; /* MPI_MIN */
        }
        C->dt_global = m * cfl;
    }
    for (b = 0; b < C->nb; b++)
        for (c = 0; c < 5; c++) orc_rk3(stage, C->B[b].N, C->B[b].rhs[c], C->B[b].qn[c], C->B[b].q[c], C->dt_global);
    if (stage == 2) { /* setRES, PDEData3d.cpp:90-122; the static res_last is per process = per block */
        C->res_global = 0.0;
        for (b = 0; b < C->nb; b++) {
            oblock *B = &C->B[b];
            double res_this = 0.0;
            long x;
            for (c = 0; c < 5; c++)
                for (x = 0; x < B->N; x++) res_this += sql((B->q[c][x] - B->qn[c][x]));
            B->res_local = fabs(sqrt(res_this / (B->N * 5)) - B->res_last) / (C->dt_global * C->dt_global);
            B->res_last = sqrt(res_this / (B->N * 5));
            if (B->res_local > C->res_global) C->res_global = B->res_local; /* MPI_MAX */
        }
    }
    return 0;
}

double orc_get_residual(const orc_case *C, int b) { return b < 0 ? C->res_global : C->B[b].res_local; }

int orc_steps(orc_case *C, int nsteps, double cfl, double *dts)
{
    int s, st, rc;
    for (s = 0; s < nsteps; s++) {
        for (st = 0; st < 3; st++)
            if ((rc = orc_stage(C, st, cfl))) return rc;
        if (dts) dts[s] = C->dt_global;
    }
    return 0;
}

/* ---------------------------------------------------------------------------
 * Cell centres, coordinate derivatives, J and metric terms from mesh nodes:
 * block::getXYZ_center + interpolation_lag (block.cpp:147-235) and
 * getJacobians + interpolation_derlag (block.cpp:237-477).  Cells are grouped
 * in tiles of 4 per direction; within a tile the five nodes 0..4 carry a
 * fourth-degree Lagrange polynomial, evaluated (LAG) and differentiated
 * (DLAG, per unit index) at the cell centres 0.5, 1.5, 2.5, 3.5.  These are
 * the weights of the reference's tables (block.h:36-47); tests check them
 * against the Lagrange formula in exact rational arithmetic. */
static const double LAG[4][5] = {
    {35.0 / 128.0, 35.0 / 32.0, -35.0 / 64.0, 7.0 / 32.0, -5.0 / 128.0},
    {-5.0 / 128.0, 15.0 / 32.0, 45.0 / 64.0, -5.0 / 32.0, 3.0 / 128.0},
    {3.0 / 128.0, -5.0 / 32.0, 45.0 / 64.0, 15.0 / 32.0, -5.0 / 128.0},
    {-5.0 / 128.0, 7.0 / 32.0, -35.0 / 64.0, 35.0 / 32.0, 35.0 / 128.0}};
static const double DLAG[4][5] = {
    {-11.0 / 12.0, 17.0 / 24.0, 3.0 / 8.0, -5.0 / 24.0, 1.0 / 24.0},
    {1.0 / 24.0, -9.0 / 8.0, 9.0 / 8.0, -1.0 / 24.0, 0.0},
    {0.0, 1.0 / 24.0, -9.0 / 8.0, 9.0 / 8.0, -1.0 / 24.0},
    {-1.0 / 24.0, 5.0 / 24.0, -3.0 / 8.0, -17.0 / 24.0, 11.0 / 12.0}};

/* triple weighted sum over the 5x5x5 nodes of a tile, zeta outermost, in
 * the accumulation order of block.cpp:212-234 */
static double tile_sum(const double *f, long sy, long sz, const double *wx, const double *wy, const double *wz)
{
    double vz = 0.0;
    int i, j, k;
    for (k = 0; k <= 4; k++) {
        double ve = 0.0;
        for (j = 0; j <= 4; j++) {
            double vx = 0.0;
            for (i = 0; i <= 4; i++) vx += wx[i] * f[i + j * sy + k * sz];
            ve += wy[j] * vx;
        }
        vz += wz[k] * ve;
    }
    return vz;
}

int orc_mesh_metrics(int nx, int ny, int nz, const double *xn, const double *yn, const double *zn, double *centre,
                     double *dxyz, double *jac, double *met)
{
    const long sy = nx + 1, sz = (long)(nx + 1) * (ny + 1), N = (long)nx * ny * nz;
    const double *node[3] = {xn, yn, zn};
    int i, j, k, c, d;
    if (nx < 4 || ny < 4 || nz < 4 || nx % 4 || ny % 4 || nz % 4) return -1;
    for (k = 0; k < nz; k++)
        for (j = 0; j < ny; j++)
            for (i = 0; i < nx; i++) {
                const long x = i + (long)j * nx + (long)k * nx * ny;
                const long base = (i / 4) * 4 + ((j / 4) * 4) * sy + ((k / 4) * 4) * sz;
                const int li = i % 4, lj = j % 4, lk = k % 4;
                double dv[9];
                for (c = 0; c < 3; c++) {
                    centre[c * N + x] = tile_sum(node[c] + base, sy, sz, LAG[li], LAG[lj], LAG[lk]);
                    for (d = 0; d < 3; d++) {
                        dv[3 * c + d] = tile_sum(node[c] + base, sy, sz, d == 0 ? DLAG[li] : LAG[li],
                                                 d == 1 ? DLAG[lj] : LAG[lj], d == 2 ? DLAG[lk] : LAG[lk]);
                        dxyz[(3 * c + d) * N + x] = dv[3 * c + d];
                    }
                }
                {
                    const double x_xi = dv[0], x_eta = dv[1], x_zeta = dv[2];
                    const double y_xi = dv[3], y_eta = dv[4], y_zeta = dv[5];
                    const double z_xi = dv[6], z_eta = dv[7], z_zeta = dv[8];
                    const double jacob = 1.0 / (x_xi * (y_eta * z_zeta - y_zeta * z_eta) +
                                                x_eta * (y_zeta * z_xi - y_xi * z_zeta) +
                                                x_zeta * (y_xi * z_eta - y_eta * z_xi));
                    met[0 * N + x] = (y_eta * z_zeta - y_zeta * z_eta) * jacob;
                    met[1 * N + x] = (x_zeta * z_eta - x_eta * z_zeta) * jacob;
                    met[2 * N + x] = (x_eta * y_zeta - x_zeta * y_eta) * jacob;
                    met[3 * N + x] = (y_zeta * z_xi - y_xi * z_zeta) * jacob;
                    met[4 * N + x] = (x_xi * z_zeta - x_zeta * z_xi) * jacob;
                    met[5 * N + x] = (x_zeta * y_xi - x_xi * y_zeta) * jacob;
                    met[6 * N + x] = (y_xi * z_eta - y_eta * z_xi) * jacob;
                    met[7 * N + x] = (x_eta * z_xi - x_xi * z_eta) * jacob;
                    met[8 * N + x] = (x_xi * y_eta - x_eta * y_xi) * jacob;
                    if (jacob < 0.0) return -2; /* block.cpp:310-314 */
                    jac[x] = jacob;
                }
            }
    return 0;
}

/* ---------------------------------------------------------------------------
 * Diagnostics of postproc::solveQ (postproc.cpp:743-848) on the current state:
 * physicsModel::getPrim (con2prim, physicsModel.cpp:153-161), velocity
 * gradients by the derivative scheme with halos and closures
 * (postproc.cpp:1794-1966, the same path as NaiverStokes3d.cpp:86-168),
 * vorticity w = (wy-vz, uz-wx, vx-uy), enstrophy density |w|^2/2, Q-criterion,
 * and the sums over all cells of |w|^2/2 and rho|u|^2/2 (per block in cell
 * order, then over blocks: the MPI_SUM over ranks). */
int orc_diagnostics(orc_case *C, double *enstrophy, double *kinetic, int bf, double *fields5)
{
    int b, d, v, rc, i, j, k;
    double ens_glb = 0.0, kin_glb = 0.0;
    for (b = 0; b < C->nb; b++) {
        oblock *B = &C->B[b];
        if (!B->rot[0])
            for (d = 0; d < 3; d++) {
                for (v = 0; v < 4; v++) B->grad[v][d] = dalloc(B->N);
                B->rot[d] = dalloc(B->N);
            }
        if ((rc = con2prim(C, B))) return rc;
    }
    for (v = 0; v < 3; v++) {
        for (b = 0; b < C->nb; b++) {
            oblock *B = &C->B[b];
            for (k = 0; k < B->n[2]; k++)
                for (j = 0; j < B->n[1]; j++)
                    for (i = 0; i < B->n[0]; i++)
                        for (d = 0; d < 3; d++) B->rot[d][ridx(B, d, i, j, k)] = B->W[1 + v][xidx(B, i, j, k)];
        }
        for (d = 0; d < 3; d++) {
            if ((rc = exchange(C, d, v, field_rot))) return rc;
            for (b = 0; b < C->nb; b++)
                if ((rc = diff_rot(C, &C->B[b], d, v, C->B[b].rot[d], C->B[b].grad[v][d]))) return rc;
        }
    }
    for (b = 0; b < C->nb; b++) {
        oblock *B = &C->B[b];
        double ens = 0.0, kin = 0.0;
        for (k = 0; k < B->n[2]; k++)
            for (j = 0; j < B->n[1]; j++)
                for (i = 0; i < B->n[0]; i++) {
                    long x = xidx(B, i, j, k), e = ridx(B, 1, i, j, k), z = ridx(B, 2, i, j, k);
                    double g[3][3];
                    for (v = 0; v < 3; v++) { /* solveGrad_x/_y/_z, postproc.cpp:1824-1951 */
                        double a = B->grad[v][0][x], bb = B->grad[v][1][e], cc = B->grad[v][2][z];
                        g[v][0] = a * B->met[0][x] + bb * B->met[3][x] + cc * B->met[6][x];
                        g[v][1] = a * B->met[1][x] + bb * B->met[4][x] + cc * B->met[7][x];
                        g[v][2] = a * B->met[2][x] + bb * B->met[5][x] + cc * B->met[8][x];
                    }
                    double ux = g[0][0], uy = g[0][1], uz = g[0][2];
                    double vx = g[1][0], vy = g[1][1], vz = g[1][2];
                    double wx = g[2][0], wy = g[2][1], wz = g[2][2];
                    double w0 = wy - vz, w1 = uz - wx, w2 = vx - uy;
                    double q0 = ux * vy + ux * wz + vy * wz - uy * vx - uz * wx - vz * wy;
                    if (b == bf && fields5) {
                        fields5[0 * B->N + x] = w0;
                        fields5[1 * B->N + x] = w1;
                        fields5[2 * B->N + x] = w2;
                        fields5[3 * B->N + x] = (w0 * w0 + w1 * w1 + w2 * w2) * 0.5;
                        fields5[4 * B->N + x] = q0;
                    }
                    ens += 0.5 * (w0 * w0 + w1 * w1 + w2 * w2);
                    kin += 0.5 * B->W[0][x] * (B->W[1][x] * B->W[1][x] + B->W[2][x] * B->W[2][x] + B->W[3][x] * B->W[3][x]);
                }
        ens_glb += ens;
        kin_glb += kin;
    }
    *enstrophy = ens_glb;
    *kinetic = kin_glb;
    return 0;
}
