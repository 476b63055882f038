/*
 * mac_oracle.c -- FP64 CPU restatement of MAESTRO's MAC projection.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links or calls
 * this file: it is the checker used by tests/, by the smoke check and by the
 * CPU-baseline leg of bench.py.  It restates, on whole-domain (single-level)
 * arrays, the algorithm of
 *   Source/macproject.f90   (macproject, divumac, mk_mac_coeffs,
 *                            mult_edge_by_1d_coeff, mkumac)
 *   Source/mac_multigrid.f90:56-67 (solver options handed to FBoxLib)
 *   Source/mg_eps.f90:22-28        (eps_mac, eps_mac_max, bottom eps)
 *   Source/define_bc_tower.f90:296-402 (press_comp elliptic BCs)
 *   Docs/mg/mg.tex:34-375          (FBoxLib cell-centred MG: GSRB smoother,
 *                                   defect, conservative-average restriction,
 *                                   linear prolongation for 2x, bottom solve)
 * FBoxLib itself is absent from the reference tree, so the MG internals here
 * are a restatement of the documented algorithm, not of FBoxLib source.
 *
 * Layout (C order, x fastest):
 *   cell arrays        (nz, ny, nx)              nz = 1 in 2D
 *   rho (1 ghost)      (nz+2gz, ny+2, nx+2)      gz = 1 in 3D, 0 in 2D
 *   x-faces            (nz, ny, nx+1)
 *   y-faces            (nz, ny+1, nx)
 *   z-faces            (nz+1, ny, nx)
 * Global 0-based indices; x-face i lies between cells i-1 and i.
 *
 * Pinning (which checks in tests/ fix each function):
 *   orc_mk_mac_coeffs   golden fixtures from an independent numpy build,
 *                       constant-rho closed form, swap symmetry
 *   orc_scale_faces     golden fixtures; constant-beta0 invariance of U_proj
 *   orc_divumac         test_projection analytic div-free fields (exact 0),
 *                       linear-field closed form, sparse-matrix fixture
 *   orc_apply/residual  periodic Laplacian eigenvalue closed form, symmetry,
 *                       constant null space, sparse -D B G fixture
 *   orc_gsrb            fixed point of the exact solution; zero black residual
 *   restriction/prolong conservation; exactness on linear/constant fields
 *   orc_mg_solve        scipy direct-solve fixtures; residual invariant
 *   cycle shapes V/W/F  the textbook error-propagation operators assembled
 *                       from dense matrices of the single operators
 *   bottom CG/BiCGStab  the textbook first step; the dense direct solution
 *   orc_mkumac          D(U_proj)=macrhs; test_projection exact recovery;
 *                       idempotence
 *   orc_sum             math.fsum (the correctly rounded sum) on vectors with
 *                       heavy cancellation, where plain and pairwise sums miss
 *                       by more than 1e6 ulp
 *   average_coeffs,     a tagged closed form: each fine face or cell carries a
 *   average_alpha       code of the coarse face or cell it belongs to plus a
 *                       sub-face tag, so the coarse mean is known exactly
 * Parity unpinned (only FBoxLib could fix them): iteration counts and
 * iterates, FBoxLib's own F/W cycle shapes and BiCGStab details, the
 * stencil_order = 2 Dirichlet stencil, and whether FBoxLib's coarse
 * coefficient is this face mean (the mean itself is pinned above).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "mac_oracle.h"

/* ---------------------------------------------------------------- helpers */

static inline long cidx(const int *n, int i, int j, int k)
{
    return (long)i + (long)n[0] * ((long)j + (long)n[1] * (long)k);
}
/* face array of direction d: extent n[d]+1 in d */
static inline long fidx(const int *n, int d, int i, int j, int k)
{
    int ex = n[0] + (d == 0), ey = n[1] + (d == 1);
    return (long)i + (long)ex * ((long)j + (long)ey * (long)k);
}
static inline long nfaces(const int *n, int d)
{
    return (long)(n[0] + (d == 0)) * (n[1] + (d == 1)) * (n[2] + (d == 2));
}
static inline long ncells(const int *n) { return (long)n[0] * n[1] * n[2]; }

/* elliptic BC kind for press_comp, define_bc_tower.f90:296-402:
 * PERIODIC -> periodic, OUTLET -> Dirichlet, everything else -> Neumann */
int orc_ell_bc(int phys)
{
    if (phys == ORC_PERIODIC) return ORC_ELL_PER;
    if (phys == ORC_OUTLET) return ORC_ELL_DIR;
    return ORC_ELL_NEU;
}

int orc_is_singular(const orc_grid *g)
{
    for (int d = 0; d < g->dim; ++d)
        for (int s = 0; s < 2; ++s)
            if (orc_ell_bc(g->bc[d][s]) == ORC_ELL_DIR) return 0;
    return 1;
}

/* ------------------------------------------------ a5: mk_mac_coeffs (3d/2d)
 * macproject.f90:698-738 (3d), :675-696 (2d):
 *   betax(i,j,k) = 2 / (rho(i,j,k) + rho(i-1,j,k)), faces lo..hi+1.       */
void orc_mk_mac_coeffs(const orc_grid *g, const double *rho, double *bx,
                       double *by, double *bz)
{
    const int *n = g->n;
    int gz = (g->dim == 3);
    long rx = n[0] + 2, ry = n[1] + 2;
#define RHO(i, j, k) rho[((long)(i) + 1) + rx * (((long)(j) + 1) + ry * ((long)(k) + gz))]
    double *b[3] = {bx, by, bz};
    for (int d = 0; d < g->dim; ++d) {
        int ex = n[0] + (d == 0), ey = n[1] + (d == 1), ez = n[2] + (d == 2);
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 0; k < ez; ++k)
            for (int j = 0; j < ey; ++j)
                for (int i = 0; i < ex; ++i) {
                    int im = i - (d == 0), jm = j - (d == 1), km = k - (d == 2);
                    b[d][fidx(n, d, i, j, k)] = 2.0 / (RHO(i, j, k) + RHO(im, jm, km));
                }
    }
#undef RHO
}

/* ------------------------------------- a2: mult_edge_by_1d_coeff (3d/2d)
 * macproject.f90:577-609 (3d), :547-575 (2d).  The radial direction is the
 * last one (z in 3d, y in 2d).  Faces normal to it use beta0_edge(r),
 * the tangential faces use the cell value beta0(r).  do_mult=1 multiplies,
 * 0 divides.                                                              */
void orc_scale_faces_1d(const orc_grid *g, double *fx, double *fy, double *fz,
                        const double *beta0, const double *beta0_edge, int do_mult)
{
    const int *n = g->n;
    int rd = g->dim - 1;
    double *f[3] = {fx, fy, fz};
    for (int d = 0; d < g->dim; ++d) {
        int ex = n[0] + (d == 0), ey = n[1] + (d == 1), ez = n[2] + (d == 2);
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 0; k < ez; ++k)
            for (int j = 0; j < ey; ++j)
                for (int i = 0; i < ex; ++i) {
                    int r = (rd == 2) ? k : j;
                    double c = (d == rd) ? beta0_edge[r] : beta0[r];
                    long q = fidx(n, d, i, j, k);
                    if (do_mult) f[d][q] = f[d][q] * c;
                    else f[d][q] = f[d][q] / c;
                }
    }
}

/* multifab_mult_mult / multifab_div_div with beta0_cart_edge,
 * macproject.f90:159-165, 194-202, 237-243                               */
void orc_scale_faces_cart(const orc_grid *g, double *fx, double *fy, double *fz,
                          const double *cx, const double *cy, const double *cz,
                          int do_mult)
{
    double *f[3] = {fx, fy, fz};
    const double *c[3] = {cx, cy, cz};
    for (int d = 0; d < g->dim; ++d) {
        long nf = nfaces(g->n, d);
#pragma omp parallel for schedule(static)
        for (long q = 0; q < nf; ++q) {
            if (do_mult) f[d][q] = f[d][q] * c[d][q];
            else f[d][q] = f[d][q] / c[d][q];
        }
    }
}

/* ------------------------------------------------------ a4: divumac
 * macproject.f90:457-480 (3d), :438-455 (2d), sign handling :389-398:
 *   D = (u(i+1)-u(i))/dx + (v(j+1)-v(j))/dy + (w(k+1)-w(k))/dz
 *   rh = macrhs - D   (or rh = -D without macrhs)                          */
void orc_divumac(const orc_grid *g, const double *u, const double *v,
                 const double *w, const double *macrhs, double *rh)
{
    const int *n = g->n;
    const double *dx = g->dx;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                double D = (u[fidx(n, 0, i + 1, j, k)] - u[fidx(n, 0, i, j, k)]) / dx[0] +
                           (v[fidx(n, 1, i, j + 1, k)] - v[fidx(n, 1, i, j, k)]) / dx[1];
                if (g->dim == 3)
                    D = D + (w[fidx(n, 2, i, j, k + 1)] - w[fidx(n, 2, i, j, k)]) / dx[2];
                long c = cidx(n, i, j, k);
                if (macrhs) rh[c] = -(D - macrhs[c]);
                else rh[c] = -D;
            }
}

/* ------------------------------------------------ stencil face evaluation
 * For cell (i,j,k) on a level with extents n and face coefficients b[d]
 * (raw beta, natural layout), returns the face weight w and the neighbour
 * value p such that the 7-point operator is
 *   (A phi)_c = sum_d ih2_d * sum_s w_s (phi_c - p_s).
 * Periodic: wrapped neighbour.  Neumann: w = 0 (no flux through the face).
 * Dirichlet (phi = 0 on the face, two-point): w = 2 beta_f, p = 0.
 * BC meaning: define_bc_tower.f90:338,349,360,371,383,394 (press_comp).   */
typedef struct {
    int n[3];
    double h[3];
    double ih2[3];
    const double *b[3];
    const double *alpha;   /* NULL: alpha = 0 */
} orc_lv;

static inline void face_pair(const orc_lv *L, const int *bc, const double *phi,
                             int i, int j, int k, int d, int side,
                             double *w, double *p)
{
    const int *n = L->n;
    int c[3] = {i, j, k};
    int fc[3] = {i, j, k};
    if (side) fc[d] += 1;             /* face index: lo face of cell c or hi */
    double beta = L->b[d][fidx(n, d, fc[0], fc[1], fc[2])];
    int nb[3] = {i, j, k};
    nb[d] += side ? 1 : -1;
    if (nb[d] >= 0 && nb[d] < n[d]) {
        *w = beta;
        *p = phi[cidx(n, nb[0], nb[1], nb[2])];
        return;
    }
    int kind = orc_ell_bc(bc[2 * d + side]);
    if (kind == ORC_ELL_PER) {
        nb[d] = (nb[d] + n[d]) % n[d];
        *w = beta;
        *p = phi[cidx(n, nb[0], nb[1], nb[2])];
    } else if (kind == ORC_ELL_DIR) {
        *w = 2.0 * beta;
        *p = 0.0;
    } else {
        *w = 0.0;
        *p = 0.0;
    }
    (void)c;
}

static void lv_from_level(const orc_level *lev, int dim, orc_lv *L)
{
    for (int d = 0; d < 3; ++d) {
        L->n[d] = lev->n[d];
        L->h[d] = lev->h[d];
        L->ih2[d] = (d < dim) ? 1.0 / (lev->h[d] * lev->h[d]) : 0.0;
        L->b[d] = lev->b[d];
    }
    L->alpha = lev->alpha;
}

/* residual r = rhs - A phi at one cell, fixed evaluation order (fma);
 * A = alpha - D beta G (the alpha term last, as one fma)                  */
static inline double cell_residual(const orc_lv *L, const int *bc, int dim,
                                   const double *phi, const double *rhs,
                                   int i, int j, int k)
{
    double wm, pm, wp, pp;
    double pc = phi[cidx(L->n, i, j, k)];
    face_pair(L, bc, phi, i, j, k, 0, 0, &wm, &pm);
    face_pair(L, bc, phi, i, j, k, 0, 1, &wp, &pp);
    double ax = fma(wm, pc - pm, wp * (pc - pp));
    face_pair(L, bc, phi, i, j, k, 1, 0, &wm, &pm);
    face_pair(L, bc, phi, i, j, k, 1, 1, &wp, &pp);
    double ay = fma(wm, pc - pm, wp * (pc - pp));
    double Aphi;
    if (dim == 3) {
        face_pair(L, bc, phi, i, j, k, 2, 0, &wm, &pm);
        face_pair(L, bc, phi, i, j, k, 2, 1, &wp, &pp);
        double az = fma(wm, pc - pm, wp * (pc - pp));
        Aphi = fma(L->ih2[0], ax, fma(L->ih2[1], ay, L->ih2[2] * az));
    } else {
        Aphi = fma(L->ih2[0], ax, L->ih2[1] * ay);
    }
    if (L->alpha) Aphi = fma(L->alpha[cidx(L->n, i, j, k)], pc, Aphi);
    return rhs[cidx(L->n, i, j, k)] - Aphi;
}

/* Gauss-Seidel point update, fixed evaluation order (fma); alpha adds to
 * the diagonal last                                                        */
static inline double cell_gs(const orc_lv *L, const int *bc, int dim,
                             const double *phi, const double *rhs,
                             int i, int j, int k)
{
    double wm, pm, wp, pp;
    face_pair(L, bc, phi, i, j, k, 0, 0, &wm, &pm);
    face_pair(L, bc, phi, i, j, k, 0, 1, &wp, &pp);
    double tx = fma(wm, pm, wp * pp), sx = wm + wp;
    face_pair(L, bc, phi, i, j, k, 1, 0, &wm, &pm);
    face_pair(L, bc, phi, i, j, k, 1, 1, &wp, &pp);
    double ty = fma(wm, pm, wp * pp), sy = wm + wp;
    double rhs_c = rhs[cidx(L->n, i, j, k)];
    double num, den;
    if (dim == 3) {
        face_pair(L, bc, phi, i, j, k, 2, 0, &wm, &pm);
        face_pair(L, bc, phi, i, j, k, 2, 1, &wp, &pp);
        double tz = fma(wm, pm, wp * pp), sz = wm + wp;
        num = fma(L->ih2[0], tx, fma(L->ih2[1], ty, fma(L->ih2[2], tz, rhs_c)));
        den = fma(L->ih2[0], sx, fma(L->ih2[1], sy, L->ih2[2] * sz));
    } else {
        num = fma(L->ih2[0], tx, fma(L->ih2[1], ty, rhs_c));
        den = fma(L->ih2[0], sx, L->ih2[1] * sy);
    }
    if (L->alpha) den = den + L->alpha[cidx(L->n, i, j, k)];
    return num / den;
}

/* --------------------------------------------------- operator utilities */

void orc_apply(const orc_grid *g, const orc_level *lev, const double *phi,
               double *out)
{
    orc_lv L;
    lv_from_level(lev, g->dim, &L);
    const int *bc = &g->bc[0][0];
    /* A phi = rhs - residual with rhs = 0 evaluated as -(0 - A phi) */
    long nc = ncells(L.n);
    double *zero = calloc((size_t)nc, sizeof(double));
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < L.n[2]; ++k)
        for (int j = 0; j < L.n[1]; ++j)
            for (int i = 0; i < L.n[0]; ++i)
                out[cidx(L.n, i, j, k)] = -cell_residual(&L, bc, g->dim, phi, zero, i, j, k);
    free(zero);
}

double orc_residual(const orc_grid *g, const orc_level *lev, const double *phi,
                    const double *rhs, double *res)
{
    orc_lv L;
    lv_from_level(lev, g->dim, &L);
    const int *bc = &g->bc[0][0];
    double mx = 0.0;
#pragma omp parallel for collapse(2) schedule(static) reduction(max : mx)
    for (int k = 0; k < L.n[2]; ++k)
        for (int j = 0; j < L.n[1]; ++j)
            for (int i = 0; i < L.n[0]; ++i) {
                double r = cell_residual(&L, bc, g->dim, phi, rhs, i, j, k);
                if (res) res[cidx(L.n, i, j, k)] = r;
                double a = fabs(r);
                if (a > mx) mx = a;
            }
    return mx;
}

/* One red-black Gauss-Seidel sweep (MG_SMOOTHER_GS_RB, mg.tex:41-44,201):
 * even-parity cells (i+j+k even) first, then odd.                         */
void orc_gsrb(const orc_grid *g, orc_level *lev) { orc_gsrb_omega(g, lev, 1.0); }

/* One red-black sweep over-relaxed by omega: each cell moves to
 * old + omega (gs - old), as fma(omega, gs - old, old); omega = 1 stores the
 * Gauss-Seidel value itself.  Over-relaxed red-black smoothing in multigrid:
 * I. Yavneh, "On red-black SOR smoothing in multigrid", SIAM J. Sci. Comput.
 * 17 (1996) 180-192.  No reference counterpart (FBoxLib's GSRB is omega = 1):
 * the smoother of the library's FP32 cycles.                               */
void orc_gsrb_omega(const orc_grid *g, orc_level *lev, double omega)
{
    orc_lv L;
    lv_from_level(lev, g->dim, &L);
    const int *bc = &g->bc[0][0];
    double *phi = lev->phi;
    for (int color = 0; color < 2; ++color) {
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 0; k < L.n[2]; ++k)
            for (int j = 0; j < L.n[1]; ++j)
                for (int i = (j + k + color) & 1; i < L.n[0]; i += 2) {
                    const long c = cidx(L.n, i, j, k);
                    const double v = cell_gs(&L, bc, g->dim, phi, lev->rhs, i, j, k);
                    phi[c] = (omega == 1.0) ? v : fma(omega, v - phi[c], phi[c]);
                }
    }
}

/* the over-relaxation of level L's sweeps in a cycle (ORC_SOR_MAX_CELLS) */
static double level_omega(const orc_mg *mg, const orc_level *L)
{
    const long cells = (long)L->n[0] * L->n[1] * L->n[2];
    const long thr = mg->g.dim == 3 ? ORC_SOR_MAX_CELLS : ORC_SOR_MAX_CELLS_2D;
    return cells > thr ? mg->opts.omega : 1.0;
}

/* ------------------------------------------------ MG transfer operators */

/* Conservative average of fine defects (mg.tex:353-354); fixed sum order */
static void restrict_defect(int dim, const orc_level *f, orc_level *c)
{
    const double *r = f->res;
    double s = (dim == 3) ? 0.125 : 0.25;
#pragma omp parallel for collapse(2) schedule(static)
    for (int K = 0; K < c->n[2]; ++K)
        for (int J = 0; J < c->n[1]; ++J)
            for (int I = 0; I < c->n[0]; ++I) {
                int i = 2 * I, j = 2 * J;
                double v;
                if (dim == 3) {
                    int k = 2 * K;
                    v = ((r[cidx(f->n, i, j, k)] + r[cidx(f->n, i + 1, j, k)]) +
                         (r[cidx(f->n, i, j + 1, k)] + r[cidx(f->n, i + 1, j + 1, k)])) +
                        ((r[cidx(f->n, i, j, k + 1)] + r[cidx(f->n, i + 1, j, k + 1)]) +
                         (r[cidx(f->n, i, j + 1, k + 1)] + r[cidx(f->n, i + 1, j + 1, k + 1)]));
                } else {
                    v = (r[cidx(f->n, i, j, 0)] + r[cidx(f->n, i + 1, j, 0)]) +
                        (r[cidx(f->n, i, j + 1, 0)] + r[cidx(f->n, i + 1, j + 1, 0)]);
                }
                c->rhs[cidx(c->n, I, J, K)] = s * v;
            }
}

/* Coarse alpha = average of the 2^dm fine cells (same sum order as the
 * defect restriction).  FBoxLib's coarsening of cell coefficients is not in
 * the reference tree; averaging is this build's documented choice.        */
static void average_alpha(int dim, const orc_level *f, orc_level *c)
{
    const double *a = f->alpha;
    double s = (dim == 3) ? 0.125 : 0.25;
    for (int K = 0; K < c->n[2]; ++K)
        for (int J = 0; J < c->n[1]; ++J)
            for (int I = 0; I < c->n[0]; ++I) {
                int i = 2 * I, j = 2 * J;
                double v;
                if (dim == 3) {
                    int k = 2 * K;
                    v = ((a[cidx(f->n, i, j, k)] + a[cidx(f->n, i + 1, j, k)]) +
                         (a[cidx(f->n, i, j + 1, k)] + a[cidx(f->n, i + 1, j + 1, k)])) +
                        ((a[cidx(f->n, i, j, k + 1)] + a[cidx(f->n, i + 1, j, k + 1)]) +
                         (a[cidx(f->n, i, j + 1, k + 1)] + a[cidx(f->n, i + 1, j + 1, k + 1)]));
                } else {
                    v = (a[cidx(f->n, i, j, 0)] + a[cidx(f->n, i + 1, j, 0)]) +
                        (a[cidx(f->n, i, j + 1, 0)] + a[cidx(f->n, i + 1, j + 1, 0)]);
                }
                c->alpha[cidx(c->n, I, J, K)] = s * v;
            }
}

/* Coarse face coefficient = average of the 2^(dm-1) fine faces covering it.
 * FBoxLib's coarse-coefficient rule is not in the reference tree; this is
 * the documented choice of this build (fixed sum order).                  */
static void average_coeffs(int dim, const orc_level *f, orc_level *c)
{
    for (int d = 0; d < dim; ++d) {
        int ex = c->n[0] + (d == 0), ey = c->n[1] + (d == 1), ez = c->n[2] + (d == 2);
        const double *fb = f->b[d];
        double *cb = c->b[d];
        int t1 = (d == 0) ? 1 : 0;           /* first tangential direction */
        int t2 = (d == 2) ? 1 : 2;           /* second tangential (3d only) */
#pragma omp parallel for collapse(2) schedule(static)
        for (int K = 0; K < ez; ++K)
            for (int J = 0; J < ey; ++J)
                for (int I = 0; I < ex; ++I) {
                    int C[3] = {I, J, K};
                    int F[3] = {2 * I, 2 * J, 2 * K};
                    (void)C;
                    double v;
                    if (dim == 3) {
                        int a[3], b2[3], c2[3], e2[3];
                        memcpy(a, F, sizeof a); memcpy(b2, F, sizeof a);
                        memcpy(c2, F, sizeof a); memcpy(e2, F, sizeof a);
                        b2[t1] += 1;
                        c2[t2] += 1;
                        e2[t1] += 1; e2[t2] += 1;
                        v = 0.25 * ((fb[fidx(f->n, d, a[0], a[1], a[2])] +
                                     fb[fidx(f->n, d, b2[0], b2[1], b2[2])]) +
                                    (fb[fidx(f->n, d, c2[0], c2[1], c2[2])] +
                                     fb[fidx(f->n, d, e2[0], e2[1], e2[2])]));
                    } else {
                        int a[3], b2[3];
                        memcpy(a, F, sizeof a); memcpy(b2, F, sizeof a);
                        a[2] = 0; b2[2] = 0;
                        b2[t1] += 1;
                        v = 0.5 * (fb[fidx(f->n, d, a[0], a[1], 0)] +
                                   fb[fidx(f->n, d, b2[0], b2[1], 0)]);
                    }
                    cb[fidx(c->n, d, I, J, K)] = v;
                }
    }
}

/* 1-D linear-interpolation weights for fine index fi on a coarse line of
 * length nc: parent P with weight 3/4, neighbour Q with weight 1/4.  Out of
 * the domain the neighbour is the periodic image, the parent itself (Neumann,
 * even reflection of the homogeneous correction) or the parent with weight
 * -1/4 (Dirichlet, odd reflection).                                       */
static inline void lin_w(int fi, int nc, int kind_lo, int kind_hi, int *P,
                         int *Q, double *wp, double *wq)
{
    int p = fi >> 1;
    int q = (fi & 1) ? p + 1 : p - 1;
    *P = p;
    *wp = 0.75;
    *wq = 0.25;
    if (q < 0 || q >= nc) {
        int kind = (q < 0) ? kind_lo : kind_hi;
        if (kind == ORC_ELL_PER) q = (q + nc) % nc;
        else if (kind == ORC_ELL_DIR) { q = p; *wq = -0.25; }
        else q = p;
    }
    *Q = q;
}

/* phi_f += P e_c, linear (bi/trilinear) cell-centred prolongation
 * ("linear reconstruction for jumps by 2x", mg.tex:363-365)               */
static void prolong_add(const orc_grid *g, const orc_level *c, orc_level *f)
{
    int dim = g->dim;
    const double *e = c->phi;
    int kx[2] = {orc_ell_bc(g->bc[0][0]), orc_ell_bc(g->bc[0][1])};
    int ky[2] = {orc_ell_bc(g->bc[1][0]), orc_ell_bc(g->bc[1][1])};
    int kz[2] = {orc_ell_bc(g->bc[2][0]), orc_ell_bc(g->bc[2][1])};
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < f->n[2]; ++k)
        for (int j = 0; j < f->n[1]; ++j)
            for (int i = 0; i < f->n[0]; ++i) {
                int xi[2], yi[2], zi[2];
                double wx[2], wy[2], wz[2];
                lin_w(i, c->n[0], kx[0], kx[1], &xi[0], &xi[1], &wx[0], &wx[1]);
                lin_w(j, c->n[1], ky[0], ky[1], &yi[0], &yi[1], &wy[0], &wy[1]);
                double v;
                if (dim == 3) {
                    lin_w(k, c->n[2], kz[0], kz[1], &zi[0], &zi[1], &wz[0], &wz[1]);
                    double vz[2];
                    for (int cz = 0; cz < 2; ++cz) {
                        double vy[2];
                        for (int cy = 0; cy < 2; ++cy) {
                            const double *row = e + cidx(c->n, 0, yi[cy], zi[cz]);
                            vy[cy] = fma(wx[1], row[xi[1]], wx[0] * row[xi[0]]);
                        }
                        vz[cz] = fma(wy[1], vy[1], wy[0] * vy[0]);
                    }
                    v = fma(wz[1], vz[1], wz[0] * vz[0]);
                } else {
                    double vy[2];
                    for (int cy = 0; cy < 2; ++cy) {
                        const double *row = e + cidx(c->n, 0, yi[cy], 0);
                        vy[cy] = fma(wx[1], row[xi[1]], wx[0] * row[xi[0]]);
                    }
                    v = fma(wy[1], vy[1], wy[0] * vy[0]);
                }
                f->phi[cidx(f->n, i, j, k)] += v;
            }
}

/* ------------------------------------------------ bottom solve (FP64 CG)
 * mg_bottom_solver = 2 (CG), mg.tex:221-262; tolerance bottom_eps
 * (mac_multigrid.f90:61 passes 1e-3).  Singular levels: the mean of the
 * right-hand side is removed first, so CG stays in the range of A.        */

/* compensated (Neumaier) sum, used for every mean in the oracle */
double orc_sum(const double *x, long n)
{
    double s = 0.0, c = 0.0;
    for (long q = 0; q < n; ++q) {
        double t = s + x[q];
        if (fabs(s) >= fabs(x[q])) c += (s - t) + x[q];
        else c += (x[q] - t) + s;
        s = t;
    }
    return s + c;
}

static double dot(const double *a, const double *b, long n)
{
    double s = 0.0;
    for (long q = 0; q < n; ++q) s += a[q] * b[q];
    return s;
}

static void project_mean(double *x, long n)
{
    double m = orc_sum(x, n) / (double)n;
    for (long q = 0; q < n; ++q) x[q] -= m;
}

static void bottom_cg(const orc_grid *g, orc_level *lev, double eps, int maxit,
                      int singular)
{
    long nc = ncells(lev->n);
    if (singular) {
        double m = orc_sum(lev->rhs, nc) / (double)nc;
        for (long q = 0; q < nc; ++q) lev->rhs[q] -= m;
    }
    double *r = malloc(sizeof(double) * nc), *p = malloc(sizeof(double) * nc),
           *Ap = malloc(sizeof(double) * nc);
    double r0 = orc_residual(g, lev, lev->phi, lev->rhs, r);
    if (singular) project_mean(r, nc);
    if (r0 > 0.0) {
        memcpy(p, r, sizeof(double) * nc);
        double rr = dot(r, r, nc);
        for (int it = 0; it < maxit; ++it) {
            orc_apply(g, lev, p, Ap);
            double pAp = dot(p, Ap, nc);
            if (!(pAp > 0.0)) break;
            double alpha = rr / pAp;
            double rmax = 0.0;
            for (long q = 0; q < nc; ++q) {
                lev->phi[q] += alpha * p[q];
                r[q] -= alpha * Ap[q];
            }
            /* keep the Krylov space orthogonal to the constant null space */
            if (singular) project_mean(r, nc);
            for (long q = 0; q < nc; ++q)
                if (fabs(r[q]) > rmax) rmax = fabs(r[q]);
            if (rmax <= eps * r0) break;
            double rr1 = dot(r, r, nc);
            double beta = rr1 / rr;
            rr = rr1;
            for (long q = 0; q < nc; ++q) p[q] = r[q] + beta * p[q];
        }
    }
    free(r); free(p); free(Ap);
}

/* BiCGStab (van der Vorst, SIAM J. Sci. Stat. Comput. 13 (1992) 631-644),
 * FBoxLib's default bottom solver (mg_bottom_solver = 1, mg.tex:242-243).
 * Same start, mean projection and stopping rule as bottom_cg; s = r -
 * alpha v overwrites r.                                                   */
static void bottom_bicgstab(const orc_grid *g, orc_level *lev, double eps, int maxit,
                            int singular)
{
    long nc = ncells(lev->n);
    if (singular) {
        double m = orc_sum(lev->rhs, nc) / (double)nc;
        for (long q = 0; q < nc; ++q) lev->rhs[q] -= m;
    }
    double *r = malloc(sizeof(double) * nc), *rh = malloc(sizeof(double) * nc),
           *p = calloc(nc, sizeof(double)), *v = calloc(nc, sizeof(double)),
           *t = malloc(sizeof(double) * nc);
    double r0 = orc_residual(g, lev, lev->phi, lev->rhs, r);
    if (singular) project_mean(r, nc);
    if (r0 > 0.0) {
        memcpy(rh, r, sizeof(double) * nc);
        double rho_old = 1.0, alpha = 1.0, omega = 1.0;
        for (int it = 0; it < maxit; ++it) {
            double rho = dot(rh, r, nc);
            if (rho == 0.0) break;
            if (it == 0) {
                memcpy(p, r, sizeof(double) * nc);
            } else {
                double beta = (rho / rho_old) * (alpha / omega);
                for (long q = 0; q < nc; ++q) p[q] = r[q] + beta * (p[q] - omega * v[q]);
            }
            orc_apply(g, lev, p, v);
            double rv = dot(rh, v, nc);
            if (rv == 0.0) break;
            alpha = rho / rv;
            double smax = 0.0;
            for (long q = 0; q < nc; ++q) r[q] -= alpha * v[q];
            if (singular) project_mean(r, nc);
            for (long q = 0; q < nc; ++q)
                if (fabs(r[q]) > smax) smax = fabs(r[q]);
            if (smax <= eps * r0) {
                for (long q = 0; q < nc; ++q) lev->phi[q] += alpha * p[q];
                break;
            }
            orc_apply(g, lev, r, t);
            double tt = dot(t, t, nc);
            if (!(tt > 0.0)) {
                for (long q = 0; q < nc; ++q) lev->phi[q] += alpha * p[q];
                break;
            }
            omega = dot(t, r, nc) / tt;
            double rmax = 0.0;
            for (long q = 0; q < nc; ++q) {
                lev->phi[q] += alpha * p[q] + omega * r[q];
                r[q] -= omega * t[q];
            }
            if (singular) project_mean(r, nc);
            for (long q = 0; q < nc; ++q)
                if (fabs(r[q]) > rmax) rmax = fabs(r[q]);
            rho_old = rho;
            if (rmax <= eps * r0 || omega == 0.0) break;
        }
    }
    free(r); free(rh); free(p); free(v); free(t);
}

/* the bottom level: nub sweeps, the chosen solver (none for smoothing
 * only, mg.tex:239-240), nub sweeps                                        */
static void bottom_solve(orc_mg *mg, orc_level *L)
{
    const orc_grid *g = &mg->g;
    for (int s = 0; s < mg->opts.nub; ++s) orc_gsrb(g, L);
    if (mg->opts.bottom_solver == ORC_BOTTOM_CG)
        bottom_cg(g, L, mg->opts.bottom_eps, mg->opts.bottom_maxit, mg->singular);
    else if (mg->opts.bottom_solver == ORC_BOTTOM_BICGSTAB)
        bottom_bicgstab(g, L, mg->opts.bottom_eps, mg->opts.bottom_maxit, mg->singular);
    for (int s = 0; s < mg->opts.nub; ++s) orc_gsrb(g, L);
}

/* -------------------------------------------------------- MG hierarchy */

int orc_mg_nlevels(const orc_grid *g, int max_levels)
{
    int n[3] = {g->n[0], g->n[1], g->n[2]};
    int nl = 1;
    for (;;) {
        if (max_levels > 0 && nl >= max_levels) break;
        int ok = 1;
        for (int d = 0; d < g->dim; ++d)
            if ((n[d] & 1) || n[d] / 2 < 2) ok = 0;
        if (!ok) break;
        for (int d = 0; d < g->dim; ++d) n[d] /= 2;
        ++nl;
    }
    return nl;
}

orc_mg *orc_mg_create(const orc_grid *g, const double *bx, const double *by,
                      const double *bz, const orc_opts *o)
{
    return orc_mg_create_alpha(g, NULL, bx, by, bz, o);
}

/* hierarchy of (alpha - D beta G); alpha NULL or all zero: alpha = 0 */
orc_mg *orc_mg_create_alpha(const orc_grid *g, const double *alpha, const double *bx,
                            const double *by, const double *bz, const orc_opts *o)
{
    orc_mg *mg = calloc(1, sizeof(orc_mg));
    mg->g = *g;
    mg->opts = *o;
    if (alpha) {
        long nc0 = ncells(g->n), q;
        for (q = 0; q < nc0; ++q)
            if (alpha[q] != 0.0) break;
        if (q == nc0) alpha = NULL;
    }
    mg->singular = orc_is_singular(g) && !alpha;
    mg->nlev = orc_mg_nlevels(g, o->max_levels);
    mg->lev = calloc((size_t)mg->nlev, sizeof(orc_level));
    int n[3] = {g->n[0], g->n[1], g->n[2]};
    double h[3] = {g->dx[0], g->dx[1], g->dx[2]};
    for (int l = 0; l < mg->nlev; ++l) {
        orc_level *L = &mg->lev[l];
        for (int d = 0; d < 3; ++d) { L->n[d] = n[d]; L->h[d] = h[d]; }
        long nc = ncells(n);
        L->phi = calloc((size_t)nc, sizeof(double));
        L->rhs = calloc((size_t)nc, sizeof(double));
        L->res = calloc((size_t)nc, sizeof(double));
        for (int d = 0; d < 3; ++d) {
            if (d < g->dim) L->b[d] = calloc((size_t)nfaces(n, d), sizeof(double));
            else L->b[d] = NULL;
        }
        if (alpha) L->alpha = calloc((size_t)nc, sizeof(double));
        if (l == 0) {
            const double *src[3] = {bx, by, bz};
            for (int d = 0; d < g->dim; ++d)
                memcpy(L->b[d], src[d], sizeof(double) * nfaces(n, d));
            if (alpha) memcpy(L->alpha, alpha, sizeof(double) * nc);
        } else {
            average_coeffs(g->dim, &mg->lev[l - 1], L);
            if (alpha) average_alpha(g->dim, &mg->lev[l - 1], L);
        }
        for (int d = 0; d < g->dim; ++d) { n[d] /= 2; h[d] *= 2.0; }
    }
    return mg;
}

void orc_mg_destroy(orc_mg *mg)
{
    if (!mg) return;
    for (int l = 0; l < mg->nlev; ++l) {
        free(mg->lev[l].phi); free(mg->lev[l].rhs); free(mg->lev[l].res); free(mg->lev[l].alpha);
        for (int d = 0; d < 3; ++d) free(mg->lev[l].b[d]);
    }
    free(mg->lev);
    free(mg);
}

/* recursive cycle (mg.tex:342-370): nu1 GSRB sweeps, defect, restrict,
 * coarse solve, prolongate and correct, nu2 sweeps.  The cycle shapes of
 * mg_cycle_type (Source/_parameters:376-377; FBoxLib's cycle routine is not
 * in the reference tree, so these are the textbook shapes of Trottenberg,
 * Oosterlee & Schueller, "Multigrid" (2001), sec. 2.4): a V visits the next
 * coarser level once, a W twice (each visit a W), an F first as an F and
 * then as a V.  A second visit starts from the first one's iterate.       */
static void cycle(orc_mg *mg, int l, int type)
{
    const orc_grid *g = &mg->g;
    orc_level *L = &mg->lev[l];
    if (l == mg->nlev - 1) {
        bottom_solve(mg, L);
        return;
    }
    const double w = level_omega(mg, L);
    for (int s = 0; s < mg->opts.nu1; ++s) orc_gsrb_omega(g, L, w);
    orc_residual(g, L, L->phi, L->rhs, L->res);
    orc_level *C = &mg->lev[l + 1];
    restrict_defect(g->dim, L, C);
    memset(C->phi, 0, sizeof(double) * ncells(C->n));
    cycle(mg, l + 1, type);
    if (type == ORC_CYCLE_W) cycle(mg, l + 1, ORC_CYCLE_W);
    else if (type == ORC_CYCLE_F) cycle(mg, l + 1, ORC_CYCLE_V);
    prolong_add(g, C, L);
    for (int s = 0; s < mg->opts.nu2; ++s) orc_gsrb_omega(g, L, w);
}

static int cycle_shape(const orc_opts *o)
{
    return o->cycle_type == 0 ? ORC_CYCLE_V : o->cycle_type;
}

/* single operators on caller-built levels (used by the kernel tests) */
void orc_level_restrict(const orc_grid *g, const orc_level *f, orc_level *c)
{
    restrict_defect(g->dim, f, c);
}

void orc_level_prolong(const orc_grid *g, const orc_level *c, orc_level *f)
{
    prolong_add(g, c, f);
}

void orc_level_coarsen(const orc_grid *g, const orc_level *f, orc_level *c)
{
    average_coeffs(g->dim, f, c);
    if (f->alpha && c->alpha) average_alpha(g->dim, f, c);
}

/* one cycle (of the shape in the options) on the hierarchy, level 0
 * holding phi/rhs already */
void orc_mg_vcycle(orc_mg *mg) { cycle(mg, 0, cycle_shape(&mg->opts)); }

/* MG solve of A phi = rh on the finest level, phi in/out (warm start).
 * Stops when max|rh - A phi| <= eps_rel * max|rh - A phi0| (the
 * "error/error0" test printed by F90mg, see the reference run log
 * Docs/managing_jobs/scaling/wdconvect/jaguarpf_031910/output.mpi.1728:
 * 177-179,213-215) or <= eps_abs.  Singular problems: rh mean removed.    */
int orc_mg_solve(orc_mg *mg, double *phi, double *rh, orc_stats *st)
{
    const orc_grid *g = &mg->g;
    orc_level *F = &mg->lev[0];
    long nc = ncells(F->n);
    if (mg->opts.cycle_type < 0 || mg->opts.cycle_type > ORC_CYCLE_V) return ORC_ERR_INVALID;
    if (mg->opts.bottom_solver < ORC_BOTTOM_SMOOTH || mg->opts.bottom_solver > ORC_BOTTOM_CG)
        return ORC_ERR_INVALID;
    if (mg->singular) {
        double m = orc_sum(rh, nc) / (double)nc;
        for (long q = 0; q < nc; ++q) rh[q] -= m;
    }
    memcpy(F->phi, phi, sizeof(double) * nc);
    memcpy(F->rhs, rh, sizeof(double) * nc);
    double r0 = orc_residual(g, F, F->phi, F->rhs, NULL);
    double r = r0;
    int it = 0, conv = (r0 == 0.0) || (mg->opts.eps_abs > 0 && r0 <= mg->opts.eps_abs);
    while (!conv && it < mg->opts.max_iter) {
        cycle(mg, 0, cycle_shape(&mg->opts));
        ++it;
        r = orc_residual(g, F, F->phi, F->rhs, NULL);
        if (st && it <= ORC_MAX_HIST) st->hist[it - 1] = r;
        conv = (r <= mg->opts.eps_rel * r0) || (mg->opts.eps_abs > 0 && r <= mg->opts.eps_abs);
    }
    memcpy(phi, F->phi, sizeof(double) * nc);
    if (st) {
        st->iters = it;
        st->res0 = r0;
        st->res = r;
        st->converged = conv;
        st->nlev = mg->nlev;
    }
    return conv ? 0 : ORC_ERR_NOT_CONVERGED;
}

/* ------------------------------------------------------- a7: mkumac
 * macproject.f90:908-975 (3d), :863-906 (2d).  Interior faces:
 *   u -= beta (phi_i - phi_{i-1}) / dx
 * Box-boundary faces use FBoxLib's fine_flx, which equals the same
 * gradient across box and periodic faces, zero at Neumann walls and the
 * Dirichlet face gradient (phi_c - 0)/(dx/2) at outlets.                  */
void orc_mkumac(const orc_grid *g, double *u, double *v, double *w,
                const double *phi, const double *bx, const double *by,
                const double *bz)
{
    const int *n = g->n;
    double *f[3] = {u, v, w};
    const double *b[3] = {bx, by, bz};
    for (int d = 0; d < g->dim; ++d) {
        int ex = n[0] + (d == 0), ey = n[1] + (d == 1), ez = n[2] + (d == 2);
        int klo = orc_ell_bc(g->bc[d][0]), khi = orc_ell_bc(g->bc[d][1]);
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 0; k < ez; ++k)
            for (int j = 0; j < ey; ++j)
                for (int i = 0; i < ex; ++i) {
                    int c[3] = {i, j, k}, m[3] = {i, j, k};
                    m[d] -= 1;
                    long q = fidx(n, d, i, j, k);
                    double beta = b[d][q];
                    double hi_v, lo_v;
                    if (c[d] == n[d] || m[d] < 0) {
                        int kind = (m[d] < 0) ? klo : khi;
                        if (kind == ORC_ELL_NEU) continue;   /* zero flux */
                        if (kind == ORC_ELL_PER) {
                            c[d] = (c[d] + n[d]) % n[d];
                            m[d] = (m[d] + n[d]) % n[d];
                            hi_v = phi[cidx(n, c[0], c[1], c[2])];
                            lo_v = phi[cidx(n, m[0], m[1], m[2])];
                        } else { /* Dirichlet: ghost = -phi_interior */
                            if (m[d] < 0) {
                                hi_v = phi[cidx(n, c[0], c[1], c[2])];
                                lo_v = -hi_v;
                            } else {
                                lo_v = phi[cidx(n, m[0], m[1], m[2])];
                                hi_v = -lo_v;
                            }
                        }
                    } else {
                        hi_v = phi[cidx(n, c[0], c[1], c[2])];
                        lo_v = phi[cidx(n, m[0], m[1], m[2])];
                    }
                    double gphi = (hi_v - lo_v) / g->dx[d];
                    f[d][q] = f[d][q] - beta * gphi;
                }
    }
}

/* ---------------------------------------------------- a1: macproject
 * Steps of macproject.f90:51-328 on whole-domain arrays (single level):
 *   umac <- beta0 umac (:154-165); rh = macrhs - D(umac) (:175-179);
 *   beta = 2/(rho_L+rho_R) (:182); beta *= beta0 (x2 with
 *   use_alt_energy_fix) (:185-203); rel eps = min(eps_mac*10^(nlev-1),
 *   eps_mac_max) (:209-217); solve (:224-226); umac -= beta G phi (:229);
 *   umac <- umac / beta0 (:232-243).                                       */
int orc_macproject(const orc_grid *g, double *u, double *v, double *w,
                   double *phi, const double *rho, const double *macrhs,
                   const double *beta0, const double *beta0_edge,
                   const double *cart_x, const double *cart_y,
                   const double *cart_z, const orc_opts *o, orc_stats *st,
                   double *bx_out, double *by_out, double *bz_out,
                   double *rh_out)
{
    const int *n = g->n;
    int use_1d = (beta0 != NULL), use_cart = (cart_x != NULL);
    if (use_1d && use_cart) return ORC_ERR_INVALID;
    if (use_1d && !beta0_edge) return ORC_ERR_INVALID;
    long nc = ncells(n);
    long nf[3] = {nfaces(n, 0), nfaces(n, 1), g->dim == 3 ? nfaces(n, 2) : 0};
    double *rh = malloc(sizeof(double) * nc);
    double *b[3] = {NULL, NULL, NULL};
    for (int d = 0; d < g->dim; ++d) b[d] = malloc(sizeof(double) * nf[d]);

    if (use_1d) orc_scale_faces_1d(g, u, v, w, beta0, beta0_edge, 1);
    else if (use_cart) orc_scale_faces_cart(g, u, v, w, cart_x, cart_y, cart_z, 1);

    orc_divumac(g, u, v, w, macrhs, rh);
    orc_mk_mac_coeffs(g, rho, b[0], b[1], b[2]);
    if (use_1d) {
        orc_scale_faces_1d(g, b[0], b[1], b[2], beta0, beta0_edge, 1);
        if (o->use_alt_energy_fix) orc_scale_faces_1d(g, b[0], b[1], b[2], beta0, beta0_edge, 1);
    } else if (use_cart) {
        orc_scale_faces_cart(g, b[0], b[1], b[2], cart_x, cart_y, cart_z, 1);
        if (o->use_alt_energy_fix)
            orc_scale_faces_cart(g, b[0], b[1], b[2], cart_x, cart_y, cart_z, 1);
    }
    if (bx_out) memcpy(bx_out, b[0], sizeof(double) * nf[0]);
    if (by_out) memcpy(by_out, b[1], sizeof(double) * nf[1]);
    if (bz_out && g->dim == 3) memcpy(bz_out, b[2], sizeof(double) * nf[2]);

    orc_mg *mg = orc_mg_create(g, b[0], b[1], b[2], o);
    int rc = orc_mg_solve(mg, phi, rh, st);
    orc_mg_destroy(mg);
    if (rh_out) memcpy(rh_out, rh, sizeof(double) * nc);

    orc_mkumac(g, u, v, w, phi, b[0], b[1], b[2]);

    if (use_1d) orc_scale_faces_1d(g, u, v, w, beta0, beta0_edge, 0);
    else if (use_cart) orc_scale_faces_cart(g, u, v, w, cart_x, cart_y, cart_z, 0);

    free(rh);
    for (int d = 0; d < 3; ++d) free(b[d]);
    return rc;
}

/* ------------------------------------- mac_multigrid with alpha (Helmholtz)
 * mac_multigrid (Source/mac_multigrid.f90:18-71) solves
 *   (alpha - D beta G) phi = rh
 * for any cell coefficient alpha (0 in macproject, rho in the thermal
 * diffusion solve, Source/thermal_conduct.f90:330-377) with the elliptic BCs
 * of the component bc_comp.  The oracle grid carries physical codes; an
 * elliptic Dirichlet side is given as OUTLET, Neumann as SLIP_WALL.
 *
 * phi_g: the solution with one ghost layer in each direction of the
 * problem, C order with x fastest ((nz+2, ny+2, nx+2) in 3D, (ny+2, nx+2)
 * in 2D).  Its ghost cells next to a Dirichlet side hold the boundary value
 * on that face (FBoxLib takes Dirichlet values from phi's ghost cells; the
 * thermal solve fills them from h, thermal_conduct.f90:343-348); the
 * two-point boundary form then adds (2 beta / h^2) g to the right side of
 * the boundary cell.  Its other ghost cells are not read.                  */
static inline long gidx(const orc_grid *g, int i, int j, int k)
{
    const long ex = g->n[0] + 2, ey = g->n[1] + 2;
    return (long)(i + 1) + ex * ((long)(j + 1) + ey * (long)(g->dim == 3 ? k + 1 : 0));
}

void orc_dirichlet_rhs(const orc_grid *g, const double *bx, const double *by, const double *bz,
                       const double *phi_g, double *rh)
{
    const int *n = g->n;
    const double *b[3] = {bx, by, bz};
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                const int c[3] = {i, j, k};
                double r = rh[cidx(n, i, j, k)];
                for (int d = 0; d < g->dim; ++d) {
                    const double ih2 = 1.0 / (g->dx[d] * g->dx[d]);
                    for (int side = 0; side < 2; ++side) {
                        if (c[d] != (side ? n[d] - 1 : 0)) continue;
                        if (orc_ell_bc(g->bc[d][side]) != ORC_ELL_DIR) continue;
                        int f[3] = {i, j, k}, q[3] = {i, j, k};
                        f[d] += side;
                        q[d] += side ? 1 : -1;
                        const double w = 2.0 * b[d][fidx(n, d, f[0], f[1], f[2])];
                        r = fma(ih2, w * phi_g[gidx(g, q[0], q[1], q[2])], r);
                    }
                }
                rh[cidx(n, i, j, k)] = r;
            }
}

int orc_solve(const orc_grid *g, const double *alpha, const double *bx, const double *by,
              const double *bz, const double *rh, double *phi_g, const orc_opts *o, orc_stats *st)
{
    const int *n = g->n;
    const long nc = ncells(n);
    double *r = malloc(sizeof(double) * nc), *phi = malloc(sizeof(double) * nc);
    memcpy(r, rh, sizeof(double) * nc);
    orc_dirichlet_rhs(g, bx, by, bz, phi_g, r);
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) phi[cidx(n, i, j, k)] = phi_g[gidx(g, i, j, k)];
    orc_mg *mg = orc_mg_create_alpha(g, alpha, bx, by, bz, o);
    int rc = orc_mg_solve(mg, phi, r, st);
    orc_mg_destroy(mg);
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) phi_g[gidx(g, i, j, k)] = phi[cidx(n, i, j, k)];
    free(r);
    free(phi);
    return rc;
}

void orc_default_opts(orc_opts *o)
{
    memset(o, 0, sizeof *o);
    o->eps_rel = 1e-10;     /* min(eps_mac*10^(nlevs-1), eps_mac_max), nlevs=1 */
    o->eps_abs = -1.0;      /* macproject.f90:209 */
    o->max_iter = 200;
    o->nu1 = 2;
    o->nu2 = 2;
    o->nub = 10;            /* mg_bottom_nu, _parameters:392 */
    o->bottom_eps = 1e-3;   /* mac_multigrid.f90:61 */
    o->bottom_maxit = 1000;
    o->bottom_solver = ORC_BOTTOM_CG;
    o->max_levels = 0;
    o->use_alt_energy_fix = 1;  /* _parameters:554 */
    o->omega = 1.0;             /* FBoxLib's GSRB */
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------ FP32 sweep, folded weights */

/* One red-black Gauss-Seidel sweep in single precision on the solver's
 * folded-weight form of a level (the form the GPU library stores): bx, by,
 * bz hold the low face of every cell with the boundary conditions already
 * folded in (Neumann 0, Dirichlet 2 beta); bxh (ny*nz), byh (nx*nz) and bzh
 * (nx*ny) hold the domain's high faces; a neighbour outside a non-periodic
 * side has value 0, a periodic one is the image.  Evaluation order is that
 * of cell_gs, in fmaf.  It checks the FP32 smoother kernels of the mixed
 * path; its own pin is the FP64 sweep orc_gsrb (agreement to FP32 rounding
 * on the same level, tests/test_oracle.py). */
void orc_gsrb_folded_f32(int dim, const int *n, const int *per, const float *ih2,
                         const float *bx, const float *by, const float *bz,
                         const float *bxh, const float *byh, const float *bzh,
                         float *phi, const float *rhs)
{
    orc_gsrb_folded_f32_omega(dim, n, per, ih2, bx, by, bz, bxh, byh, bzh, phi, rhs, 1.0f);
}

/* the same sweep over-relaxed by omega, as orc_gsrb_omega: a cell moves to
 * fmaf(omega, gs - old, old); omega = 1 stores gs itself                    */
void orc_gsrb_folded_f32_omega(int dim, const int *n, const int *per, const float *ih2,
                               const float *bx, const float *by, const float *bz,
                               const float *bxh, const float *byh, const float *bzh,
                               float *phi, const float *rhs, float omega)
{
    const int nx = n[0], ny = n[1], nz = (dim == 3) ? n[2] : 1;
    const long sy = nx, sz = (long)nx * ny;
    for (int color = 0; color < 2; ++color) {
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = (j + k + color) & 1; i < nx; i += 2) {
                    const long c = i + sy * j + sz * k;
                    float w[6], p[6];
                    w[0] = bx[c];
                    w[1] = (i < nx - 1) ? bx[c + 1] : bxh[j + (long)ny * k];
                    w[2] = by[c];
                    w[3] = (j < ny - 1) ? by[c + sy] : byh[i + (long)nx * k];
                    p[0] = (i > 0) ? phi[c - 1] : (per[0] ? phi[c + nx - 1] : 0.0f);
                    p[1] = (i < nx - 1) ? phi[c + 1] : (per[0] ? phi[c - (nx - 1)] : 0.0f);
                    p[2] = (j > 0) ? phi[c - sy] : (per[1] ? phi[c + (long)(ny - 1) * sy] : 0.0f);
                    p[3] = (j < ny - 1) ? phi[c + sy] : (per[1] ? phi[c - (long)(ny - 1) * sy] : 0.0f);
                    float tx = fmaf(w[0], p[0], w[1] * p[1]), sx = w[0] + w[1];
                    float ty = fmaf(w[2], p[2], w[3] * p[3]), sy2 = w[2] + w[3];
                    float num, den;
                    if (dim == 3) {
                        w[4] = bz[c];
                        w[5] = (k < nz - 1) ? bz[c + sz] : bzh[i + (long)nx * j];
                        p[4] = (k > 0) ? phi[c - sz] : (per[2] ? phi[c + (long)(nz - 1) * sz] : 0.0f);
                        p[5] = (k < nz - 1) ? phi[c + sz] : (per[2] ? phi[c - (long)(nz - 1) * sz] : 0.0f);
                        float tz = fmaf(w[4], p[4], w[5] * p[5]), szz = w[4] + w[5];
                        num = fmaf(ih2[0], tx, fmaf(ih2[1], ty, fmaf(ih2[2], tz, rhs[c])));
                        den = fmaf(ih2[0], sx, fmaf(ih2[1], sy2, ih2[2] * szz));
                    } else {
                        num = fmaf(ih2[0], tx, fmaf(ih2[1], ty, rhs[c]));
                        den = fmaf(ih2[0], sx, ih2[1] * sy2);
                    }
                    const float v = num / den;
                    phi[c] = (omega == 1.0f) ? v : fmaf(omega, v - phi[c], phi[c]);
                }
    }
}
