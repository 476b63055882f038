// mac_ml.hip -- kernels of the two-level (AMR) MAC projection, MAESTRO's
// macproject with nlevs = 2 (Source/macproject.f90:51-328):
//   k_ml_gather     the caller's boxes of one level -> whole-level arrays:
//                   beta0 U (:154-158), beta = beta0^p / rho (mk_mac_coeffs
//                   :698-738, then :178-193), macrhs and the warm start
//   k_ml_edge_restrict  ml_edge_restriction (:357-361, :829-833): coarse
//                   faces under the fine region <- mean of the fine faces
//   k_ml_div        divumac (:457-480): rh = macrhs - D per level
//   k_ml_cell_restrict  ml_cc_restriction (:401-404): covered coarse cells
//                   <- mean of their fine cells
//   k_ml_resid      residual of the composite operator (FBoxLib ml_cc_solve,
//                   Docs/mg/mg.tex:9-19,124-156) and its max norm
//   k_ml_prolong    the coarse correction of one composite iteration, added
//                   on both levels
//   k_ml_mkumac_*   mkumac (:908-975) with the composite fluxes at the
//                   coarse-fine (C-F) faces
//   k_ml_unscale, k_ml_scatter, k_ml_grown  U / beta0 (:229-240), the
//                   caller's boxes, and the fine ghost faces outside the
//                   fine region from the coarse faces (create_umac_grown)
//
// Whole-level arrays are C order, x fastest, no ghosts: coarse cells nc,
// coarse faces nc + e_d; fine arrays cover the region, the fine boxes'
// bounding rectangle (region cells nr, faces nr + e_d), whose lowest fine
// cell is rlo; MlGeom::cov marks the region's coarse cells a box covers
// (the others are holes: no unknowns, never written).
//
// The composite operator at a C-F face: the fine cell f0 next to it sees
// the ghost value
//     g = 8/15 Cint + 2/3 phi(f0) - 1/5 phi(f1)
// (the quadratic along the face normal through the coarse value, one fine
// cell beyond the face, and the first two fine cells inside), Cint the
// coarse cell across the face moved to the ghost's transverse position with
// its linear slopes (central differences; one-sided next to a wall).  The
// fine flux is beta (phi(f0) - g) / h_f (sign by side); the coarse cell
// across takes the mean of the fine fluxes through its face.  FBoxLib's own
// interpolation is not in the reference tree; this one is second order.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mac_launch.h"

namespace mac {

static constexpr int MTPB = 256;
static constexpr double W_COARSE = 8.0 / 15.0, W_FINE0 = 2.0 / 3.0, W_FINE1 = -1.0 / 5.0;

static inline unsigned mblk(long n) { return (unsigned)((n + MTPB - 1) / MTPB); }

__device__ __forceinline__ long nat_cell(const int *n, int i, int j, int k)
{
    return i + (long)n[0] * (j + (long)n[1] * k);
}
__device__ __forceinline__ long nat_face(const int *n, int d, int i, int j, int k)
{
    const long ex = n[0] + (d == 0), ey = n[1] + (d == 1);
    return i + ex * (j + ey * (long)k);
}
// offset of global index (i,j,k) in a caller's box array (ghost width ng,
// face direction fd or -1 for cells)
__device__ __forceinline__ long ml_boxoff(const BoxDev &B, int ng, int fd, int dim, int i, int j, int k)
{
    const long ex = B.hi[0] - B.lo[0] + 1 + 2 * ng + (fd == 0);
    const long ey = B.hi[1] - B.lo[1] + 1 + 2 * ng + (fd == 1);
    const long li = i - B.lo[0] + ng, lj = j - B.lo[1] + ng;
    const long lk = (dim == 3) ? (k - B.lo[2] + ng) : 0;
    return li + ex * (lj + ey * lk);
}

// coarse neighbour of c along d: false beyond a non-periodic side
__device__ __forceinline__ bool ml_cnb(const MlGeom &G, int c[3], int d, int step)
{
    c[d] += step;
    if (c[d] >= 0 && c[d] < G.nc[d]) return true;
    if (G.ckind[d][0] != ELL_PER) return false;
    c[d] = (c[d] + G.nc[d]) % G.nc[d];
    return true;
}
// whether the region cell (region-relative coarse index q[], inside the
// region) is covered by a fine box
__device__ __forceinline__ bool ml_cov_at(const MlGeom &G, int q0, int q1, int q2)
{
    if (!G.cov) return true;
    return G.cov[q0 + (G.nr[0] >> 1) * (q1 + (G.nr[1] >> 1) * q2)] != 0;
}
__device__ __forceinline__ bool ml_covered(const MlGeom &G, const int c[3])
{
    int q[3] = {0, 0, 0};
    for (int d = 0; d < G.dim; ++d) {
        q[d] = c[d] - (G.rlo[d] >> 1);
        if (q[d] < 0 || q[d] >= (G.nr[d] >> 1)) return false;
    }
    return ml_cov_at(G, q[0], q[1], q[2]);
}
// whether fine cell f (global fine index) belongs to a fine box
__device__ __forceinline__ bool ml_in_region(const MlGeom &G, const int f[3])
{
    int q[3] = {0, 0, 0};
    for (int d = 0; d < G.dim; ++d) {
        const int r = f[d] - G.rlo[d];
        if (r < 0 || r >= G.nr[d]) return false;
        q[d] = r >> 1;
    }
    return ml_cov_at(G, q[0], q[1], q[2]);
}
// fine neighbour of f along d in the fine domain (periodic wrap); false
// beyond a non-periodic side
__device__ __forceinline__ bool ml_fnb(const MlGeom &G, int f[3], int d, int step)
{
    f[d] += step;
    const int nf = 2 * G.nc[d];
    if (f[d] >= 0 && f[d] < nf) return true;
    if (G.ckind[d][0] != ELL_PER) return false;
    f[d] = (f[d] + nf) % nf;
    return true;
}
__device__ __forceinline__ double ml_phic(const MlState &S, const int c[3])
{
    return S.phc[nat_cell(S.G.nc, c[0], c[1], c[2])];
}
__device__ __forceinline__ double ml_phif(const MlState &S, const int f[3])
{
    return S.phf[nat_cell(S.G.nr, f[0] - S.G.rlo[0], f[1] - S.G.rlo[1], f[2] - S.G.rlo[2])];
}

// the coarse value of c moved by sig[t] coarse cells along each transverse
// direction t != d (d = -1: every direction), linear slopes of x (central,
// one-sided where a neighbour is missing)
__device__ double ml_lin(const MlGeom &G, const double *x, const int c[3], int d, const double sig[3],
                         bool uncovered_only)
{
    const double xc = x[nat_cell(G.nc, c[0], c[1], c[2])];
    double v = xc;
    for (int t = 0; t < G.dim; ++t) {
        if (t == d) continue;
        int cp[3] = {c[0], c[1], c[2]}, cm[3] = {c[0], c[1], c[2]};
        bool hp = ml_cnb(G, cp, t, 1), hm = ml_cnb(G, cm, t, -1);
        // the interface ghost uses composite unknowns only: a transverse
        // neighbour under the fine level (a re-entrant corner) is skipped
        if (uncovered_only) {
            hp = hp && !ml_covered(G, cp);
            hm = hm && !ml_covered(G, cm);
        }
        double sl = 0.0;
        if (hp && hm) sl = 0.5 * (x[nat_cell(G.nc, cp[0], cp[1], cp[2])] - x[nat_cell(G.nc, cm[0], cm[1], cm[2])]);
        else if (hp) sl = x[nat_cell(G.nc, cp[0], cp[1], cp[2])] - xc;
        else if (hm) sl = xc - x[nat_cell(G.nc, cm[0], cm[1], cm[2])];
        v = dfma(sig[t], sl, v);
    }
    return v;
}

// flux beta d(phi)/dx_d (positive along +d) through the face on side s of
// fine cell f (global fine indices)
__device__ double ml_fine_flux(const MlState &S, const int f[3], int d, int s)
{
    const MlGeom &G = S.G;
    int fr[3] = {f[0] - G.rlo[0], f[1] - G.rlo[1], f[2] - G.rlo[2]};
    fr[d] += s;
    const double beta = S.bf[d][nat_face(G.nr, d, fr[0], fr[1], fr[2])];
    const double ih = G.ihf[d];
    int nb[3] = {f[0], f[1], f[2]};
    const double p0 = ml_phif(S, f);
    if (!ml_fnb(G, nb, d, s ? 1 : -1)) {
        // a physical side of the domain
        if (G.ckind[d][s] == ELL_DIR) return (s ? -2.0 : 2.0) * beta * ih * p0;
        return 0.0;
    }
    if (!ml_in_region(G, nb)) {
        {
            // coarse-fine face
            const int inward = s ? -1 : 1;
            int f1[3] = {f[0], f[1], f[2]};
            f1[d] += inward;
            int C[3] = {f[0] >> 1, f[1] >> 1, f[2] >> 1};
            ml_cnb(G, C, d, -inward);
            double sig[3];
            for (int t = 0; t < 3; ++t) sig[t] = (f[t] & 1) ? 0.25 : -0.25;
            const double ci = ml_lin(G, S.phc, C, d, sig, true);
            const double g = dfma(W_COARSE, ci, dfma(W_FINE0, p0, W_FINE1 * ml_phif(S, f1)));
            return s ? beta * ih * (g - p0) : beta * ih * (p0 - g);
        }
    }
    const double pn = ml_phif(S, nb);
    return s ? beta * ih * (pn - p0) : beta * ih * (p0 - pn);
}

// flux through the face on side s of uncovered coarse cell c
__device__ double ml_coarse_flux(const MlState &S, const int c[3], int d, int s)
{
    const MlGeom &G = S.G;
    int fc[3] = {c[0], c[1], c[2]};
    fc[d] += s;
    const double beta = S.bc[d][nat_face(G.nc, d, fc[0], fc[1], fc[2])];
    const double ih = G.ihc[d];
    int nb[3] = {c[0], c[1], c[2]};
    const double p0 = ml_phic(S, c);
    if (!ml_cnb(G, nb, d, s ? 1 : -1)) {
        if (G.ckind[d][s] == ELL_DIR) return (s ? -2.0 : 2.0) * beta * ih * p0;
        return 0.0;
    }
    if (!ml_covered(G, nb)) {
        const double pn = ml_phic(S, nb);
        return s ? beta * ih * (pn - p0) : beta * ih * (p0 - pn);
    }
    // C-F face: the mean of the fine fluxes through it, from the fine cells
    // just inside the region (their face on the far side from c)
    const int fs = s ? 0 : 1;
    const int fd = s ? 2 * nb[d] : 2 * nb[d] + 1;
    const int t0 = (d == 0) ? 1 : 0, t1 = (d == 2) ? 1 : 2;
    const int cnt = (G.dim == 3) ? 4 : 2;
    double F = 0.0;
    for (int m = 0; m < cnt; ++m) {
        int f[3];
        f[d] = fd;
        f[t0] = 2 * c[t0] + (m & 1);
        f[t1] = (G.dim == 3) ? 2 * c[t1] + (m >> 1) : 0;
        F += ml_fine_flux(S, f, d, fs);
    }
    return F / cnt;
}

// ---------------------------------------------------------------- gather
// One thread per valid cell of each caller box of a level: the low face of
// the cell in every direction (and the high face where the cell ends the
// level's rectangle): beta0 U and the face coefficient 2 / (rho + rho_nb)
// times beta0^p; macrhs and phi's valid cells.  Each face of the rectangle
// is written by exactly one box.
template <int DIM>
__global__ void __launch_bounds__(MTPB) k_ml_gather(MlGather A)
{
    const int b = blockIdx.y;
    const BoxDev B = A.boxes[b];
    const int ex = B.hi[0] - B.lo[0] + 1, ey = B.hi[1] - B.lo[1] + 1;
    const int ez = (DIM == 3) ? B.hi[2] - B.lo[2] + 1 : 1;
    const long q = (long)blockIdx.x * MTPB + threadIdx.x;
    if (q >= (long)ex * ey * ez) return;
    const int g[3] = {B.lo[0] + (int)(q % ex), B.lo[1] + (int)((q / ex) % ey),
                      (DIM == 3) ? B.lo[2] + (int)(q / ((long)ex * ey)) : 0};
    const int r[3] = {g[0] - A.org[0], g[1] - A.org[1], g[2] - A.org[2]};
    const int rd = DIM - 1;
    bool bad = false;
    const double *rho = A.rho[b];
    const double rc = rho[ml_boxoff(B, A.ng_rho, -1, DIM, g[0], g[1], g[2])];
    bad |= !isfinite(rc);
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const double *u = A.umac[b * DIM + d];
        // the high face too where the cell ends the level's rectangle, or
        // ends its box with no fine box after it (a face on a hole)
        bool last = r[d] == A.n[d] - 1;
        if (!last && A.cov && g[d] == B.hi[d]) {
            int q[3] = {r[0] >> 1, r[1] >> 1, r[2] >> 1};
            q[d] = (r[d] + 1) >> 1;
            last = A.cov[q[0] + (A.n[0] >> 1) * (q[1] + (A.n[1] >> 1) * q[2])] == 0;
        }
        for (int s = 0; s < 2; ++s) {
            if (s == 1 && !last) break;
            int f[3] = {g[0], g[1], g[2]};
            f[d] += s;
            int nb[3] = {g[0], g[1], g[2]};
            nb[d] += s ? 1 : -1;
            double fac = 1.0;
            if (A.b0) fac = (d == rd) ? A.b0e[f[rd]] : A.b0[f[rd]];
            const double uv = u[ml_boxoff(B, A.ng_u, d, DIM, f[0], f[1], f[2])];
            const double rn = rho[ml_boxoff(B, A.ng_rho, -1, DIM, nb[0], nb[1], nb[2])];
            double beta = 2.0 / (rc + rn);
            if (A.b0) {
                beta = beta * fac;
                if (A.alt) beta = beta * fac;
            }
            bad |= !isfinite(uv) || !isfinite(beta);
            const long o = nat_face(A.n, d, r[0] + (d == 0) * s, r[1] + (d == 1) * s, r[2] + (d == 2) * s);
            A.u[d][o] = A.b0 ? uv * fac : uv;
            A.beta[d][o] = beta;
        }
    }
    const long c = nat_cell(A.n, r[0], r[1], r[2]);
    if (A.macrhs) {
        const double m = A.macrhs[b][ml_boxoff(B, A.ng_rhs, -1, DIM, g[0], g[1], g[2])];
        bad |= !isfinite(m);
        A.rhs[c] = m;
    }
    const double p = A.phi[b][ml_boxoff(B, A.ng_phi, -1, DIM, g[0], g[1], g[2])];
    bad |= !isfinite(p);
    A.phi_out[c] = p;
    if (bad) atomicOr(A.err, 1);
}

// --------------------------------------------------------- level transfers
// coarse faces under the region (its boundary faces included) <- mean of the
// 2^(dim-1) fine faces; where the region meets a periodic seam without
// spanning the direction, the face's periodic image as well
template <int DIM>
__global__ void __launch_bounds__(MTPB) k_ml_edge_restrict(MlGeom G, double *uc0, double *uc1, double *uc2,
                                                          const double *uf0, const double *uf1, const double *uf2)
{
    const int d = blockIdx.y;
    double *uc = d == 0 ? uc0 : (d == 1 ? uc1 : uc2);
    const double *uf = d == 0 ? uf0 : (d == 1 ? uf1 : uf2);
    int m[3];
    for (int e = 0; e < 3; ++e) m[e] = (e < DIM) ? (G.nr[e] >> 1) + (e == d) : 1;
    const long q = (long)blockIdx.x * MTPB + threadIdx.x;
    if (q >= (long)m[0] * m[1] * m[2]) return;
    const int l[3] = {(int)(q % m[0]), (int)((q / m[0]) % m[1]), (int)(q / ((long)m[0] * m[1]))};
    const int t0 = (d == 0) ? 1 : 0, t1 = (d == 2) ? 1 : 2;
    const int cnt = (DIM == 3) ? 4 : 2;
    double s = 0.0;
    for (int k = 0; k < cnt; ++k) {
        int f[3];
        f[d] = 2 * l[d];
        f[t0] = 2 * l[t0] + (k & 1);
        f[t1] = (DIM == 3) ? 2 * l[t1] + (k >> 1) : 0;
        s += uf[nat_face(G.nr, d, f[0], f[1], f[2])];
    }
    const double v = s / cnt;
    int c[3] = {(G.rlo[0] >> 1) + l[0], (G.rlo[1] >> 1) + l[1], (DIM == 3) ? (G.rlo[2] >> 1) + l[2] : 0};
    // only faces with a covered coarse cell on a side (the low one through
    // the periodic wrap)
    if (G.cov) {
        int hi_c[3] = {c[0], c[1], c[2]}, lo_c[3] = {c[0], c[1], c[2]};
        bool cov = false;
        if (hi_c[d] < G.nc[d]) cov = ml_covered(G, hi_c);
        if (!cov && ml_cnb(G, lo_c, d, -1)) cov = ml_covered(G, lo_c);
        if (!cov) return;
    }
    uc[nat_face(G.nc, d, c[0], c[1], c[2])] = v;
    if (G.ckind[d][0] == ELL_PER && G.side[d][0] == ML_SIDE_CF) {
        if (c[d] == 0) c[d] = G.nc[d];
        else if (c[d] == G.nc[d]) c[d] = 0;
        else return;
        uc[nat_face(G.nc, d, c[0], c[1], c[2])] = v;
    }
}

// rh = macrhs - D (rh = -D without macrhs), whole level
template <int DIM>
__global__ void __launch_bounds__(MTPB) k_ml_div(int n0, int n1, int n2, double dx0, double dx1, double dx2,
                                                 const double *u, const double *v, const double *w,
                                                 const double *macrhs, double *rh)
{
    const long q = (long)blockIdx.x * MTPB + threadIdx.x;
    const long nc = (long)n0 * n1 * n2;
    if (q >= nc) return;
    const int i = (int)(q % n0), j = (int)((q / n0) % n1), k = (int)(q / ((long)n0 * n1));
    const long fx = i + (long)(n0 + 1) * (j + (long)n1 * k);
    const long fy = i + (long)n0 * (j + (long)(n1 + 1) * k);
    double D = (u[fx + 1] - u[fx]) / dx0 + (v[fy + n0] - v[fy]) / dx1;
    if (DIM == 3) D = D + (w[q + (long)n0 * n1] - w[q]) / dx2;
    rh[q] = macrhs ? -(D - macrhs[q]) : -D;
}

// covered coarse cells <- mean of their 2^dim fine cells
template <int DIM>
__global__ void __launch_bounds__(MTPB) k_ml_cell_restrict(MlGeom G, double *xc, const double *xf)
{
    const int m[3] = {G.nr[0] >> 1, G.nr[1] >> 1, (DIM == 3) ? G.nr[2] >> 1 : 1};
    const long q = (long)blockIdx.x * MTPB + threadIdx.x;
    if (q >= (long)m[0] * m[1] * m[2]) return;
    const int l[3] = {(int)(q % m[0]), (int)((q / m[0]) % m[1]), (int)(q / ((long)m[0] * m[1]))};
    if (!ml_cov_at(G, l[0], l[1], l[2])) return;
    double s = 0.0;
    for (int k = 0; k < (DIM == 3 ? 2 : 1); ++k)
        for (int j = 0; j < 2; ++j)
            for (int i = 0; i < 2; ++i)
                s += xf[nat_cell(G.nr, 2 * l[0] + i, 2 * l[1] + j, (DIM == 3) ? 2 * l[2] + k : 0)];
    xc[nat_cell(G.nc, (G.rlo[0] >> 1) + l[0], (G.rlo[1] >> 1) + l[1], (DIM == 3) ? (G.rlo[2] >> 1) + l[2] : 0)] =
        s / (DIM == 3 ? 8 : 4);
}

// the composite vector whose sum is the volume-weighted sum of rh over the
// composite cells (fine-cell units): 2^dim rh on uncovered coarse cells, 0
// on covered ones, then the fine cells
template <int DIM>
__global__ void __launch_bounds__(MTPB) k_ml_comp_vec(MlGeom G, const double *rhc, const double *rhf, double *out)
{
    const long q = (long)blockIdx.x * MTPB + threadIdx.x;
    const long ncc = (long)G.nc[0] * G.nc[1] * G.nc[2], nfc = (long)G.nr[0] * G.nr[1] * G.nr[2];
    if (q >= ncc + nfc) return;
    if (q < ncc) {
        const int c[3] = {(int)(q % G.nc[0]), (int)((q / G.nc[0]) % G.nc[1]), (int)(q / ((long)G.nc[0] * G.nc[1]))};
        out[q] = ml_covered(G, c) ? 0.0 : (DIM == 3 ? 8.0 : 4.0) * rhc[q];
    } else {
        const long p = q - ncc;
        const int f[3] = {G.rlo[0] + (int)(p % G.nr[0]), G.rlo[1] + (int)((p / G.nr[0]) % G.nr[1]),
                          G.rlo[2] + (int)(p / ((long)G.nr[0] * G.nr[1]))};
        out[q] = ml_in_region(G, f) ? rhf[p] : 0.0;
    }
}

// ------------------------------------------------------ composite residual
// r = rh - A phi on the uncovered coarse cells (into rc; covered cells left
// alone) and the fine cells (into rf), max |r| into *nrm.  fine_only: the
// fine cells only.  Cells whose six neighbours are plain cells of their own
// level (no wall, wrap, covered cell or region boundary) take the 7-point
// form directly; the others go through the flux functions.  32-bit index
// arithmetic (a level holds fewer than 2^31 cells).
template <int DIM>
__global__ void __launch_bounds__(MTPB) k_ml_resid(MlState S, double *rc, double *rf, unsigned long long *nrm,
                                                   int fine_only)
{
    const MlGeom &G = S.G;
    const int ncc = G.nc[0] * G.nc[1] * G.nc[2], nfc = G.nr[0] * G.nr[1] * G.nr[2];
    const int qend = ncc + nfc;
    double amax = 0.0;
    // grid-stride over the cells: one norm atomic per block, not per wave
    for (int q = blockIdx.x * MTPB + threadIdx.x + (fine_only ? ncc : 0); q < qend; q += gridDim.x * MTPB) {
    double a = 0.0;
    if (q < ncc) {
        const int t = q / G.nc[0];
        const int c[3] = {q - t * G.nc[0], t % G.nc[1], t / G.nc[1]};
        if (!ml_covered(G, c)) {
            bool plain = true;
            for (int d = 0; d < DIM; ++d) {
                if (c[d] == 0 || c[d] == G.nc[d] - 1) {
                    plain = false;
                    break;
                }
                if (G.cov) {
                    int cp[3] = {c[0], c[1], c[2]}, cm[3] = {c[0], c[1], c[2]};
                    cp[d] += 1;
                    cm[d] -= 1;
                    if (ml_covered(G, cp) || ml_covered(G, cm)) plain = false;
                } else {
                    // one rectangle: a covered neighbour along d means the
                    // cell lies just outside its coarse range in d
                    const int lo = G.rlo[d] >> 1, hi = (G.rlo[d] + G.nr[d] - 1) >> 1;
                    if (c[d] == lo - 1 || c[d] == hi + 1) plain = false;
                }
            }
            double Ap = 0.0;
            if (plain) {
                const double pc = S.phc[q];
                const int st[3] = {1, G.nc[0], G.nc[0] * G.nc[1]};
                for (int d = 0; d < DIM; ++d) {
                    const long f0 = nat_face(G.nc, d, c[0], c[1], c[2]);
                    const long f1 = f0 + ((d == 0) ? 1 : (d == 1 ? (G.nc[0]) : (long)G.nc[0] * (G.nc[1])));
                    const double bl = S.bc[d][f0], bh = S.bc[d][f1];
                    Ap -= (bh * G.ihc[d] * (S.phc[q + st[d]] - pc) - bl * G.ihc[d] * (pc - S.phc[q - st[d]])) *
                          G.ihc[d];
                }
            } else {
                for (int d = 0; d < DIM; ++d)
                    Ap -= (ml_coarse_flux(S, c, d, 1) - ml_coarse_flux(S, c, d, 0)) * G.ihc[d];
            }
            const double r = S.rhc[q] - Ap;
            rc[q] = r;
            a = fabs(r);
        }
    } else if (q < ncc + nfc) {
        const int p = q - ncc;
        const int t = p / G.nr[0];
        const int r3[3] = {p - t * G.nr[0], t % G.nr[1], t / G.nr[1]};
        const int fg[3] = {G.rlo[0] + r3[0], G.rlo[1] + r3[1], G.rlo[2] + r3[2]};
        if (!ml_in_region(G, fg)) continue;   // a hole of the region
        bool plain = true;
        for (int d = 0; d < DIM; ++d) {
            if (r3[d] == 0 || r3[d] == G.nr[d] - 1) {
                plain = false;
                break;
            }
            if (G.cov) {
                int fp[3] = {fg[0], fg[1], fg[2]}, fm[3] = {fg[0], fg[1], fg[2]};
                fp[d] += 1;
                fm[d] -= 1;
                if (!ml_in_region(G, fp) || !ml_in_region(G, fm)) plain = false;
            }
        }
        double Ap = 0.0;
        if (plain) {
            const double pc = S.phf[p];
            const int st[3] = {1, G.nr[0], G.nr[0] * G.nr[1]};
            for (int d = 0; d < DIM; ++d) {
                const long f0 = nat_face(G.nr, d, r3[0], r3[1], r3[2]);
                const long f1 = f0 + ((d == 0) ? 1 : (d == 1 ? (G.nr[0]) : (long)G.nr[0] * (G.nr[1])));
                const double bl = S.bf[d][f0], bh = S.bf[d][f1];
                Ap -= (bh * G.ihf[d] * (S.phf[p + st[d]] - pc) - bl * G.ihf[d] * (pc - S.phf[p - st[d]])) * G.ihf[d];
            }
        } else {
            const int f[3] = {G.rlo[0] + r3[0], G.rlo[1] + r3[1], G.rlo[2] + r3[2]};
            for (int d = 0; d < DIM; ++d) Ap -= (ml_fine_flux(S, f, d, 1) - ml_fine_flux(S, f, d, 0)) * G.ihf[d];
        }
        const double r = S.rhf[p] - Ap;
        rf[p] = r;
        a = fabs(r);
    }
    if (!isfinite(a)) a = __builtin_huge_val();
    amax = fmax(amax, a);
    }
    amax = wave_max(amax);
    __shared__ double wm[MTPB / 64];
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = amax;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = wm[0];
        for (int w = 1; w < MTPB / 64; ++w) m = fmax(m, wm[w]);
        if (m > 0.0) atomicMax(nrm, (unsigned long long)__double_as_longlong(m));
    }
}

// the coarse correction e (whole coarse level) added to phi on both levels:
// coarse cells directly, fine cells through the linear interpolation of e
template <int DIM>
__global__ void __launch_bounds__(MTPB) k_ml_prolong(MlGeom G, const double *e, double *phc, double *phf)
{
    const int ncc = G.nc[0] * G.nc[1] * G.nc[2], nfc = G.nr[0] * G.nr[1] * G.nr[2];
    const int q = blockIdx.x * MTPB + threadIdx.x;
    if (q < ncc) {
        phc[q] += e[q];
    } else if (q < ncc + nfc) {
        const int p = q - ncc;
        const int t = p / G.nr[0];
        const int f[3] = {G.rlo[0] + p - t * G.nr[0], G.rlo[1] + t % G.nr[1], G.rlo[2] + t / G.nr[1]};
        const int C[3] = {f[0] >> 1, f[1] >> 1, f[2] >> 1};
        double sig[3];
        for (int t = 0; t < 3; ++t) sig[t] = (f[t] & 1) ? 0.25 : -0.25;
        if (ml_in_region(G, f)) phf[p] += ml_lin(G, e, C, -1, sig, false);
    }
}

__global__ void __launch_bounds__(MTPB) k_ml_axpy(long n, const double *x, double *y)
{
    const long q = (long)blockIdx.x * MTPB + threadIdx.x;
    if (q < n) y[q] += x[q];
}

// The fine correction solve's coarse-fine sides: with the coarse values
// held, the composite flux through such a face changes by beta (1/3 e_0 +
// 1/5 e_1) / h for a change e of the first two fine cells, about 8/15 beta
// e_0 / h for a smooth e, where the Dirichlet folding of the plan (zero on
// the face) has 2 beta e_0 / h.  Scaling those folded weights by 4/15 makes
// the inner problem match the composite one near the interface, which the
// composite iteration's rate depends on (the answer does not: the composite
// residual decides convergence).
__global__ void __launch_bounds__(MTPB) k_ml_cf_weights(MlGeom G, Lev<double> L, double f)
{
    const int side = blockIdx.y, d = side >> 1, s = side & 1;
    if (d >= G.dim || G.side[d][s] != ML_SIDE_CF) return;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int t0 = (d == 0) ? ny : nx, t1 = (d == 2) ? ny : nz;
    const long q = (long)blockIdx.x * MTPB + threadIdx.x;
    if (q >= (long)t0 * t1) return;
    const int a = (int)(q % t0), b = (int)(q / t0);
    if (d == 0) {
        if (s) L.bxh[a + (long)ny * b] *= f;
        else L.bx[(long)nx * (a + (long)ny * b)] *= f;
    } else if (d == 1) {
        if (s) L.byh[a + (long)nx * b] *= f;
        else L.by[a + (long)nx * ny * b] *= f;
    } else {
        if (s) L.bzh[a + (long)nx * b] *= f;
        else L.bz[a + (long)nx * b] *= f;
    }
}

// Holes of the region in the fine correction plan's level 0: alpha 0 on
// the boxes' cells and, on a hole, 1e4 times the largest diagonal its faces
// could give (so its correction stays ~1e-4 of its neighbours'); the face
// between a box cell and a hole scaled by f (the interface weight, as on the
// region's own coarse-fine sides).  Each thread owns its cell's low faces
// (and the high slab faces across a periodic wrap).
__global__ void __launch_bounds__(MTPB) k_ml_holes(MlGeom G, Lev<double> L, double f)
{
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int q = blockIdx.x * MTPB + threadIdx.x;
    if (q >= nx * ny * nz) return;
    const int t = q / nx;
    const int r[3] = {q - t * nx, t % ny, t / ny};
    const int dim = G.dim;
    auto inb = [&](int a, int b, int c) { return ml_cov_at(G, a >> 1, b >> 1, c >> 1); };
    const bool in = inb(r[0], r[1], r[2]);
    const int n[3] = {nx, ny, nz};
    const int per[3] = {L.perx, L.pery, L.perz};
    double *lo_w[3] = {L.bx, L.by, L.bz};
    double wmax = 0.0;
    for (int d = 0; d < dim; ++d) {
        int m[3] = {r[0], r[1], r[2]};
        if (r[d] > 0 || per[d]) {
            m[d] = (r[d] + n[d] - 1) % n[d];
            if (inb(m[0], m[1], m[2]) != in) lo_w[d][q] *= f;
        }
        wmax = fmax(wmax, lo_w[d][q]);
        if (r[d] == n[d] - 1 && per[d]) {
            int p[3] = {r[0], r[1], r[2]};
            p[d] = 0;
            double *sl = (d == 0) ? &L.bxh[r[1] + (long)ny * r[2]]
                                  : (d == 1 ? &L.byh[r[0] + (long)nx * r[2]] : &L.bzh[r[0] + (long)nx * r[1]]);
            if (inb(p[0], p[1], p[2]) != in) *sl *= f;
        }
    }
    const double ih2 = (double)L.ih2x + (double)L.ih2y + (double)L.ih2z;
    L.alpha[q] = in ? 0.0 : 1e4 * 2.0 * ih2 * fmax(wmax, 1.0);
}

// ------------------------------------------------------------------ mkumac
// fine faces: U -= beta G phi with the composite flux (C-F ghosts at the
// region's C-F faces, the physical rule on its physical sides)
template <int DIM>
__global__ void __launch_bounds__(MTPB) k_ml_mkumac_fine(MlState S, double *u0, double *u1, double *u2)
{
    const MlGeom &G = S.G;
    const int d = blockIdx.y;
    double *u = d == 0 ? u0 : (d == 1 ? u1 : u2);
    int m[3];
    for (int e = 0; e < 3; ++e) m[e] = G.nr[e] + (e == d);
    const long q = (long)blockIdx.x * MTPB + threadIdx.x;
    if (q >= (long)m[0] * m[1] * m[2]) return;
    int r[3] = {(int)(q % m[0]), (int)((q / m[0]) % m[1]), (int)(q / ((long)m[0] * m[1]))};
    // the face's fine cells: the flux from the one in a fine box (the high
    // one's low face, else the low one's high face); faces between holes of
    // the region are no faces of the level
    int fh[3] = {G.rlo[0] + r[0], G.rlo[1] + r[1], G.rlo[2] + r[2]};
    int fl[3] = {fh[0], fh[1], fh[2]};
    fl[d] -= 1;
    if (r[d] < G.nr[d] && ml_in_region(G, fh)) u[q] -= ml_fine_flux(S, fh, d, 0);
    else if (r[d] > 0 && ml_in_region(G, fl)) u[q] -= ml_fine_flux(S, fl, d, 1);
}

// coarse faces: U -= beta G phi with the single-level boundary rule (faces
// under the region are overwritten by the edge restriction afterwards)
template <int DIM>
__global__ void __launch_bounds__(MTPB) k_ml_mkumac_coarse(MlState S, double *u0, double *u1, double *u2)
{
    const MlGeom &G = S.G;
    const int d = blockIdx.y;
    double *u = d == 0 ? u0 : (d == 1 ? u1 : u2);
    int m[3];
    for (int e = 0; e < 3; ++e) m[e] = G.nc[e] + (e == d);
    const long q = (long)blockIdx.x * MTPB + threadIdx.x;
    if (q >= (long)m[0] * m[1] * m[2]) return;
    int c[3] = {(int)(q % m[0]), (int)((q / m[0]) % m[1]), (int)(q / ((long)m[0] * m[1]))};
    int s = 0;
    if (c[d] == G.nc[d]) { c[d] -= 1; s = 1; }
    const double beta = S.bc[d][q];
    const double ih = G.ihc[d];
    const double p0 = ml_phic(S, c);
    int nb[3] = {c[0], c[1], c[2]};
    double F;
    if (!ml_cnb(G, nb, d, s ? 1 : -1)) {
        F = (G.ckind[d][s] == ELL_DIR) ? (s ? -2.0 : 2.0) * beta * ih * p0 : 0.0;
    } else {
        const double pn = ml_phic(S, nb);
        F = s ? beta * ih * (pn - p0) : beta * ih * (p0 - pn);
    }
    u[q] -= F;
}

// U / beta0 (the level's planar profile) on a whole-level face array
template <int DIM>
__global__ void __launch_bounds__(MTPB) k_ml_unscale(int d, int n0, int n1, int n2, int org_r, double *u,
                                                     const double *b0, const double *b0e)
{
    int m[3] = {n0, n1, n2};
    m[d] += 1;
    const long q = (long)blockIdx.x * MTPB + threadIdx.x;
    if (q >= (long)m[0] * m[1] * m[2]) return;
    const int rd = DIM - 1;
    const int idx[3] = {(int)(q % m[0]), (int)((q / m[0]) % m[1]), (int)(q / ((long)m[0] * m[1]))};
    const int r = org_r + idx[rd];
    u[q] = u[q] / ((d == rd) ? b0e[r] : b0[r]);
}

// whole-level arrays -> the caller's boxes: every valid face and cell
template <int DIM>
__global__ void __launch_bounds__(MTPB) k_ml_scatter(MlScatter A)
{
    const int b = blockIdx.y, d = blockIdx.z;   // d == DIM: the cells (phi)
    const BoxDev B = A.boxes[b];
    int e[3];
    for (int t = 0; t < 3; ++t) e[t] = (t < DIM) ? B.hi[t] - B.lo[t] + 1 + (t == d) : 1;
    const long q = (long)blockIdx.x * MTPB + threadIdx.x;
    if (q >= (long)e[0] * e[1] * e[2]) return;
    const int g[3] = {B.lo[0] + (int)(q % e[0]), B.lo[1] + (int)((q / e[0]) % e[1]),
                      (DIM == 3) ? B.lo[2] + (int)(q / ((long)e[0] * e[1])) : 0};
    const int r[3] = {g[0] - A.org[0], g[1] - A.org[1], g[2] - A.org[2]};
    if (d == DIM) {
        A.phi[b][ml_boxoff(B, A.ng_phi, -1, DIM, g[0], g[1], g[2])] = A.phi_in[nat_cell(A.n, r[0], r[1], r[2])];
    } else {
        A.umac[b * DIM + d][ml_boxoff(B, A.ng_u, d, DIM, g[0], g[1], g[2])] =
            A.u[d][nat_face(A.n, d, r[0], r[1], r[2])];
    }
}

// fine ghost faces outside the region (inside the domain, or its periodic
// image) from the final coarse faces: a fine face on a coarse face takes its
// value, one between two coarse faces their mean; transversely the coarse
// face containing it (FBoxLib's create_umac_grown is not in the reference
// tree; this is its plain form)
template <int DIM>
__global__ void __launch_bounds__(MTPB) k_ml_grown(MlGrown A)
{
    const int b = blockIdx.y, d = blockIdx.z;
    const BoxDev B = A.boxes[b];
    const int ng = A.ng_u;
    int e[3];
    for (int t = 0; t < 3; ++t) e[t] = (t < DIM) ? B.hi[t] - B.lo[t] + 1 + 2 * ng + (t == d) : 1;
    const long q = (long)blockIdx.x * MTPB + threadIdx.x;
    if (q >= (long)e[0] * e[1] * e[2]) return;
    const int g[3] = {B.lo[0] - ng + (int)(q % e[0]), B.lo[1] - ng + (int)((q / e[0]) % e[1]),
                      (DIM == 3) ? B.lo[2] - ng + (int)(q / ((long)e[0] * e[1])) : 0};
    // map into the fine domain (periodic wrap); beyond a wall: physbc's
    int w[3];
    for (int t = 0; t < DIM; ++t) {
        const int nf = 2 * A.G.nc[t];
        int x = g[t];
        const int top = nf - 1 + (t == d);
        if (x < 0 || x > top) {
            if (A.G.ckind[t][0] != ELL_PER) return;
            x = ((x % nf) + nf) % nf;
        }
        w[t] = x;
    }
    if (DIM == 2) w[2] = 0;
    // a face of a fine cell is a face of the fine level (the box images
    // fill it); only the others take the coarse faces
    {
        int hi_f[3] = {w[0], w[1], w[2]}, lo_f[3] = {w[0], w[1], w[2]};
        if (hi_f[d] < 2 * A.G.nc[d] && ml_in_region(A.G, hi_f)) return;
        if (ml_fnb(A.G, lo_f, d, -1) && ml_in_region(A.G, lo_f)) return;
    }
    int c[3] = {w[0] >> 1, w[1] >> 1, w[2] >> 1};
    double v;
    if ((w[d] & 1) == 0) {
        c[d] = w[d] >> 1;
        v = A.uc[d][nat_face(A.G.nc, d, c[0], c[1], c[2])];
    } else {
        c[d] = w[d] >> 1;
        const double lo = A.uc[d][nat_face(A.G.nc, d, c[0], c[1], c[2])];
        c[d] += 1;
        const double hi = A.uc[d][nat_face(A.G.nc, d, c[0], c[1], c[2])];
        v = 0.5 * (lo + hi);
    }
    A.umac[b * DIM + d][ml_boxoff(B, ng, d, DIM, g[0], g[1], g[2])] = v;
}

// ----------------------------------------------------------------- launchers
#define ML_DIM(dim, ...)                        \
    do {                                        \
        if ((dim) == 3) {                       \
            constexpr int D_ = 3;               \
            __VA_ARGS__;                        \
        } else {                                \
            constexpr int D_ = 2;               \
            __VA_ARGS__;                        \
        }                                       \
    } while (0)

void launch_ml_gather(int dim, const MlGather &A, int nboxes, int maxcells, hipStream_t s)
{
    ML_DIM(dim, hipLaunchKernelGGL(k_ml_gather<D_>, dim3(mblk(maxcells), nboxes), dim3(MTPB), 0, s, A));
}

void launch_ml_edge_restrict(const MlGeom &G, double *const uc[3], const double *const uf[3], hipStream_t s)
{
    long m = 1;
    for (int e = 0; e < G.dim; ++e) m *= (G.nr[e] >> 1) + 1;
    ML_DIM(G.dim, hipLaunchKernelGGL(k_ml_edge_restrict<D_>, dim3(mblk(m), G.dim), dim3(MTPB), 0, s, G, uc[0],
                                     uc[1], uc[2], uf[0], uf[1], uf[2]));
}

void launch_ml_div(int dim, const int n[3], const double dx[3], const double *const u[3], const double *macrhs,
                   double *rh, hipStream_t s)
{
    const long nc = (long)n[0] * n[1] * n[2];
    ML_DIM(dim, hipLaunchKernelGGL(k_ml_div<D_>, dim3(mblk(nc)), dim3(MTPB), 0, s, n[0], n[1], n[2], dx[0], dx[1],
                                   dx[2], u[0], u[1], u[2], macrhs, rh));
}

void launch_ml_cell_restrict(const MlGeom &G, double *xc, const double *xf, hipStream_t s)
{
    long m = 1;
    for (int e = 0; e < G.dim; ++e) m *= G.nr[e] >> 1;
    ML_DIM(G.dim, hipLaunchKernelGGL(k_ml_cell_restrict<D_>, dim3(mblk(m)), dim3(MTPB), 0, s, G, xc, xf));
}

void launch_ml_comp_vec(const MlGeom &G, const double *rhc, const double *rhf, double *out, hipStream_t s)
{
    const long n = (long)G.nc[0] * G.nc[1] * G.nc[2] + (long)G.nr[0] * G.nr[1] * G.nr[2];
    ML_DIM(G.dim, hipLaunchKernelGGL(k_ml_comp_vec<D_>, dim3(mblk(n)), dim3(MTPB), 0, s, G, rhc, rhf, out));
}

void launch_ml_resid(const MlState &S, double *rc, double *rf, unsigned long long *nrm, bool fine_only,
                     hipStream_t s)
{
    const MlGeom &G = S.G;
    const long ncc = (long)G.nc[0] * G.nc[1] * G.nc[2], nfc = (long)G.nr[0] * G.nr[1] * G.nr[2];
    const long n = fine_only ? nfc : ncc + nfc;
    const unsigned nb = std::min<unsigned>(mblk(n), 2048u);
    ML_DIM(G.dim, hipLaunchKernelGGL(k_ml_resid<D_>, dim3(nb), dim3(MTPB), 0, s, S, rc, rf, nrm,
                                     fine_only ? 1 : 0));
}

void launch_ml_prolong(const MlGeom &G, const double *e, double *phc, double *phf, hipStream_t s)
{
    const long n = (long)G.nc[0] * G.nc[1] * G.nc[2] + (long)G.nr[0] * G.nr[1] * G.nr[2];
    ML_DIM(G.dim, hipLaunchKernelGGL(k_ml_prolong<D_>, dim3(mblk(n)), dim3(MTPB), 0, s, G, e, phc, phf));
}

void launch_ml_cf_weights(const MlGeom &G, const Lev<double> &L, double f, hipStream_t s)
{
    const long plane = std::max({(long)L.ny * L.nz, (long)L.nx * L.nz, (long)L.nx * L.ny});
    hipLaunchKernelGGL(k_ml_cf_weights, dim3(mblk(plane), 2 * G.dim), dim3(MTPB), 0, s, G, L, f);
}

void launch_ml_holes(const MlGeom &G, const Lev<double> &L, double f, hipStream_t s)
{
    const long n = (long)L.nx * L.ny * L.nz;
    hipLaunchKernelGGL(k_ml_holes, dim3(mblk(n)), dim3(MTPB), 0, s, G, L, f);
}

void launch_ml_axpy(long n, const double *x, double *y, hipStream_t s)
{
    hipLaunchKernelGGL(k_ml_axpy, dim3(mblk(n)), dim3(MTPB), 0, s, n, x, y);
}

void launch_ml_mkumac(const MlState &S, double *const uc[3], double *const uf[3], hipStream_t s)
{
    const MlGeom &G = S.G;
    long mf = 0, mc = 0;
    for (int d = 0; d < G.dim; ++d) {
        long a = 1, b = 1;
        for (int e = 0; e < 3; ++e) {
            a *= G.nr[e] + (e == d);
            b *= G.nc[e] + (e == d);
        }
        mf = a > mf ? a : mf;
        mc = b > mc ? b : mc;
    }
    ML_DIM(G.dim, hipLaunchKernelGGL(k_ml_mkumac_fine<D_>, dim3(mblk(mf), G.dim), dim3(MTPB), 0, s, S, uf[0], uf[1],
                                     uf[2]));
    ML_DIM(G.dim, hipLaunchKernelGGL(k_ml_mkumac_coarse<D_>, dim3(mblk(mc), G.dim), dim3(MTPB), 0, s, S, uc[0],
                                     uc[1], uc[2]));
}

void launch_ml_unscale(int dim, const int n[3], int org_r, double *const u[3], const double *b0, const double *b0e,
                       hipStream_t s)
{
    for (int d = 0; d < dim; ++d) {
        long m = 1;
        for (int e = 0; e < 3; ++e) m *= n[e] + (e == d);
        ML_DIM(dim, hipLaunchKernelGGL(k_ml_unscale<D_>, dim3(mblk(m)), dim3(MTPB), 0, s, d, n[0], n[1], n[2], org_r,
                                       u[d], b0, b0e));
    }
}

void launch_ml_scatter(int dim, const MlScatter &A, int nboxes, int maxfaces, hipStream_t s)
{
    ML_DIM(dim, hipLaunchKernelGGL(k_ml_scatter<D_>, dim3(mblk(maxfaces), nboxes, dim + 1), dim3(MTPB), 0, s, A));
}

void launch_ml_grown(int dim, const MlGrown &A, int nboxes, int maxgrown, hipStream_t s)
{
    ML_DIM(dim, hipLaunchKernelGGL(k_ml_grown<D_>, dim3(mblk(maxgrown), nboxes, dim), dim3(MTPB), 0, s, A));
}

}  // namespace mac
