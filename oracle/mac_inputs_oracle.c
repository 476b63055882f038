/*
 * mac_inputs_oracle.c -- CPU restatement (FP64) of the producers of the MAC
 * projection's inputs: cell_to_edge, make_macrhs and make_s0mac.
 *
 * TEST INFRASTRUCTURE ONLY: the checker the GPU kernels are compared with.
 * Built by oracle/Makefile with -ffp-contract=off, so every expression is
 * evaluated in the reference's left-to-right order with one rounding per
 * operation, as in the HIP kernels (built the same way).
 *
 * Arrays here are whole-domain, C order with x fastest (cells (nz, ny, nx)).
 */
#include <math.h>
#include <stddef.h>

#include "mac_oracle.h"

/* cell_to_edge, one level, planar (Source/cell_to_edge.f90:35-62):
 * edge 0 and nr copy the end cells, edges 1 and nr-1 average their two
 * cells, the others take the fourth-order average limited to the range of
 * their two cells. */
void orc_cell_to_edge(const double *c, int nr, double *e)
{
    for (int r = 0; r <= nr; ++r) {
        double v;
        if (r == 0) {
            v = c[0];
        } else if (r == 1) {
            v = 0.5 * (c[r - 1] + c[r]);
        } else if (r == nr) {
            v = c[r - 1];
        } else if (r == nr - 1) {
            v = 0.5 * (c[r] + c[r - 1]);
        } else {
            const double tmp = 7.0 / 12.0 * (c[r] + c[r - 1]) - 1.0 / 12.0 * (c[r + 1] + c[r - 2]);
            const double lo = fmin(c[r], c[r - 1]), hi = fmax(c[r], c[r - 1]);
            v = fmin(fmax(tmp, lo), hi);
        }
        e[r] = v;
    }
}

/* quad_interp (Source/fill_3d_data.f90:535-546): the parabola through
 * (x0,y0), (x1,y1), (x2,y2) at x, limited to the range of y0, y1, y2 */
static double quad_interp(double x, double x0, double x1, double x2, double y0, double y1, double y2)
{
    double y = y0 + (y1 - y0) / (x1 - x0) * (x - x0) +
               ((y2 - y1) / (x2 - x1) - (y1 - y0) / (x1 - x0)) / (x2 - x0) * (x - x0) * (x - x1);
    const double mx = fmax(fmax(y0, y1), y2), mn = fmin(fmin(y0, y1), y2);
    if (y > mx) y = mx;
    if (y < mn) y = mn;
    return y;
}

/* s0 (radial bins of width dr, centres (i + 1/2) dr, i < nr) at the centre
 * of cell (i, j, k): put_1d_array_on_cart_3d_sphr, cell-centred input,
 * scalar output (Source/fill_3d_data.f90:406-531); type 1 piecewise
 * constant, 2 piecewise linear, 3 quadratic */
double orc_sphr_cell_value(const double *s0, int nr, double dr, const double *center, const double *prob_lo,
                           const double *dx, int type, int i, int j, int k)
{
    const double x = prob_lo[0] + ((double)i + 0.5) * dx[0] - center[0];
    const double y = prob_lo[1] + ((double)j + 0.5) * dx[1] - center[1];
    const double z = prob_lo[2] + ((double)k + 0.5) * dx[2] - center[2];
    const double radius = sqrt(x * x + y * y + z * z);
    int index = (int)(radius / dr);
    if (type == 1) return s0[index];
    if (type == 2) {
        const double rc = ((double)index + 0.5) * dr;
        if (radius >= rc) {
            if (index >= nr - 1) return s0[nr - 1];
            const double rc1 = ((double)(index + 1) + 0.5) * dr;
            return s0[index + 1] * (radius - rc) / dr + s0[index] * (rc1 - radius) / dr;
        }
        if (index == 0) return s0[index];
        if (index > nr - 1) return s0[nr - 1];
        const double rcm = ((double)(index - 1) + 0.5) * dr;
        return s0[index] * (radius - rcm) / dr + s0[index - 1] * (rc - radius) / dr;
    }
    if (index == 0) index = 1;
    else if (index >= nr - 1) index = nr - 2;
    return quad_interp(radius, ((double)(index - 1) + 0.5) * dr, ((double)index + 0.5) * dr,
                       ((double)(index + 1) + 0.5) * dr, s0[index - 1], s0[index], s0[index + 1]);
}

/* make_macrhs (Source/make_macrhs.f90): planar (dim 2 or 3, 1-D arrays
 * indexed by the last coordinate, make_macrhs_2d/3d :149-279) when sph_dr <=
 * 0, else spherical 3D (make_macrhs_3d_sphr :281-340; 1-D arrays in radial
 * bins, put on the cells by orc_sphr_cell_value).  delta_p / delta_chi may
 * be NULL when the dpdt term is off (dpdt_factor > 0 planar, >= 0
 * spherical, as in the reference). */
void orc_make_macrhs(int dim, const int *n, const double *S, const double *dg, const double *Sbar,
                     const double *beta0, const double *gamma1bar, const double *p0, const double *rho0,
                     const double *dp, double *dchi, double dt, double dpdt_factor, int base_cutoff_coord,
                     double base_cutoff_density, int is_predictor, double sph_dr, int sph_nr,
                     const double *center, const double *prob_lo, const double *dx, int interp, double *rhs)
{
    const int nx = n[0], ny = n[1], nz = dim == 3 ? n[2] : 1;
    const int sph = sph_dr > 0.0;
    const int dpdt_on = sph ? (dpdt_factor >= 0.0) : (dpdt_factor > 0.0);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const long c = i + (long)nx * (j + (long)ny * k);
                double b0, sb, g1 = 0.0, pp = 0.0, r0 = 0.0;
                int below;
                if (sph) {
                    b0 = orc_sphr_cell_value(beta0, sph_nr, sph_dr, center, prob_lo, dx, interp, i, j, k);
                    sb = orc_sphr_cell_value(Sbar, sph_nr, sph_dr, center, prob_lo, dx, interp, i, j, k);
                    if (dpdt_on) {
                        g1 = orc_sphr_cell_value(gamma1bar, sph_nr, sph_dr, center, prob_lo, dx, interp, i, j, k);
                        pp = orc_sphr_cell_value(p0, sph_nr, sph_dr, center, prob_lo, dx, interp, i, j, k);
                        r0 = orc_sphr_cell_value(rho0, sph_nr, sph_dr, center, prob_lo, dx, interp, i, j, k);
                    }
                    below = r0 > base_cutoff_density;
                } else {
                    const int r = dim == 3 ? k : j;
                    b0 = beta0[r];
                    sb = Sbar[r];
                    if (dpdt_on) {
                        g1 = gamma1bar[r];
                        pp = p0[r];
                    }
                    below = r < base_cutoff_coord;
                }
                double v = b0 * (S[c] - sb + dg[c]);
                if (dpdt_on) {
                    double x = is_predictor ? 0.0 : dchi[c];
                    if (below) {
                        x = x + dpdt_factor * dp[c] / (dt * g1 * pp);
                        v = v + b0 * x;
                    }
                    dchi[c] = x;
                }
                rhs[c] = v;
            }
}

/* make_s0mac, s0mac_interp_type 1 (Source/fill_3d_data.f90:942-1075): s0
 * on the cell centres (orc_sphr_cell_value), cells outside the domain by
 * first-order extrapolation (the nearest domain cell), then each face the
 * mean of its two cells.  Output: faces of direction d over the grown range
 * (face index -1 .. n_d + 1 along d, -1 .. n_e along the others), C order
 * with x fastest: extents (nx+3, ny+2, nz+2) for x faces and so on. */
void orc_make_s0mac(const int *n, const double *s0, int nr, double dr, const double *center,
                    const double *prob_lo, const double *dx, int interp, double *fx, double *fy, double *fz)
{
    double *f[3] = {fx, fy, fz};
    for (int d = 0; d < 3; ++d) {
        int e[3];
        for (int q = 0; q < 3; ++q) e[q] = n[q] + 2 + (q == d);
        for (int kk = 0; kk < e[2]; ++kk)
            for (int jj = 0; jj < e[1]; ++jj)
                for (int ii = 0; ii < e[0]; ++ii) {
                    const int g[3] = {ii - 1, jj - 1, kk - 1};
                    double v[2];
                    for (int s = 0; s < 2; ++s) {
                        int c[3] = {g[0], g[1], g[2]};
                        c[d] -= s;   /* s = 0: the face's high cell, 1: its low cell */
                        for (int q = 0; q < 3; ++q) c[q] = c[q] < 0 ? 0 : (c[q] > n[q] - 1 ? n[q] - 1 : c[q]);
                        v[s] = orc_sphr_cell_value(s0, nr, dr, center, prob_lo, dx, interp, c[0], c[1], c[2]);
                    }
                    f[d][ii + (long)e[0] * (jj + (long)e[1] * kk)] = 0.5 * (v[0] + v[1]);
                }
    }
}
