// mac_inputs.hip -- gfx950 kernels producing the MAC projection's inputs:
//   cell_to_edge  Source/cell_to_edge.f90:14-68        (beta0 on radial edges)
//   make_macrhs   Source/make_macrhs.f90:24-342        (the projection's macrhs)
//   make_s0mac    Source/fill_3d_data.f90:942-1075     (spherical beta0 on faces,
//                 s0mac_interp_type 1, with put_1d_array_on_cart_3d_sphr
//                 :269-533 for the cell values)
// Each is one streaming pass over the per-box arrays (Fortran order with the
// caller's ghost widths), with the reference's expression order.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mac_launch.h"

namespace mac {

static constexpr int ITPB = 256;

__device__ __forceinline__ long box_index(const BoxDev &B, int ng, int fd, int dim, int i, int j, int k)
{
    const long ex = B.hi[0] - B.lo[0] + 1 + 2 * ng + (fd == 0);
    const long ey = B.hi[1] - B.lo[1] + 1 + 2 * ng + (fd == 1);
    const long li = i - B.lo[0] + ng, lj = j - B.lo[1] + ng;
    const long lk = (dim == 3) ? (k - B.lo[2] + ng) : 0;
    return li + ex * (lj + ey * lk);
}

// ---------------------------------------------------------- cell_to_edge
__global__ void __launch_bounds__(ITPB) k_cell_to_edge(const double *__restrict__ c, int nr, double *__restrict__ e)
{
    for (int r = blockIdx.x * ITPB + threadIdx.x; r <= nr; r += gridDim.x * ITPB) {
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

void launch_cell_to_edge(const double *c, int nr, double *e, hipStream_t s)
{
    hipLaunchKernelGGL(k_cell_to_edge, dim3((nr + ITPB) / ITPB), dim3(ITPB), 0, s, c, nr, e);
}

// ------------------------------------- 1-D radial profile on a 3D cell
// put_1d_array_on_cart_3d_sphr for a bin-centred, scalar profile: type 1
// piecewise constant, 2 piecewise linear, 3 quadratic (quad_interp,
// fill_3d_data.f90:535-546, limited to the range of its three points)
__device__ double sphr_cell_value(const double *__restrict__ s0, const SphrGeom &G, int i, int j, int k)
{
    const double x = G.prob_lo[0] + ((double)i + 0.5) * G.dx[0] - G.center[0];
    const double y = G.prob_lo[1] + ((double)j + 0.5) * G.dx[1] - G.center[1];
    const double z = G.prob_lo[2] + ((double)k + 0.5) * G.dx[2] - G.center[2];
    const double radius = sqrt(x * x + y * y + z * z);
    const double dr = G.dr;
    const int nr = G.nr;
    int index = (int)(radius / dr);
    if (G.interp == 1) return s0[index];
    if (G.interp == 2) {
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
    const double x0 = ((double)(index - 1) + 0.5) * dr, x1 = ((double)index + 0.5) * dr,
                 x2 = ((double)(index + 1) + 0.5) * dr;
    const double y0 = s0[index - 1], y1 = s0[index], y2 = s0[index + 1];
    double v = y0 + (y1 - y0) / (x1 - x0) * (radius - x0) +
               ((y2 - y1) / (x2 - x1) - (y1 - y0) / (x1 - x0)) / (x2 - x0) * (radius - x0) * (radius - x1);
    const double mx = fmax(fmax(y0, y1), y2), mn = fmin(fmin(y0, y1), y2);
    if (v > mx) v = mx;
    if (v < mn) v = mn;
    return v;
}

// ----------------------------------------------------------- make_macrhs
// one thread per valid cell of box blockIdx.y
template <int DIM, bool SPH>
__global__ void __launch_bounds__(ITPB) k_make_macrhs(MacrhsArgs A)
{
    const int b = blockIdx.y;
    const BoxDev B = A.boxes[b];
    const int ex = B.hi[0] - B.lo[0] + 1, ey = B.hi[1] - B.lo[1] + 1;
    const int ez = (DIM == 3) ? B.hi[2] - B.lo[2] + 1 : 1;
    const long q = (long)blockIdx.x * ITPB + threadIdx.x;
    if (q >= (long)ex * ey * ez) return;
    const int i = B.lo[0] + (int)(q % ex), j = B.lo[1] + (int)((q / ex) % ey);
    const int k = (DIM == 3) ? B.lo[2] + (int)(q / ((long)ex * ey)) : 0;
    const int I = i - A.dlo[0], J = j - A.dlo[1], K = (DIM == 3) ? k - A.dlo[2] : 0;
    double b0, sb, g1 = 0.0, pp = 0.0;
    bool below;
    if constexpr (SPH) {
        b0 = sphr_cell_value(A.beta0, A.geo, I, J, K);
        sb = sphr_cell_value(A.Sbar, A.geo, I, J, K);
        double r0 = 0.0;
        if (A.dpdt_on) {
            g1 = sphr_cell_value(A.gamma1bar, A.geo, I, J, K);
            pp = sphr_cell_value(A.p0, A.geo, I, J, K);
            r0 = sphr_cell_value(A.rho0, A.geo, I, J, K);
        }
        below = r0 > A.cutoff_density;
    } else {
        const int r = (DIM == 3) ? K : J;
        b0 = A.beta0[r];
        sb = A.Sbar[r];
        if (A.dpdt_on) {
            g1 = A.gamma1bar[r];
            pp = A.p0[r];
        }
        below = r < A.cutoff_coord;
    }
    const double S = A.S[b][box_index(B, A.ng_S, -1, DIM, i, j, k)];
    const double dg = A.dg[b][box_index(B, A.ng_dg, -1, DIM, i, j, k)];
    double v = b0 * (S - sb + dg);
    if (A.dpdt_on) {
        double *dchi = A.dchi[b] + box_index(B, A.ng_dchi, -1, DIM, i, j, k);
        double x = A.predictor ? 0.0 : *dchi;
        if (below) {
            const double dp = A.dp[b][box_index(B, A.ng_dp, -1, DIM, i, j, k)];
            x = x + A.dpdt * dp / (A.dt * g1 * pp);
            v = v + b0 * x;
        }
        *dchi = x;
    }
    A.rhs[b][box_index(B, A.ng_rhs, -1, DIM, i, j, k)] = v;
}

// the reference zeroes a predictor's delta_chi over the whole array, ghost
// cells included; the kernel above covers the valid cells
__global__ void __launch_bounds__(ITPB) k_zero_ghosts(MacrhsArgs A, int dim)
{
    const int b = blockIdx.y;
    const BoxDev B = A.boxes[b];
    const int ng = A.ng_dchi;
    const long ex = B.hi[0] - B.lo[0] + 1 + 2 * ng, ey = B.hi[1] - B.lo[1] + 1 + 2 * ng;
    const long ez = (dim == 3) ? B.hi[2] - B.lo[2] + 1 + 2 * ng : 1;
    for (long q = (long)blockIdx.x * ITPB + threadIdx.x; q < ex * ey * ez; q += (long)gridDim.x * ITPB) {
        const long li = q % ex, lj = (q / ex) % ey, lk = q / (ex * ey);
        const bool ghost = li < ng || li >= ex - ng || lj < ng || lj >= ey - ng ||
                           (dim == 3 && (lk < ng || lk >= ez - ng));
        if (ghost) A.dchi[b][q] = 0.0;
    }
}

void launch_make_macrhs(const MacrhsArgs &A, int maxcells, bool sph, hipStream_t s)
{
    dim3 g((unsigned)((maxcells + ITPB - 1) / ITPB), A.nboxes, 1);
    if (A.dim == 3) {
        if (sph) hipLaunchKernelGGL((k_make_macrhs<3, true>), g, dim3(ITPB), 0, s, A);
        else hipLaunchKernelGGL((k_make_macrhs<3, false>), g, dim3(ITPB), 0, s, A);
    } else {
        hipLaunchKernelGGL((k_make_macrhs<2, false>), g, dim3(ITPB), 0, s, A);
    }
    if (A.dpdt_on && A.predictor && A.ng_dchi > 0)
        hipLaunchKernelGGL(k_zero_ghosts, dim3(64, A.nboxes, 1), dim3(ITPB), 0, s, A, A.dim);
}

// ------------------------------------------------------------ make_s0mac
// (1) the profile on every cell of the domain (whole-domain scratch, C order
// with x fastest), one interpolation per cell
__global__ void __launch_bounds__(ITPB) k_s0_cells(S0macArgs A)
{
    const long nc = (long)A.n[0] * A.n[1] * A.n[2];
    for (long q = (long)blockIdx.x * ITPB + threadIdx.x; q < nc; q += (long)gridDim.x * ITPB) {
        const int i = (int)(q % A.n[0]), j = (int)((q / A.n[0]) % A.n[1]), k = (int)(q / ((long)A.n[0] * A.n[1]));
        A.cells[q] = sphr_cell_value(A.s0, A.geo, i, j, k);
    }
}

// (2) one thread per element of box blockIdx.y's face array of direction
// blockIdx.z (one ghost layer: every face of the grown box): the mean of the
// face's two cells, clamped into the domain (first-order extrapolation of
// the cell values outside it)
__global__ void __launch_bounds__(ITPB) k_make_s0mac(S0macArgs A)
{
    const int b = blockIdx.y, d = blockIdx.z;
    const BoxDev B = A.boxes[b];
    const int ex = B.hi[0] - B.lo[0] + 3 + (d == 0), ey = B.hi[1] - B.lo[1] + 3 + (d == 1);
    const int ez = B.hi[2] - B.lo[2] + 3 + (d == 2);
    const long q = (long)blockIdx.x * ITPB + threadIdx.x;
    if (q >= (long)ex * ey * ez) return;
    const int g[3] = {B.lo[0] - 1 + (int)(q % ex) - A.dlo[0], B.lo[1] - 1 + (int)((q / ex) % ey) - A.dlo[1],
                      B.lo[2] - 1 + (int)(q / ((long)ex * ey)) - A.dlo[2]};
    double v[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        int c[3] = {g[0], g[1], g[2]};
        c[d] -= s;   // s = 0: the face's high cell, 1: its low cell
#pragma unroll
        for (int e = 0; e < 3; ++e) c[e] = c[e] < 0 ? 0 : (c[e] > A.n[e] - 1 ? A.n[e] - 1 : c[e]);
        v[s] = A.cells[c[0] + (long)A.n[0] * (c[1] + (long)A.n[1] * c[2])];
    }
    A.out[b * 3 + d][q] = 0.5 * (v[0] + v[1]);
}

void launch_make_s0mac(const S0macArgs &A, int maxfaces, hipStream_t s)
{
    const long nc = (long)A.n[0] * A.n[1] * A.n[2];
    hipLaunchKernelGGL(k_s0_cells, dim3((unsigned)std::min<long>((nc + ITPB - 1) / ITPB, 65536)), dim3(ITPB), 0, s,
                       A);
    dim3 g((unsigned)((maxfaces + ITPB - 1) / ITPB), A.nboxes, 3);
    hipLaunchKernelGGL(k_make_s0mac, g, dim3(ITPB), 0, s, A);
}

}  // namespace mac
