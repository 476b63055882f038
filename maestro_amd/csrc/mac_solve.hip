// mac_solve.hip -- setup kernel of the solver-level entry point mac_solve:
// FBoxLib's cell-centred solve as MAESTRO calls it through
//   mac_multigrid(mla, rh, phi, fine_flx, alpha, beta, dx, the_bc_tower,
//                 bc_comp, stencil_order, rel_eps, abs_eps)
//     Source/mac_multigrid.f90:18-71
// i.e. (alpha - D beta G) phi = rh with caller-built alpha (cells) and beta
// (faces): alpha = 0 in macproject, alpha = rho and beta = (dt/2) k/c_p in
// the thermal diffusion solve (Source/thermal_conduct.f90:316-377).  One
// pass over the caller's boxes fills level 0 of the solver: folded face
// weights, alpha, the right side with the Dirichlet boundary terms, and the
// warm start.
#include <hip/hip_runtime.h>

#include "mac_launch.h"

namespace mac {

static constexpr int STPB = 256;

__device__ __forceinline__ long sbox_index(const BoxDev &B, int ng, int fd, int dim, int i, int j, int k)
{
    const long ex = B.hi[0] - B.lo[0] + 1 + 2 * ng + (fd == 0);
    const long ey = B.hi[1] - B.lo[1] + 1 + 2 * ng + (fd == 1);
    const long li = i - B.lo[0] + ng, lj = j - B.lo[1] + ng;
    const long lk = (dim == 3) ? (k - B.lo[2] + ng) : 0;
    return li + ex * (lj + ey * lk);
}

// Per cell of every box: face weights beta with the boundary folding of the
// projection (Neumann 0, Dirichlet 2 beta: the two-point form with the
// boundary value on the face), alpha, and
//   rh' = rh + sum over Dirichlet sides of the cell (ih2_d * w) g,
// g the boundary value held in phi's ghost cell beyond that side, in the
// order x lo, x hi, y lo, ... (one fma each).  Also: phi's valid cells as
// the warm start, a non-finite flag and block compensated partial sums of
// rh' (the solvability mean of singular problems).
template <int DIM>
__global__ void __launch_bounds__(STPB) k_solve_setup(SolveArgs A)
{
    const int b = blockIdx.y;
    const BoxDev B = A.boxes[b];
    const int ex = B.hi[0] - B.lo[0] + 1, ey = B.hi[1] - B.lo[1] + 1;
    const int ez = (DIM == 3) ? B.hi[2] - B.lo[2] + 1 : 1;
    const long q = (long)blockIdx.x * STPB + threadIdx.x;
    double rsum = 0.0;
    bool bad = false;
    if (q < (long)ex * ey * ez) {
        const int i = B.lo[0] + (int)(q % ex);
        const int j = B.lo[1] + (int)((q / ex) % ey);
        const int k = (DIM == 3) ? B.lo[2] + (int)(q / ((long)ex * ey)) : 0;
        const int I = i - A.dlo[0], J = j - A.dlo[1], K = (DIM == 3) ? k - A.dlo[2] : 0;
        const int nx = A.n[0], ny = A.n[1];
        const long c = I + (long)nx * (J + (long)ny * K);
        const double *ph = A.phi[b];
        const double ih2[3] = {A.L.ih2x, A.L.ih2y, A.L.ih2z};
        double r = A.rh[b][sbox_index(B, A.ng_rh, -1, DIM, i, j, k)];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const int di = (d == 0), dj = (d == 1), dk = (d == 2);
            const int Id = (d == 0) ? I : (d == 1 ? J : K);
            const double *bt = A.beta[b * DIM + d];
            double w = bt[sbox_index(B, A.ng_beta, d, DIM, i, j, k)];
            bad |= !isfinite(w);
            if (Id == 0) {
                const int kind = A.bc.k[d][0];
                w = kind == ELL_NEU ? 0.0 : (kind == ELL_DIR ? 2.0 * w : w);
                if (kind == ELL_DIR) {
                    const double g = ph[sbox_index(B, A.ng_phi, -1, DIM, i - di, j - dj, k - dk)];
                    bad |= !isfinite(g);
                    r = fma(ih2[d], w * g, r);
                }
            }
            double *bl = (d == 0) ? A.L.bx : (d == 1 ? A.L.by : A.L.bz);
            bl[c] = w;
            if (Id == A.n[d] - 1) {
                double wh = bt[sbox_index(B, A.ng_beta, d, DIM, i + di, j + dj, k + dk)];
                bad |= !isfinite(wh);
                const int kind = A.bc.k[d][1];
                wh = kind == ELL_NEU ? 0.0 : (kind == ELL_DIR ? 2.0 * wh : wh);
                if (kind == ELL_DIR) {
                    const double g = ph[sbox_index(B, A.ng_phi, -1, DIM, i + di, j + dj, k + dk)];
                    bad |= !isfinite(g);
                    r = fma(ih2[d], wh * g, r);
                }
                if (d == 0) A.L.bxh[J + (long)ny * K] = wh;
                else if (d == 1) A.L.byh[I + (long)nx * K] = wh;
                else A.L.bzh[I + (long)nx * J] = wh;
            }
        }
        A.L.rhs[c] = r;
        bad |= !isfinite(r);
        rsum = r;
        if (A.alpha) {
            const double a = A.alpha[b][sbox_index(B, A.ng_alpha, -1, DIM, i, j, k)];
            A.L.alpha[c] = a;
            bad |= !isfinite(a);
        }
        const double p0 = ph[sbox_index(B, A.ng_phi, -1, DIM, i, j, k)];
        bad |= !isfinite(p0);
        A.L.phi[c] = p0;
    }
    if (bad) atomicOr(A.err, 1);
    // block compensated partial sum of rh'
    double s = rsum, cmp = 0.0;
    wave_comp_sum(s, cmp);
    __shared__ double ss[STPB / 64], sc[STPB / 64];
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { ss[wv] = s; sc[wv] = cmp; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double S = ss[0], Cc = sc[0];
        for (int w = 1; w < STPB / 64; ++w) comp_add(S, Cc, ss[w], sc[w]);
        const long pid = (long)b * gridDim.x + blockIdx.x;
        A.part[2 * pid] = S;
        A.part[2 * pid + 1] = Cc;
    }
}

void launch_solve_setup(const SolveArgs &A, int maxcells, hipStream_t s)
{
    const dim3 g((unsigned)((maxcells + STPB - 1) / STPB), (unsigned)A.nboxes, 1);
    if (A.dim == 3)
        hipLaunchKernelGGL(k_solve_setup<3>, g, dim3(STPB), 0, s, A);
    else
        hipLaunchKernelGGL(k_solve_setup<2>, g, dim3(STPB), 0, s, A);
}

}  // namespace mac
