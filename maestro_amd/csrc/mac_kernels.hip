// mac_kernels.hip -- gfx950 kernels of the MAC projection (libmaestro_mac.so).
//
// Reference routines each kernel replaces (Source/macproject.f90 unless noted):
//   k_setup            mult_edge_by_1d_coeff (:577-609), divumac_3d (:457-480),
//                      mk_mac_coeffs_3d (:698-738), the beta0 scaling of the
//                      coefficients (:185-203) and the copy of phi into the
//                      solver, fused into one pass over the caller's boxes
//   k_mean_*           solvability: compensated FP64 mean of rh (singular BCs)
//   k_coarsen          coarse-level face coefficients (FBoxLib stencil fill,
//                      Docs/mg/mg.tex:333-336; averaging rule of this build)
//   k_gsrb             red-black Gauss-Seidel (MG_SMOOTHER_GS_RB, mg.tex:201)
//   k_residual         compute_defect (mg.tex:351)
//   k_restrict         conservative-average restriction (mg.tex:353-354)
//   k_prolong          linear prolongation + correction (mg.tex:363-365)
//   k_coarse_cycle     the cycle (V, W or F) on the coarse levels plus the
//                      FP64 CG bottom solve (mg.tex:221-262) in one workgroup
//   k_correct_valid    mkumac_3d (:908-975) fused with the division by beta0
//                      (:232-236)
//   k_fill_ghost       multifab_fill_boundary of umac (:245-251)
//   k_physbc           multifab_physbc_edgevel (:254; rules of the in-tree
//                      sample Exec/SCIENCE/flame/multifab_physbc_edgevel.f90:
//                      211-321)
#include <algorithm>
#include <type_traits>

#include "mac_launch.h"

namespace mac {

static constexpr int TPB = 256;

// ------------------------------------------------------ per-box indexing
__device__ __forceinline__ long boxoff(const BoxDev &B, int ng, int fd, int dim, int i, int j,
                                       int k)
{
    long ex = B.hi[0] - B.lo[0] + 1 + 2 * ng + (fd == 0);
    long ey = B.hi[1] - B.lo[1] + 1 + 2 * ng + (fd == 1);
    long li = i - B.lo[0] + ng, lj = j - B.lo[1] + ng;
    long lk = (dim == 3) ? (k - B.lo[2] + ng) : 0;
    return li + ex * (lj + ey * lk);
}

__device__ __forceinline__ double fold_bc(double beta, int kind)
{
    if (kind == ELL_NEU) return 0.0;
    if (kind == ELL_DIR) return 2.0 * beta;
    return beta;
}

// A planar beta0 seen from one plane of a column: beta0(r) on the tangential
// faces, beta0_edge(r) and beta0_edge(r + 1) on the radial low and high faces
// (r follows k in 3D, j in 2D); ones without a beta0.  The column kernels load
// it a plane ahead like the caller arrays: a load issued at its use stalls the
// march.
struct B0 {
    double t, lo, hi;
};
__device__ __forceinline__ B0 planar_b0(const double *b0, const double *b0e, int r)
{
    B0 o = {1.0, 1.0, 1.0};
    if (b0) {
        o.t = b0[r];
        o.lo = b0e[r];
        o.hi = b0e[r + 1];
    }
    return o;
}

// ------------------------------------------------- column addressing
// A box array (ghost width ng, nodal in direction fd, or -1 for cells) seen
// from the column (li, lj) of the box's valid region: the element offset of
// the column's first valid plane and the stride between planes.  32-bit: a
// box array holds fewer than 2^31 elements.
struct Col {
    int off, sy, pz;   // first valid element of the column, row and plane strides
    // element offset of the neighbour (di, dj, dk) of the column's plane kk
    __device__ __forceinline__ int at(int kk, int di = 0, int dj = 0, int dk = 0) const
    {
        return off + (kk + dk) * pz + dj * sy + di;
    }
};
template <int DIM>
__device__ __forceinline__ Col col_of(const BoxDev &B, int ng, int fd, int li, int lj)
{
    const int ex = B.hi[0] - B.lo[0] + 1 + 2 * ng + (fd == 0);
    const int ey = B.hi[1] - B.lo[1] + 1 + 2 * ng + (fd == 1);
    Col c;
    c.sy = ex;
    c.pz = ex * ey;
    c.off = (li + ng) + ex * (lj + ng) + (DIM == 3 ? c.pz * ng : 0);
    return c;
}

// Planes of a box that one thread of the column kernels marches through
// (3D; the grid's z dimension counts the chunks).  Eight keep some 32K waves
// in flight at 256^3 while the carried values save all but 1/8 of the
// re-reads of the plane below.
static constexpr int ZCH = 8;

// ------------------------------------------------------------- k_setup
// One thread per (i, j) column of a box marching k through a chunk of ZCH
// planes (3D): the scaled z faces and rho of the plane below are carried in
// registers, so each caller array element is read once from HBM (the x and y
// neighbours come from the same cache lines the neighbouring lanes read),
// and the indexing is 32-bit.  Per cell: rh = macrhs - D(beta0 U), the folded face
// coefficients 2/(rho_L + rho_R) beta0^p (with the domain's high faces into
// the slabs) and the warm start.  The column's rh values are summed
// compensated, then per block (the solvability mean).
template <int DIM>
__global__ void __launch_bounds__(TPB) k_setup(SetupArgs A)
{
    const int b = blockIdx.y;
    const BoxDev B = A.boxes[b];
    const int ex = B.hi[0] - B.lo[0] + 1, ey = B.hi[1] - B.lo[1] + 1;
    const int ez = (DIM == 3) ? B.hi[2] - B.lo[2] + 1 : 1;
    const int q = blockIdx.x * TPB + threadIdx.x;
    const int kz0 = (DIM == 3) ? (int)blockIdx.z * ZCH : 0, kz1 = (DIM == 3) ? min(kz0 + ZCH, ez) : 1;
    double s = 0.0, cmp = 0.0;
    bool bad = false;
    if (q < ex * ey && kz0 < kz1) {
        const int li = q % ex, lj = q / ex;
        const int i = B.lo[0] + li, j = B.lo[1] + lj;
        const int I = i - A.dlo[0], J = j - A.dlo[1];
        const int nx = A.n[0], ny = A.n[1];
        const int rd = DIM - 1;   // radial direction of a planar beta0
        const Col cr = col_of<DIM>(B, A.ng_rho, -1, li, lj), cp = col_of<DIM>(B, A.ng_phi, -1, li, lj);
        const Col cm = col_of<DIM>(B, A.ng_rhs, -1, li, lj);
        Col cu[3];
#pragma unroll
        for (int d = 0; d < DIM; ++d) cu[d] = col_of<DIM>(B, A.ng_u, d, li, lj);
        const double *__restrict__ rho = A.rho[b];
        const double *__restrict__ phi = A.phi[b];
        const double *__restrict__ mrh = A.macrhs ? A.macrhs[b] : nullptr;
        double *__restrict__ lrhs = A.L.rhs;
        double *__restrict__ lphi = A.L.phi;
        const double *u[3];
        const double *cart[3];
        Col cc[3];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            u[d] = A.umac[b * DIM + d];
            cart[d] = A.cart ? A.cart[b * DIM + d] : nullptr;
            cc[d] = col_of<DIM>(B, A.ng_cart, d, li, lj);
        }
        auto b0load = [&](int kk) -> B0 {
            return planar_b0(A.beta0, A.beta0e, (rd == 2 ? B.lo[2] + kk : j) - A.dlo[rd]);
        };
        // beta0 of face s (0 low, 1 high) of direction d at plane kk, whose
        // planar values are pb; has: any
        auto b0f = [&](int d, int sd, int kk, const B0 &pb, bool &has) -> double {
            if (A.beta0) {
                has = true;
                return (d == rd) ? (sd ? pb.hi : pb.lo) : pb.t;
            }
            if (A.cart) {
                has = true;
                return cart[d][cc[d].at(kk, d == 0 ? sd : 0, d == 1 ? sd : 0, d == 2 ? sd : 0)];
            }
            has = false;
            return 1.0;
        };
        // the caller values one plane needs (raw face velocities: [0, 1] x low
        // and high, [2, 3] y, [4] the z high face; rho of the cell and of its
        // x and y low neighbours; macrhs; the warm start).  The next plane's
        // are loaded before this plane's results are stored, so the march
        // never waits on its own stores.
        struct PIn {
            double u[5], r, rx, ry, m, p;
        };
        auto load = [&](int kk) -> PIn {
            PIn o;
            o.u[0] = u[0][cu[0].at(kk)];
            o.u[1] = u[0][cu[0].at(kk, 1)];
            o.u[2] = u[1][cu[1].at(kk)];
            o.u[3] = u[1][cu[1].at(kk, 0, 1)];
            o.u[4] = (DIM == 3) ? u[2][cu[2].at(kk, 0, 0, 1)] : 0.0;
            o.r = rho[cr.at(kk)];
            o.rx = rho[cr.at(kk, -1)];
            o.ry = rho[cr.at(kk, 0, -1)];
            o.m = mrh ? mrh[cm.at(kk)] : 0.0;
            o.p = phi[cp.at(kk)];
            return o;
        };
        // scaled z faces and rho of the previous plane (3D carries)
        double uz_lo = 0.0, r_below = 0.0;
        B0 bcur = b0load(kz0);
        if (DIM == 3) {
            bool has;
            const double b0 = b0f(2, 0, kz0, bcur, has);
            uz_lo = u[2][cu[2].at(kz0)];
            if (has) uz_lo = uz_lo * b0;
            r_below = rho[cr.at(kz0, 0, 0, -1)];
        }
        PIn cur = load(kz0);
        for (int kk = kz0; kk < kz1; ++kk) {
            const PIn nxt = load(min(kk + 1, kz1 - 1));
            const B0 bnxt = b0load(min(kk + 1, kz1 - 1));
            const int k = (DIM == 3) ? B.lo[2] + kk : 0, K = (DIM == 3) ? k - A.dlo[2] : 0;
            const int c = I + nx * (J + ny * K);
            double uf[DIM][2];
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
#pragma unroll
                for (int sd = 0; sd < 2; ++sd) {
                    if (DIM == 3 && d == 2 && sd == 0) {
                        uf[2][0] = uz_lo;
                        continue;
                    }
                    double v = cur.u[d == 2 ? 4 : 2 * d + sd];
                    bool has;
                    const double b0 = b0f(d, sd, kk, bcur, has);
                    if (has) v = v * b0;
                    uf[d][sd] = v;
                }
            }
            double D = (uf[0][1] - uf[0][0]) / A.dx[0] + (uf[1][1] - uf[1][0]) / A.dx[1];
            if (DIM == 3) D = D + (uf[2][1] - uf[2][0]) / A.dx[2];
            double rh;
            if (mrh) rh = -(D - cur.m);
            else rh = -D;
            lrhs[c] = rh;
            bad |= !isfinite(rh);
            comp_add(s, cmp, rh, 0.0);
            // ---- face coefficients (mk_mac_coeffs, beta0 scaling, BC folding)
            const double rc = cur.r;
            if (A.rho32) A.rho32[c] = (float)rc;   // the FP32 solver's rho form
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                const int Id = (d == 0) ? I : (d == 1 ? J : K);
                const double rl = (DIM == 3 && d == 2) ? r_below : (d == 0 ? cur.rx : cur.ry);
                double beta = 2.0 / (rc + rl);
                bool has;
                const double b0 = b0f(d, 0, kk, bcur, has);
                if (has) {
                    beta = beta * b0;
                    if (A.alt) beta = beta * b0;
                }
                if (Id == 0) beta = fold_bc(beta, A.bc.k[d][0]);
                bad |= !isfinite(beta);
                double *bl = (d == 0) ? A.L.bx : (d == 1 ? A.L.by : A.L.bz);
                bl[c] = beta;
                if (Id == A.n[d] - 1) {
                    double bh = 2.0 / (rho[cr.at(kk, d == 0, d == 1, d == 2)] + rc);
                    const double b0h = b0f(d, 1, kk, bcur, has);
                    if (has) {
                        bh = bh * b0h;
                        if (A.alt) bh = bh * b0h;
                    }
                    bh = fold_bc(bh, A.bc.k[d][1]);
                    bad |= !isfinite(bh);
                    if (d == 0) A.L.bxh[J + ny * K] = bh;
                    else if (d == 1) A.L.byh[I + nx * K] = bh;
                    else A.L.bzh[I + nx * J] = bh;
                }
            }
            // ---- warm start
            bad |= !isfinite(cur.p);
            lphi[c] = cur.p;
            if (DIM == 3) {
                uz_lo = uf[2][1];
                r_below = rc;
            }
            cur = nxt;
            bcur = bnxt;
        }
    }
    if (bad) atomicOr(A.err, 1);
    // ---- block compensated partial sum of rh (solvability mean)
    wave_comp_sum(s, cmp);
    __shared__ double ss[TPB / 64], sc[TPB / 64];
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { ss[wv] = s; sc[wv] = cmp; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double S = ss[0], Cc = sc[0];
        for (int w = 1; w < TPB / 64; ++w) comp_add(S, Cc, ss[w], sc[w]);
        long pid = ((long)b * gridDim.z + blockIdx.z) * gridDim.x + blockIdx.x;
        A.part[2 * pid] = S;
        A.part[2 * pid + 1] = Cc;
    }
}

// ------------------------------------ compensated mean of partial sums
__global__ void __launch_bounds__(1024) k_mean_finalize(const double *part, long nparts, double ncell,
                                                        double *mean_out)
{
    // batches of 16 pairs per thread are loaded before they are added, so
    // the loads overlap instead of waiting one by one
    constexpr int BT = 16;
    double s = 0.0, c = 0.0;
    const long step = (long)blockDim.x;
    for (long p0 = threadIdx.x; p0 < nparts; p0 += BT * step) {
        double vs[BT], vc[BT];
#pragma unroll
        for (int t = 0; t < BT; ++t) {
            const long p = p0 + t * step;
            vs[t] = p < nparts ? part[2 * p] : 0.0;
            vc[t] = p < nparts ? part[2 * p + 1] : 0.0;
        }
#pragma unroll
        for (int t = 0; t < BT; ++t) comp_add(s, c, vs[t], vc[t]);
    }
    wave_comp_sum(s, c);
    __shared__ double ss[16], sc[16];
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { ss[wv] = s; sc[wv] = c; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double S = ss[0], C = sc[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) comp_add(S, C, ss[w], sc[w]);
        *mean_out = (S + C) / ncell;
    }
}

template <typename T>
__global__ void __launch_bounds__(TPB) k_sub_mean(T *x, long n, const double *mean)
{
    long q = (long)blockIdx.x * TPB + threadIdx.x;
    if (q < n) x[q] = (T)((double)x[q] - *mean);
}

// ------------------------------------------------- block partial sums
template <typename T>
__global__ void __launch_bounds__(TPB) k_partial_sum(const T *x, long n, double *part)
{
    long q = (long)blockIdx.x * TPB + threadIdx.x;
    double s = (q < n) ? (double)x[q] : 0.0, c = 0.0;
    wave_comp_sum(s, c);
    __shared__ double ss[TPB / 64], sc[TPB / 64];
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { ss[wv] = s; sc[wv] = c; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double S = ss[0], C = sc[0];
        for (int w = 1; w < TPB / 64; ++w) comp_add(S, C, ss[w], sc[w]);
        part[2 * blockIdx.x] = S;
        part[2 * blockIdx.x + 1] = C;
    }
}

// ------------------------------------------------ coarse coefficients
// coarse face = average of the 2^(dim-1) fine faces covering it
template <int DIM, typename T>
__global__ void __launch_bounds__(TPB) k_coarsen(Lev<T> F, Lev<T> C)
{
    long q = (long)blockIdx.x * TPB + threadIdx.x;
    long nc = (long)C.nx * C.ny * C.nz;
    if (q >= nc) return;
    int I = (int)(q % C.nx), J = (int)((q / C.nx) % C.ny), K = (int)(q / ((long)C.nx * C.ny));
    int i = 2 * I, j = 2 * J, k = (DIM == 3) ? 2 * K : 0;
    auto fc = [&](int ii, int jj, int kk) { return ii + (long)F.nx * (jj + (long)F.ny * kk); };
    if (C.alpha) {   // cell coefficient: mean of the 2^DIM fine cells (restriction's sum order)
        const T *a = F.alpha;
        T v = (a[fc(i, j, k)] + a[fc(i + 1, j, k)]) + (a[fc(i, j + 1, k)] + a[fc(i + 1, j + 1, k)]);
        if (DIM == 3)
            C.alpha[q] = T(0.125) * (v + ((a[fc(i, j, k + 1)] + a[fc(i + 1, j, k + 1)]) +
                                          (a[fc(i, j + 1, k + 1)] + a[fc(i + 1, j + 1, k + 1)])));
        else
            C.alpha[q] = T(0.25) * v;
    }
    if (DIM == 3) {
        const T q4 = T(0.25);
        C.bx[q] = q4 * ((F.bx[fc(i, j, k)] + F.bx[fc(i, j + 1, k)]) +
                        (F.bx[fc(i, j, k + 1)] + F.bx[fc(i, j + 1, k + 1)]));
        C.by[q] = q4 * ((F.by[fc(i, j, k)] + F.by[fc(i + 1, j, k)]) +
                        (F.by[fc(i, j, k + 1)] + F.by[fc(i + 1, j, k + 1)]));
        C.bz[q] = q4 * ((F.bz[fc(i, j, k)] + F.bz[fc(i + 1, j, k)]) +
                        (F.bz[fc(i, j + 1, k)] + F.bz[fc(i + 1, j + 1, k)]));
        if (I == C.nx - 1)
            C.bxh[J + (long)C.ny * K] =
                q4 * ((F.bxh[j + (long)F.ny * k] + F.bxh[j + 1 + (long)F.ny * k]) +
                      (F.bxh[j + (long)F.ny * (k + 1)] + F.bxh[j + 1 + (long)F.ny * (k + 1)]));
        if (J == C.ny - 1)
            C.byh[I + (long)C.nx * K] =
                q4 * ((F.byh[i + (long)F.nx * k] + F.byh[i + 1 + (long)F.nx * k]) +
                      (F.byh[i + (long)F.nx * (k + 1)] + F.byh[i + 1 + (long)F.nx * (k + 1)]));
        if (K == C.nz - 1)
            C.bzh[I + (long)C.nx * J] =
                q4 * ((F.bzh[i + (long)F.nx * j] + F.bzh[i + 1 + (long)F.nx * j]) +
                      (F.bzh[i + (long)F.nx * (j + 1)] + F.bzh[i + 1 + (long)F.nx * (j + 1)]));
    } else {
        const T h = T(0.5);
        C.bx[q] = h * (F.bx[fc(i, j, 0)] + F.bx[fc(i, j + 1, 0)]);
        C.by[q] = h * (F.by[fc(i, j, 0)] + F.by[fc(i + 1, j, 0)]);
        if (I == C.nx - 1) C.bxh[J] = h * (F.bxh[j] + F.bxh[j + 1]);
        if (J == C.ny - 1) C.byh[I] = h * (F.byh[i] + F.byh[i + 1]);
    }
}

// FP64 -> FP32 copy of a level's face weights (mixed precision hierarchy)
template <int DIM>
__global__ void __launch_bounds__(TPB) k_weights_to_f32(Lev<double> S, Lev<float> D)
{
    long q = (long)blockIdx.x * TPB + threadIdx.x;
    long n = (long)S.nx * S.ny * S.nz;
    if (q >= n) return;
    D.bx[q] = (float)S.bx[q];
    D.by[q] = (float)S.by[q];
    if (D.alpha) D.alpha[q] = (float)S.alpha[q];
    if (DIM == 3) D.bz[q] = (float)S.bz[q];
    int I = (int)(q % S.nx), J = (int)((q / S.nx) % S.ny), K = (int)(q / ((long)S.nx * S.ny));
    if (I == 0) D.bxh[J + (long)S.ny * K] = (float)S.bxh[J + (long)S.ny * K];
    if (J == 0) D.byh[I + (long)S.nx * K] = (float)S.byh[I + (long)S.nx * K];
    if (DIM == 3 && K == 0) D.bzh[I + (long)S.nx * J] = (float)S.bzh[I + (long)S.nx * J];
}

// Compact level 0 and FP32 level 1 from the FP64 level 0 in one pass (3D):
// one thread per coarse cell reads the weights of its 2x2x2 fine cells,
// writes them as bfloat16 of their FP32 rounding (fine column pairs as one
// 32-bit store), and writes the coarse weights with k_coarsen's FP32
// arithmetic on the same FP32 values.
__global__ void __launch_bounds__(TPB) k_compact_level0(Lev<double> F, Compact16 K, Lev<float> C)
{
    const long q = (long)blockIdx.x * TPB + threadIdx.x;
    const long nc = (long)C.nx * C.ny * C.nz;
    if (q >= nc) return;
    const int I = (int)(q % C.nx), J = (int)((q / C.nx) % C.ny), Kc = (int)(q / ((long)C.nx * C.ny));
    const int i = 2 * I, j = 2 * J, k = 2 * Kc;
    auto fc = [&](int jj, int kk) { return i + (long)F.nx * (jj + (long)F.ny * kk); };
    float wx[2][2][2], wy[2][2][2], wz[2][2][2];   // [dk][dj][di]
#pragma unroll
    for (int dk = 0; dk < 2; ++dk)
#pragma unroll
        for (int dj = 0; dj < 2; ++dj) {
            const long f = fc(j + dj, k + dk);
            const double2 bx = *(const double2 *)(F.bx + f), by = *(const double2 *)(F.by + f),
                          bz = *(const double2 *)(F.bz + f);
            wx[dk][dj][0] = (float)bx.x; wx[dk][dj][1] = (float)bx.y;
            wy[dk][dj][0] = (float)by.x; wy[dk][dj][1] = (float)by.y;
            wz[dk][dj][0] = (float)bz.x; wz[dk][dj][1] = (float)bz.y;
            *(unsigned *)(K.bx + f) = bf16_pack(wx[dk][dj][0], wx[dk][dj][1]);
            *(unsigned *)(K.by + f) = bf16_pack(wy[dk][dj][0], wy[dk][dj][1]);
            *(unsigned *)(K.bz + f) = bf16_pack(wz[dk][dj][0], wz[dk][dj][1]);
        }
    const float q4 = 0.25f;
    C.bx[q] = q4 * ((wx[0][0][0] + wx[0][1][0]) + (wx[1][0][0] + wx[1][1][0]));
    C.by[q] = q4 * ((wy[0][0][0] + wy[0][0][1]) + (wy[1][0][0] + wy[1][0][1]));
    C.bz[q] = q4 * ((wz[0][0][0] + wz[0][0][1]) + (wz[0][1][0] + wz[0][1][1]));
    if (I == C.nx - 1) {   // x-high slab entries of rows j, j+1 on planes k, k+1
        float h[2][2];
#pragma unroll
        for (int dk = 0; dk < 2; ++dk)
#pragma unroll
            for (int dj = 0; dj < 2; ++dj) {
                const long s = (j + dj) + (long)F.ny * (k + dk);
                h[dk][dj] = (float)F.bxh[s];
                K.bxh[s] = f32_to_bf16(h[dk][dj]);
            }
        C.bxh[J + (long)C.ny * Kc] = q4 * ((h[0][0] + h[0][1]) + (h[1][0] + h[1][1]));
    }
    if (J == C.ny - 1) {   // y-high slab entries of columns i, i+1 on planes k, k+1
        float h[2][2];
#pragma unroll
        for (int dk = 0; dk < 2; ++dk)
#pragma unroll
            for (int di = 0; di < 2; ++di) {
                const long s = (i + di) + (long)F.nx * (k + dk);
                h[dk][di] = (float)F.byh[s];
                K.byh[s] = f32_to_bf16(h[dk][di]);
            }
        C.byh[I + (long)C.nx * Kc] = q4 * ((h[0][0] + h[0][1]) + (h[1][0] + h[1][1]));
    }
    if (Kc == C.nz - 1) {   // z-high slab entries of columns i, i+1 on rows j, j+1
        float h[2][2];
#pragma unroll
        for (int dj = 0; dj < 2; ++dj)
#pragma unroll
            for (int di = 0; di < 2; ++di) {
                const long s = (i + di) + (long)F.nx * (j + dj);
                h[dj][di] = (float)F.bzh[s];
                K.bzh[s] = f32_to_bf16(h[dj][di]);
            }
        C.bzh[I + (long)C.nx * J] = q4 * ((h[0][0] + h[0][1]) + (h[1][0] + h[1][1]));
    }
}

// The rho form's face factors: cxy[k] = 2 beta0(k)^p on the planes, cz[f] =
// 2 beta0_edge(f)^p on the z faces (p = 2 with use_alt_energy_fix), the
// FP64 coefficient formula's factors (2 / (rho_L + rho_R) beta0 beta0)
// rounded once to FP32.
__global__ void __launch_bounds__(TPB) k_rho_coef(const double *b0, const double *b0e, int alt, int nz, RhoLev R)
{
    const int t = blockIdx.x * TPB + threadIdx.x;
    if (t < nz) R.cxy[t] = (float)(alt ? 2.0 * b0[t] * b0[t] : 2.0 * b0[t]);
    if (t <= nz) R.cz[t] = (float)(alt ? 2.0 * b0e[t] * b0e[t] : 2.0 * b0e[t]);
}

// The FP32 level 1 from the FP64 level-0 weights (one thread per coarse
// cell, the averaging of k_coarsen on the FP32-rounded fine weights) and the
// rho form's z slabs: the FP32 folded weights of the domain's low z faces
// (R.bzl) and high ones (bzh0, the FP32 level 0's slab).
__global__ void __launch_bounds__(TPB) k_rho_level0(Lev<double> F, RhoLev R, float *bzh0, Lev<float> C)
{
    const long q = (long)blockIdx.x * TPB + threadIdx.x;
    const long nc = (long)C.nx * C.ny * C.nz;
    if (q >= nc) return;
    const int I = (int)(q % C.nx), J = (int)((q / C.nx) % C.ny), Kc = (int)(q / ((long)C.nx * C.ny));
    const int i = 2 * I, j = 2 * J, k = 2 * Kc;
    auto fc = [&](int jj, int kk) { return i + (long)F.nx * (jj + (long)F.ny * kk); };
    float wx[2][2][2], wy[2][2][2], wz[2][2][2];   // [dk][dj][di]
#pragma unroll
    for (int dk = 0; dk < 2; ++dk)
#pragma unroll
        for (int dj = 0; dj < 2; ++dj) {
            const long f = fc(j + dj, k + dk);
            const double2 bx = *(const double2 *)(F.bx + f), by = *(const double2 *)(F.by + f),
                          bz = *(const double2 *)(F.bz + f);
            wx[dk][dj][0] = (float)bx.x; wx[dk][dj][1] = (float)bx.y;
            wy[dk][dj][0] = (float)by.x; wy[dk][dj][1] = (float)by.y;
            wz[dk][dj][0] = (float)bz.x; wz[dk][dj][1] = (float)bz.y;
        }
    const float q4 = 0.25f;
    C.bx[q] = q4 * ((wx[0][0][0] + wx[0][1][0]) + (wx[1][0][0] + wx[1][1][0]));
    C.by[q] = q4 * ((wy[0][0][0] + wy[0][0][1]) + (wy[1][0][0] + wy[1][0][1]));
    C.bz[q] = q4 * ((wz[0][0][0] + wz[0][0][1]) + (wz[0][1][0] + wz[0][1][1]));
    if (Kc == 0)   // the fine level's low z faces: plane 0's z weights
#pragma unroll
        for (int dj = 0; dj < 2; ++dj)
#pragma unroll
            for (int di = 0; di < 2; ++di) R.bzl[(i + di) + (long)F.nx * (j + dj)] = wz[0][dj][di];
    if (I == C.nx - 1) {
        float h[2][2];
#pragma unroll
        for (int dk = 0; dk < 2; ++dk)
#pragma unroll
            for (int dj = 0; dj < 2; ++dj) h[dk][dj] = (float)F.bxh[(j + dj) + (long)F.ny * (k + dk)];
        C.bxh[J + (long)C.ny * Kc] = q4 * ((h[0][0] + h[0][1]) + (h[1][0] + h[1][1]));
    }
    if (J == C.ny - 1) {
        float h[2][2];
#pragma unroll
        for (int dk = 0; dk < 2; ++dk)
#pragma unroll
            for (int di = 0; di < 2; ++di) h[dk][di] = (float)F.byh[(i + di) + (long)F.nx * (k + dk)];
        C.byh[I + (long)C.nx * Kc] = q4 * ((h[0][0] + h[0][1]) + (h[1][0] + h[1][1]));
    }
    if (Kc == C.nz - 1) {
        float h[2][2];
#pragma unroll
        for (int dj = 0; dj < 2; ++dj)
#pragma unroll
            for (int di = 0; di < 2; ++di) {
                const long sidx = (i + di) + (long)F.nx * (j + dj);
                h[dj][di] = (float)F.bzh[sidx];
                bzh0[sidx] = h[dj][di];
            }
        C.bzh[I + (long)C.nx * J] = q4 * ((h[0][0] + h[0][1]) + (h[1][0] + h[1][1]));
    }
}

__global__ void __launch_bounds__(TPB) k_to_bf16(const float *__restrict__ in, unsigned short *__restrict__ out,
                                                 long n)
{
    const long q = (long)blockIdx.x * TPB + threadIdx.x;
    if (q < n) out[q] = f32_to_bf16(in[q]);
}

template <typename A, typename B>
__global__ void __launch_bounds__(TPB) k_convert(const A *__restrict__ in, B *__restrict__ out, long n)
{
    long q = (long)blockIdx.x * TPB + threadIdx.x;
    if (q < n) out[q] = (B)in[q];
}

// --------------------------------------------------------------- GSRB
// one colour of a red-black sweep; thread per (i-pair, j, k)
template <int DIM, typename T, bool SOR = false>
__global__ void __launch_bounds__(TPB) k_gsrb(Lev<T> L, int color)
{
    const int hx = (L.nx + 1) >> 1;
    long q = (long)blockIdx.x * TPB + threadIdx.x;
    long nq = (long)hx * L.ny * L.nz;
    if (q >= nq) return;
    int m = (int)(q % hx);
    int j = (int)((q / hx) % L.ny);
    int k = (int)(q / ((long)hx * L.ny));
    int i = 2 * m + ((j + k + color) & 1);
    if (i >= L.nx) return;
    long c = i + (long)L.nx * (j + (long)L.ny * k);
    T w[6], p[6];
    stencil<DIM, T>(L, L.phi, i, j, k, c, w, p);
    const T v = gs_value<DIM, T>(L, w, p, L.rhs[c], alpha_at(L, c));
    L.phi[c] = relax<SOR>(v, L.phi[c], L.omega);
}

__device__ __forceinline__ void wrap_coord(int x, int n, int per, int &xw, bool &ok)
{
    if (x >= 0 && x < n) { xw = x; ok = true; return; }
    if (per) { xw = ((x % n) + n) % n; ok = true; return; }
    xw = 0;
    ok = false;
}

// ------------------------------- fused red-black sweep, deep prefetch (3D)
// One launch performs a whole GSRB sweep (red then black).  A block owns a
// 64 x EY tile of x-y columns (the inner 62 x (EY-2) are written; the ring's
// red values are what the tile's black cells need) and marches through a
// z-chunk: at step k it updates the red cells of plane k and the black cells
// of plane k-1, whose red neighbours are then all final.  The sweep reads pin
// and writes pout, so overlapping tiles never observe each other's writes;
// the result equals the two-pass red/black update exactly.  Global loads are
// issued PD planes ahead (face weights, rhs, ring-2 values) and the old
// iterate PD+1 planes ahead, held in register queues.  One column per thread,
// 32-bit indexing (levels below 2^31 cells).  MODE 0: plain; 1: zero old
// iterate.  AL: the level has alpha (its value travels with the face weights)
template <typename T, int MODE, int EY, int PD, bool AL = false, bool SOR = false>
__global__ void __launch_bounds__(64 * EY) k_gsrb_deep(Lev<T> L, const T *__restrict__ pin,
                                                      T *__restrict__ pout, FusedGrid G)
{
    constexpr int NW = AL ? 7 : 6;   // face weights (+ alpha) per cell
    constexpr int EX = 64, FX = EX - 2, FY = EY - 2, PX = EX + 2, PY = EY + 2, NT = EX * EY;
    constexpr int QF = PD;   // old-iterate queue: planes k+2 .. k+1+QF
    const int b = blockIdx.x;
    const int idx = (b & 7) * G.per_xcd + (b >> 3);    // XCD-contiguous tiles
    if (idx >= G.total) return;
    const int ntile = G.ntx * G.nty;
    const int chunk = idx / ntile, tile = idx - chunk * ntile;
    const int tyi = tile / G.ntx, txi = tile - tyi * G.ntx;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int plane = nx * ny;
    const int z0 = chunk * G.tz, z1 = min(z0 + G.tz, nz);
    const int tx = threadIdx.x & (EX - 1), ty = threadIdx.x >> 6;
    const int X = txi * FX - 1 + tx, Y = tyi * FY - 1 + ty;
    int Xw, Yw;
    bool okx, oky;
    wrap_coord(X, nx, L.perx, Xw, okx);
    wrap_coord(Y, ny, L.pery, Yw, oky);
    const bool active = okx && oky;
    const bool outcol = tx >= 1 && tx <= FX && ty >= 1 && ty <= FY && X < nx && Y < ny;
    const int col = Xw + nx * Yw;
    const int pos = (tx + 1) + PX * (ty + 1);
    int r2a = -1, r2b = -1, c2a = 0, c2b = 0;
    if (ty == 0 || ty == EY - 1) {
        int yw;
        bool ok;
        wrap_coord(ty == 0 ? Y - 1 : Y + 1, ny, L.pery, yw, ok);
        if (ok && okx) { r2a = (tx + 1) + PX * (ty == 0 ? 0 : EY + 1); c2a = Xw + nx * yw; }
    }
    if (tx == 0 || tx == EX - 1) {
        int xw;
        bool ok;
        wrap_coord(tx == 0 ? X - 1 : X + 1, nx, L.perx, xw, ok);
        if (ok && oky) { r2b = (tx == 0 ? 0 : EX + 1) + PX * (ty + 1); c2b = xw + nx * Yw; }
    }
    __shared__ T tile_s[3][PX * PY];
    for (int q = threadIdx.x; q < 3 * PX * PY; q += NT) (&tile_s[0][0])[q] = T(0);

    auto kwrap = [&](int k, int &kw) -> bool {
        if (k >= 0 && k < nz) { kw = k; return true; }
        if (L.perz) {
            while (k < 0) k += nz;
            while (k >= nz) k -= nz;
            kw = k;
            return true;
        }
        kw = 0;
        return false;
    };
    auto ld = [&](int c, int k) -> T {   // old iterate, zero outside / in MODE 1
        int kw;
        if (MODE == 1 || !kwrap(k, kw)) return T(0);
        return pin[c + plane * kw];
    };
    const int pxy = Xw + Yw;
    const int kb = z0 - 1, ke = z1;
    // old iterate of this column: pm2, pm1 (updated), p0 (plane k), pp1, then
    // the queue fq[i] = plane k+2+i
    T pm2 = T(0), pm1 = T(0), p0 = T(0), pp1 = T(0), fq[QF];
    if (active) {
        pm1 = ld(col, kb - 1);
        p0 = ld(col, kb);
        pp1 = ld(col, kb + 1);
    }
#pragma unroll
    for (int i = 0; i < QF; ++i) fq[i] = active ? ld(col, kb + 2 + i) : T(0);
    // per-plane inputs queue: sw/srhs/sa/sb[i] = plane k+1+i at step k
    T cw[NW], crhs = T(0), pw[NW], prhs = T(0);
    T sw[PD][NW], srhs[PD], sa[PD], sb[PD];
#pragma unroll
    for (int s = 0; s < NW; ++s) { cw[s] = T(0); pw[s] = T(0); }
#pragma unroll
    for (int i = 0; i < PD; ++i) {
        srhs[i] = sa[i] = sb[i] = T(0);
#pragma unroll
        for (int s = 0; s < NW; ++s) sw[i][s] = T(0);
    }
    auto alpha_of = [](const T *w) -> T {
        if constexpr (AL) return w[6];
        else return T(-0.0);
    };
    auto fetch = [&](int k, T *nw, T &nrhs, T &na, T &nb) {
        int kw;
        const bool kok = kwrap(k, kw);
        if (active && kok) {
            const int c = col + plane * kw;
            nw[0] = L.bx[c];
            nw[1] = (Xw < nx - 1) ? L.bx[c + 1] : L.bxh[Yw + ny * kw];
            nw[2] = L.by[c];
            nw[3] = (Yw < ny - 1) ? L.by[c + nx] : L.byh[Xw + nx * kw];
            nw[4] = L.bz[c];
            nw[5] = (kw < nz - 1) ? L.bz[c + plane] : L.bzh[col];
            if constexpr (AL) nw[6] = L.alpha[c];
            nrhs = L.rhs[c];
        }
        na = (r2a >= 0) ? ld(c2a, k) : T(0);
        nb = (r2b >= 0) ? ld(c2b, k) : T(0);
    };
    // prologue: plane kb goes to (cw, crhs) via the queue head
#pragma unroll
    for (int i = 0; i < PD; ++i)
        if (kb + i <= ke) fetch(kb + i, sw[i], srhs[i], sa[i], sb[i]);
    __syncthreads();
    T *buf = tile_s[0], *pbuf = tile_s[2], *qbuf = tile_s[1];
    for (int k = kb; k <= ke; ++k) {
        if (k > kb) {
            T *t = qbuf;
            qbuf = pbuf;
            pbuf = buf;
            buf = t;
        }
        int kw;
        const bool kok = kwrap(k, kw);
        // advance the input queue: head -> current plane, fetch plane k+PD
#pragma unroll
        for (int s = 0; s < NW; ++s) { pw[s] = cw[s]; cw[s] = sw[0][s]; }
        prhs = crhs;
        crhs = srhs[0];
        const T r2va = sa[0], r2vb = sb[0];
#pragma unroll
        for (int i = 0; i + 1 < PD; ++i) {
#pragma unroll
            for (int s = 0; s < NW; ++s) sw[i][s] = sw[i + 1][s];
            srhs[i] = srhs[i + 1];
            sa[i] = sa[i + 1];
            sb[i] = sb[i + 1];
        }
        if (k + PD <= ke) fetch(k + PD, sw[PD - 1], srhs[PD - 1], sa[PD - 1], sb[PD - 1]);
        // (a) publish the old values of plane k
        if (active) buf[pos] = kok ? p0 : T(0);
        if (r2a >= 0) buf[r2a] = r2va;
        if (r2b >= 0) buf[r2b] = r2vb;
        __syncthreads();
        const bool red = kok && ((pxy + kw) & 1) == 0;
        // (b) red cells of plane k (tile and ring)
        if (active && red) {
            T p[6] = {buf[pos - 1], buf[pos + 1], buf[pos - PX], buf[pos + PX], pm1, pp1};
            T v = relax<SOR>(gs_value<3, T>(L, cw, p, crhs, alpha_of(cw)), p0, L.omega);
            buf[pos] = v;
            p0 = v;
        }
        // (c) black cells of plane k-1 (tile only), (d) write plane k-1
        const int km = k - 1;
        if (outcol && km >= z0 && km < z1) {
            if ((pxy + km) & 1) {
                T p[6] = {pbuf[pos - 1], pbuf[pos + 1], pbuf[pos - PX], pbuf[pos + PX], pm2, p0};
                pm1 = relax<SOR>(gs_value<3, T>(L, pw, p, prhs, alpha_of(pw)), pm1, L.omega);
            }
            pout[col + plane * km] = pm1;
        }
        // rotate the column and fetch the old iterate of plane k+2+QF
        pm2 = pm1;
        pm1 = p0;
        p0 = pp1;
        pp1 = fq[0];
#pragma unroll
        for (int i = 0; i + 1 < QF; ++i) fq[i] = fq[i + 1];
        fq[QF - 1] = (active && k + 2 + QF <= ke + 1) ? ld(col, k + 2 + QF) : T(0);
    }
}

// ------------------------------------------ fused red-black sweep, full rows
// Each lane owns two neighbouring columns x0, x0+1 (x0 = 2*(lane + 64*w)), and
// W waves cover a whole x-row (nx = 128*W), so there is no x-ring and every
// lane has one red and one black cell per plane (no idle lanes in either
// phase).  The LDS rows carry a pad column on each side: zero outside a
// non-periodic side, the periodic image otherwise (mirrored by the end lanes).
// EY rows per block, the first and last being the y-ring (red only, not
// written); z-march as k_gsrb_deep with PD planes of prefetch.  Loads and
// stores are two-element vectors.  MODE 0: plain; 1: zero old iterate.
template <typename T> struct Vec2;
template <> struct Vec2<float> { using type = float2; };
template <> struct Vec2<double> { using type = double2; };

struct RowSweepGrid {
    int W;         // waves per x-row
    int nty;       // y-tiles
    int tz;        // planes per z-chunk
    int total;     // y-tiles x z-chunks
    int per_xcd;   // ceil(total / 8)
};

// Load/store at a wave-uniform base plus a per-lane 32-bit byte offset,
// through a raw buffer resource built from the base (the resource sits in
// SGPRs and the offset is one VGPR; a plain pointer form lets the compiler
// hoist a 64-bit per-lane address per array out of the march, which costs
// two VGPRs per array).  The range is not limited (num_records = 2^32-1);
// every offset used stays inside its array.
static constexpr int BUF_CFG = 0x00020000;   // gfx9-family raw buffer dword 3

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *base)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)0xffffffff, BUF_CFG);
}
template <typename V>
__device__ __forceinline__ V ld_off(const void *base, unsigned boff)
{
    const __amdgpu_buffer_rsrc_t r = buf_rsrc(base);
    static_assert(sizeof(V) == 4 || sizeof(V) == 8 || sizeof(V) == 16, "buffer load width");
    if constexpr (sizeof(V) == 4)
        return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b32(r, boff, 0, 0));
    else if constexpr (sizeof(V) == 8)
        return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b64(r, boff, 0, 0));
    else
        return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b128(r, boff, 0, 0));
}
template <typename V>
__device__ __forceinline__ void st_off(void *base, unsigned boff, const V &v)
{
    const __amdgpu_buffer_rsrc_t r = buf_rsrc(base);
    static_assert(sizeof(V) == 4 || sizeof(V) == 8 || sizeof(V) == 16, "buffer store width");
    if constexpr (sizeof(V) == 4)
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, boff, 0, 0);
    else if constexpr (sizeof(V) == 8)
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, v),
                                              r, boff, 0, 0);
    else
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(__attribute__((ext_vector_type(4))) unsigned, v),
                                               r, boff, 0, 0);
}

__device__ __forceinline__ unsigned short ld_u16(const void *base, unsigned boff)
{
    return __builtin_amdgcn_raw_buffer_load_b16(buf_rsrc(base), boff, 0, 0);
}

template <typename V>
__device__ __forceinline__ V ld_rs(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    static_assert(sizeof(V) == 2 || sizeof(V) == 4 || sizeof(V) == 8 || sizeof(V) == 16, "buffer load width");
    if constexpr (sizeof(V) == 2)
        return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, 0));
    else if constexpr (sizeof(V) == 4)
        return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
    else if constexpr (sizeof(V) == 8)
        return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
    else
        return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
}
// Cache policy of a buffer access: 0 default; CPOL_NT the non-temporal
// (streaming) hint.  The sweeps write their output with it: the new iterate
// is written once and read by the next launch, and the streaming stores
// measured 2-3% faster per sweep at 256^3 (FP32 111 -> 108.6 us, FP64 173 ->
// 167.8 us).  Their loads keep the default: the y-ring rows a neighbouring
// tile re-reads come from L2 only if the first read allocated them there.
static constexpr int CPOL_NT = 2;

template <typename V, int CPOL = 0>
__device__ __forceinline__ void st_rs(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, const V &v)
{
    static_assert(sizeof(V) == 8 || sizeof(V) == 16, "buffer store width");
    if constexpr (sizeof(V) == 8)
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, v),
                                              r, voff, soff, CPOL);
    else
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(__attribute__((ext_vector_type(4))) unsigned, v),
                                               r, voff, soff, CPOL);
}

// Full-row sweep, nx = 128 W (W a template parameter, so every LDS offset is
// an immediate).  A wave covers 128 columns of one row, two per lane (one red
// and one black cell on every plane).  Each array's buffer resource is built
// once; a plane's byte offset goes in the scalar offset of the load (the
// arrays are below 4 GiB).  The LDS rows carry a zero pad column on each
// side; an x-periodic neighbour across the row end is read from the opposite
// end of the same LDS row (per-lane offsets fixed for the whole march), so no
// pad is ever written after the initial clear.  Plane values rotate through
// four LDS buffers (plane m in buffer m & 3, a compile-time slot of the
// 8-step unroll), one barrier per plane.
// B16: the face weights and right side come from compact (bfloat16) storage
// K instead of L; the arithmetic is the same FP32 code on the widened values.
// Byte offsets of the arrays a row sweep touches: the iterates from bp, the
// weights and rhs from bw (one allocation each, so one buffer resource each)
struct RowBufs {
    const char *bp, *bw;
    unsigned in, out;
    unsigned bx, by, bz, rhs, bxh, byh, bzh;
    unsigned alpha;   // AL only (alpha is part of the weights' allocation)
};

// AL: the level has alpha (a column pair per plane beside the weights).
// PDW: how many planes ahead the weights, rhs and ring-2 values are fetched
// (the z faces always two): 2, or 1, which keeps one slot of the input ring
// fewer in registers (the FP32 form of 6 rows then fits 80 VGPRs, two blocks
// per CU).
template <typename T, int MODE, int EY, int NTH, bool B16, int W, bool AL = false, int PDW = 2, bool SOR = false>
__global__ void __launch_bounds__(NTH) k_gsrb_rows(Lev<T> L, RowBufs B, RowSweepGrid G)
{
    static_assert(!B16 || sizeof(T) == 4, "compact storage widens to FP32");
    static_assert(!(AL && B16), "compact storage carries no alpha");
    using T2 = typename Vec2<T>::type;
    constexpr int NX = 128 * W, FY = EY - 2, PY = EY + 2, NI = AL ? 15 : 13;
    constexpr int PXL = NX + 4;      // LDS row: pad (x = -1) at 1, column x at 2 + x, pad (x = nx) at NX + 2
    constexpr int PS = PY * PXL;     // LDS plane
    constexpr unsigned ES = sizeof(T), WS = B16 ? 2u : (unsigned)sizeof(T);   // iterate / weight bytes
    const int b = blockIdx.x;
    const int idx = (b & 7) * G.per_xcd + (b >> 3);    // XCD-contiguous tiles
    if (idx >= G.total) return;
    const int chunk = idx / G.nty, tyi = idx - chunk * G.nty;
    const int ny = L.ny, nz = L.nz, plane = NX * ny;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int w = wv % W, ty = wv / W;                  // wave-uniform
    const int x0 = 2 * (lane + 64 * w);
    const int Y = tyi * FY - 1 + ty;
    int Yw;
    bool oky;
    wrap_coord(Y, ny, L.pery, Yw, oky);
    const int z0 = chunk * G.tz, z1 = min(z0 + G.tz, nz);
    const bool outrow = oky && ty >= 1 && ty <= FY && Y < ny;
    const bool ring = (ty == 0) || (ty == EY - 1);
    int y2w = 0;
    bool ok2 = false;
    if (ring) wrap_coord(ty == 0 ? Y - 1 : Y + 1, ny, L.pery, y2w, ok2);
    const bool ylast = Yw == ny - 1, xend = x0 + 2 == NX;
    // per-lane byte offsets within a plane
    const unsigned col = (unsigned)(x0 + NX * Yw);
    const unsigned vw = col * WS, vp = col * ES, vp2 = (unsigned)(x0 + NX * y2w) * ES;
    const unsigned vyn = (ylast ? (unsigned)x0 : col + NX) * WS;    // y-high faces: by or the byh slab
    const unsigned vx2 = (xend ? col : col + 2u) * WS;             // third x face (bxh replaces it on xend)
    const unsigned bynstride = ylast ? (unsigned)NX : (unsigned)plane;
    const unsigned obyn = ylast ? B.byh : B.by;
    const __amdgpu_buffer_rsrc_t Rp = buf_rsrc(B.bp), Rw = buf_rsrc(B.bw);
    // a column pair / one value of a weight or rhs array
    auto ldw = [&](__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) -> T2 {
        if constexpr (B16) {
            const unsigned u = ld_rs<unsigned>(r, voff, soff);
            T2 v;
            v.x = bf16_lo(u);
            v.y = bf16_hi(u);
            return v;
        } else {
            return ld_rs<T2>(r, voff, soff);
        }
    };
    auto ldw1 = [&](__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) -> T {
        if constexpr (B16) return bf16_lo((unsigned)ld_rs<unsigned short>(r, voff, soff));
        else return ld_rs<T>(r, voff, soff);
    };
    // LDS element indices: own column pair, ring-2 row, x-neighbours x0-1 and x0+2
    const int lp = (ty + 1) * PXL + 2 + x0;
    const int lp2 = (ty == 0 ? 0 : EY + 1) * PXL + 2 + x0;
    const int lxl = (x0 == 0 && L.perx) ? lp + NX - 1 : lp - 1;
    const int lxr = (xend && L.perx) ? lp - (NX - 2) : lp + 2;
    extern __shared__ __align__(16) unsigned char rows_smem[];
    T *lds = (T *)rows_smem;
    for (int q = threadIdx.x; q < 4 * PS; q += blockDim.x) lds[q] = T(0);

    // plane k (at most one period outside the domain) -> stored plane kw;
    // false outside a non-periodic side (kw = 0 stands in)
    auto zwrap = [&](int k, int &kw) -> bool {
        const bool in = (unsigned)k < (unsigned)nz;
        kw = in ? k : (L.perz ? (k < 0 ? k + nz : k - nz) : 0);
        return in || L.perz;
    };
    // zero pairs are built from immediates (a zero object in memory would
    // turn "cond ? load : zero" into a flat load through a selected pointer)
    auto z2 = []() {
        T2 v;
        v.x = T(0);
        v.y = T(0);
        return v;
    };
    // old iterate of the column pair on plane k (zero outside / in MODE 1)
    auto ldphi = [&](int k) -> T2 {
        if constexpr (MODE == 1) {
            return z2();
        } else {
            int kw;
            const bool ok = zwrap(k, kw) && oky;
            const T2 v = ld_rs<T2>(Rp, vp, B.in + (unsigned)kw * (unsigned)plane * ES);
            return ok ? v : z2();
        }
    };
    // per-plane inputs: [0..2] x faces at x0, x0+1, x0+2; [3,4] y-low faces;
    // [5,6] y-high faces; [7,8] z-low faces; [9,10] rhs; [11,12] ring-2 old
    // values (ring waves).  The z-high faces of plane k are the z-low faces of
    // plane k+1 (next slot of the ring), except on the domain's top plane,
    // whose high faces come from the bzh slab.  Every entry is assigned and every
    // load is made from an in-range address (plane 0 stands in for a plane
    // outside a non-periodic side, row 0 for a row outside; those values are
    // never used), so the queue stays in registers.
    struct In {
        T v[NI];
    };
    auto fetch = [&](int k) -> In {
        In o;
        int kw;
        const bool kok = zwrap(k, kw);
        const unsigned sw = (unsigned)kw * (unsigned)plane * WS;
        const T2 vx = ldw(Rw, vw, B.bx + sw);
        o.v[0] = vx.x;
        o.v[1] = vx.y;
        const T bx2 = ldw1(Rw, vx2, B.bx + sw);
        const T bxh = ldw1(Rw, 0u, B.bxh + (unsigned)(Yw + ny * kw) * WS);
        o.v[2] = xend ? bxh : bx2;
        const T2 vy = ldw(Rw, vw, B.by + sw);
        o.v[3] = vy.x;
        o.v[4] = vy.y;
        const T2 vn = ldw(Rw, vyn, obyn + (unsigned)kw * bynstride * WS);
        o.v[5] = vn.x;
        o.v[6] = vn.y;
        const T2 vz = ldw(Rw, vw, B.bz + sw);
        o.v[7] = vz.x;
        o.v[8] = vz.y;
        const T2 vr = ldw(Rw, vw, B.rhs + sw);
        o.v[9] = vr.x;
        o.v[10] = vr.y;
        T2 va = z2();
        if (MODE == 0 && ring) va = ld_rs<T2>(Rp, vp2, B.in + (unsigned)kw * (unsigned)plane * ES);
        const bool use2 = kok && ok2;
        o.v[11] = use2 ? va.x : T(0);
        o.v[12] = use2 ? va.y : T(0);
        if constexpr (AL) {   // [13, 14]: alpha of the column pair
            const T2 al = ldw(Rw, vw, B.alpha + sw);
            o.v[13] = al.x;
            o.v[14] = al.y;
        }
        return o;
    };
    // PDW = 1: everything but the z faces one plane ahead, into the slot of
    // plane k+1 (whose z faces were fetched a step earlier); z faces two ahead
    auto fetch_w = [&](int k, In &o) {
        const In f = fetch(k);
#pragma unroll
        for (int t = 0; t < NI; ++t)
            if (t != 7 && t != 8) o.v[t] = f.v[t];
    };
    auto fetch_z = [&](int k, In &o) {
        int kw;
        zwrap(k, kw);
        const T2 vz = ldw(Rw, vw, B.bz + (unsigned)kw * (unsigned)plane * WS);
        o.v[7] = vz.x;
        o.v[8] = vz.y;
    };
    const int kb = z0 - 1, ke = z1;
    const int nsteps = ((ke - kb + 1) + 7) & ~7;   // a multiple of the unroll
    // rings indexed by the step m = k - kb: inputs of plane m in wq[m & 3],
    // old/updated iterate of plane m in ph[m & 7], LDS plane m in buffer m & 3
    In wq[4];
    struct Ring {
        T v[16];
    } phr;
#define PH(slot, c) phr.v[2 * (slot) + (c)]
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int t = 0; t < NI; ++t) wq[s].v[t] = T(0);
    // planes kb-1 .. kb+5 into slots 7, 0 .. 5 (written out: a loop here may
    // stay rolled and turn the ring into a stack array)
    auto put = [&](auto SC, const T2 &e) __attribute__((always_inline)) {
        constexpr int S = decltype(SC)::value;
        PH(S, 0) = e.x;
        PH(S, 1) = e.y;
    };
    put(std::integral_constant<int, 7>{}, ldphi(kb - 1));
    put(std::integral_constant<int, 0>{}, ldphi(kb));
    put(std::integral_constant<int, 1>{}, ldphi(min(kb + 1, ke + 1)));
    put(std::integral_constant<int, 2>{}, ldphi(min(kb + 2, ke + 1)));
    put(std::integral_constant<int, 3>{}, ldphi(min(kb + 3, ke + 1)));
    put(std::integral_constant<int, 4>{}, ldphi(min(kb + 4, ke + 1)));
    put(std::integral_constant<int, 5>{}, ldphi(min(kb + 5, ke + 1)));
    PH(6, 0) = PH(6, 1) = T(0);
    wq[0] = fetch(kb);
    if constexpr (PDW == 2) wq[1] = fetch(min(kb + 1, ke + 1));
    else fetch_z(min(kb + 1, ke + 1), wq[1]);
    __syncthreads();

    auto step = [&](auto SC, int m) __attribute__((always_inline)) {
        constexpr int S = decltype(SC)::value;
        constexpr int C0 = S & 3, CP = (S + 3) & 3, C1 = (S + 1) & 3, CN = (S + 2) & 3;   // m, m-1, m+1, m+2
        constexpr int F0 = S, FM1 = (S + 7) & 7, FM2 = (S + 6) & 7, FP1 = (S + 1) & 7;
        T *const buf = lds + (S & 3) * PS, *const pbuf = lds + ((S + 3) & 3) * PS;
        const int k = kb + m;
        int kw;
        const bool kok = zwrap(min(k, ke + 1), kw) && k <= ke;
        if constexpr (PDW == 2) {
            wq[CN] = fetch(min(k + 2, ke + 1));
        } else {
            fetch_w(min(k + 1, ke + 1), wq[C1]);
            fetch_z(min(k + 2, ke + 1), wq[CN]);
        }
        // (a) publish the old values of plane k (and the ring-2 row)
        if (oky) {
            T2 v;
            v.x = kok ? PH(F0, 0) : T(0);
            v.y = kok ? PH(F0, 1) : T(0);
            *(T2 *)(buf + lp) = v;
        }
        if (ring) {
            T2 v;
            v.x = kok ? wq[C0].v[11] : T(0);
            v.y = kok ? wq[C0].v[12] : T(0);
            *(T2 *)(buf + lp2) = v;
        }
        __syncthreads();
        // r is uniform over the wave (one row, one plane): each colour case
        // is its own straight-line code behind a scalar branch.  The parity
        // uses k itself (a plane outside a non-periodic side has kw = 0; a
        // periodic extent is even, so kw and k agree in parity).
        const int r = (Yw + k) & 1;
        const int km = k - 1;
        const bool do_red = kok && oky, do_black = outrow && km >= z0 && km < z1;
        // z-high faces of plane k (zh) and plane k-1 (zhb): the next plane's
        // z-low faces, or the bzh slab on the domain's top plane
        T zh0, zh1, zb0, zb1;
        if (kw == nz - 1) {
            const T2 t = ldw(Rw, vw, B.bzh);
            zh0 = t.x;
            zh1 = t.y;
        } else {
            zh0 = wq[C1].v[7];
            zh1 = wq[C1].v[8];
        }
        if (km == nz - 1) {
            const T2 t = ldw(Rw, vw, B.bzh);
            zb0 = t.x;
            zb1 = t.y;
        } else {
            zb0 = wq[C0].v[7];
            zb1 = wq[C0].v[8];
        }
        // the ring values this step reads, as values: the colour branches
        // below then touch no ring slot, so the ring stays in registers (a
        // slot chosen per branch would be merged into a pointer select)
        const T f0a = PH(F0, 0), f0b = PH(F0, 1), m1a = PH(FM1, 0), m1b = PH(FM1, 1);
        const T p1a = PH(FP1, 0), p1b = PH(FP1, 1), m2a = PH(FM2, 0), m2b = PH(FM2, 1);
        T red_v = T(0), blk_v = T(0);
        auto update = [&](auto RC) __attribute__((always_inline)) {
            constexpr int R = decltype(RC)::value;
            const In &cw = wq[C0], &pw = wq[CP];
            auto alpha_of = [](const In &q) -> T {
                if constexpr (AL) return q.v[13 + R];
                else return T(-0.0);
            };
            T f0R = R ? f0b : f0a;   // column R of plane k, after the red update
            if (do_red) {
                // (b) the red cell of plane k: column R
                const T wt[6] = {cw.v[R], cw.v[R + 1], cw.v[3 + R], cw.v[5 + R], cw.v[7 + R], R ? zh1 : zh0};
                T xm, xp;
                if constexpr (R == 1) {
                    xm = f0a;
                    xp = buf[lxr];
                } else {
                    xm = buf[lxl];
                    xp = f0b;
                }
                const T nb[6] = {xm, xp, buf[lp - PXL + R], buf[lp + PXL + R], R ? m1b : m1a, R ? p1b : p1a};
                red_v = relax<SOR>(gs_value<3, T>(L, wt, nb, cw.v[9 + R], alpha_of(cw)), f0R, L.omega);
                buf[lp + R] = red_v;
                f0R = red_v;
            }
            if (do_black) {
                // (c) the black cell of plane k-1 (same column R); the other
                // column is red there; (d) write plane k-1
                const T wt[6] = {pw.v[R], pw.v[R + 1], pw.v[3 + R], pw.v[5 + R], pw.v[7 + R], R ? zb1 : zb0};
                T xm, xp;
                if constexpr (R == 1) {
                    xm = m1a;
                    xp = pbuf[lxr];
                } else {
                    xm = pbuf[lxl];
                    xp = m1b;
                }
                const T nb[6] = {xm, xp, pbuf[lp - PXL + R], pbuf[lp + PXL + R], R ? m2b : m2a, f0R};
                blk_v = relax<SOR>(gs_value<3, T>(L, wt, nb, pw.v[9 + R], alpha_of(pw)), R ? m1b : m1a, L.omega);
                T2 o;
                o.x = R ? m1a : blk_v;
                o.y = R ? blk_v : m1b;
                st_rs<T2, CPOL_NT>(Rp, vp, B.out + (unsigned)km * (unsigned)plane * ES, o);
            }
        };
        if (r) update(std::integral_constant<int, 1>{});
        else update(std::integral_constant<int, 0>{});
        // write the updated values back to their (fixed) ring slots
        if (do_red) {
            PH(F0, 0) = r ? f0a : red_v;
            PH(F0, 1) = r ? red_v : f0b;
        }
        if (do_black) {
            PH(FM1, 0) = r ? m1a : blk_v;
            PH(FM1, 1) = r ? blk_v : m1b;
        }
        // the slot of plane k-2 takes the old iterate of plane k+6
        {
            const T2 e = (k + 6 <= ke + 1) ? ldphi(k + 6) : z2();
            PH(FM2, 0) = e.x;
            PH(FM2, 1) = e.y;
        }
    };
    for (int m = 0; m < nsteps; m += 8) {
        step(std::integral_constant<int, 0>{}, m);
        step(std::integral_constant<int, 1>{}, m + 1);
        step(std::integral_constant<int, 2>{}, m + 2);
        step(std::integral_constant<int, 3>{}, m + 3);
        step(std::integral_constant<int, 4>{}, m + 4);
        step(std::integral_constant<int, 5>{}, m + 5);
        step(std::integral_constant<int, 6>{}, m + 6);
        step(std::integral_constant<int, 7>{}, m + 7);
    }
}
#undef PH

// -------------------- fused red-black sweep, full rows, weights from rho
// The FP32 full-row sweep of a plane-parallel projection's finest level with
// the face weights formed from the cell density instead of read: a face
// between cells of densities rho_L and rho_R has w = c / (rho_L + rho_R),
// c = 2 beta0^p of the face (x and y faces: the plane's cell-centred beta0;
// z faces: the edge value; p = 2 with use_alt_energy_fix).  A cell update
// then reads rho and rhs (8 B) where the stored form reads three weights and
// rhs (16 B).  x and y are periodic; the z faces on the domain's low and high
// sides come from FP32 slabs of the folded FP64 weights (walls, outlets or
// the periodic image).  The reciprocal is the hardware's (1 ulp): these
// weights drive the FP32 preconditioner inside FP64 refinement, whose FP64
// residual uses the FP64 weights.  Structure as k_gsrb_rows (one column pair
// per lane, W waves per row, EY rows, z-march, one barrier per plane); rho
// rotates through LDS beside phi.  The z faces of both columns are formed
// once per plane, into a four-slot register ring.  PZ: z periodic, a template
// parameter because the wrap logic per step costs 5% of the sweep at run time.
struct RhoBufs {      // byte offsets in the weights' allocation (with B.rhs)
    unsigned rho;     // FP32 cell density
    unsigned cxy;     // per plane: 2 beta0^p (x and y faces)
    unsigned cz;      // per z face (nz + 1): 2 beta0_edge^p
    unsigned bzl;     // FP32 folded weights of the domain's low z faces (nx ny)
    unsigned bzh;     // and of its high z faces
};

template <int MODE, int EY, int NTH, int W, bool SOR = false, bool PZ = false>
__global__ void __launch_bounds__(NTH) k_gsrb_rows_rho(Lev<float> L, RowBufs B, RhoBufs RB, RowSweepGrid G)
{
    using T = float;
    using T2 = float2;
    constexpr int NX = 128 * W, FY = EY - 2, PY = EY + 2, NI = 8;
    constexpr int PXL = NX + 4, PS = PY * PXL;
    constexpr unsigned ES = 4u;
    const int b = blockIdx.x;
    const int idx = (b & 7) * G.per_xcd + (b >> 3);    // XCD-contiguous tiles
    if (idx >= G.total) return;
    const int chunk = idx / G.nty, tyi = idx - chunk * G.nty;
    const int ny = L.ny, nz = L.nz, plane = NX * ny;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int w = wv % W, ty = wv / W;                  // wave-uniform
    const int x0 = 2 * (lane + 64 * w);
    const int Y = tyi * FY - 1 + ty;
    int Yw;
    bool oky;
    wrap_coord(Y, ny, 1, Yw, oky);                      // y periodic
    const int z0 = chunk * G.tz, z1 = min(z0 + G.tz, nz);
    const bool outrow = ty >= 1 && ty <= FY && Y < ny;
    const bool ring = (ty == 0) || (ty == EY - 1);
    int y2w = 0;
    bool ok2 = false;
    if (ring) wrap_coord(ty == 0 ? Y - 1 : Y + 1, ny, 1, y2w, ok2);
    const bool xend = x0 + 2 == NX;
    const unsigned col = (unsigned)(x0 + NX * Yw);
    const unsigned vw = col * ES, vp2 = (unsigned)(x0 + NX * y2w) * ES;
    const __amdgpu_buffer_rsrc_t Rp = buf_rsrc(B.bp), Rw = buf_rsrc(B.bw);
    const float *cxy = (const float *)(B.bw + RB.cxy), *cz = (const float *)(B.bw + RB.cz);
    // LDS element indices: own column pair, ring-2 row, x-neighbours x0-1 and
    // x0+2 (x periodic: across the row end from the other end of the row)
    const int lp = (ty + 1) * PXL + 2 + x0;
    const int lp2 = (ty == 0 ? 0 : EY + 1) * PXL + 2 + x0;
    const int lxl = x0 == 0 ? lp + NX - 1 : lp - 1;
    const int lxr = xend ? lp - (NX - 2) : lp + 2;
    extern __shared__ __align__(16) unsigned char rows_smem[];
    T *lds = (T *)rows_smem;          // phi: four planes
    T *ldr = lds + 4 * PS;            // rho: four planes
    for (int q = threadIdx.x; q < 8 * PS; q += blockDim.x) lds[q] = T(0);

    auto zwrap = [&](int k, int &kw) -> bool {
        const bool in = (unsigned)k < (unsigned)nz;
        kw = in ? k : (PZ ? (k < 0 ? k + nz : k - nz) : 0);
        return in || PZ;
    };
    auto z2 = []() {
        T2 v;
        v.x = T(0);
        v.y = T(0);
        return v;
    };
    auto ldphi = [&](int k) -> T2 {
        if constexpr (MODE == 1) {
            return z2();
        } else {
            int kw;
            const bool ok = zwrap(k, kw);
            const T2 v = ld_rs<T2>(Rp, vw, B.in + (unsigned)kw * (unsigned)plane * ES);
            return ok ? v : z2();
        }
    };
    // per-plane inputs: [0, 1] rho of the column pair, [2, 3] rhs, [4, 5]
    // ring-2 old phi and [6, 7] ring-2 rho (ring waves)
    struct In {
        T v[NI];
    };
    auto fetch = [&](int k) -> In {
        In o;
        int kw;
        const bool kok = zwrap(k, kw);
        const unsigned sw = (unsigned)kw * (unsigned)plane * ES;
        const T2 r = ld_rs<T2>(Rw, vw, RB.rho + sw);
        o.v[0] = r.x;
        o.v[1] = r.y;
        const T2 h = ld_rs<T2>(Rw, vw, B.rhs + sw);
        o.v[2] = h.x;
        o.v[3] = h.y;
        T2 va = z2(), vr = z2();
        if (ring) {
            if (MODE == 0) va = ld_rs<T2>(Rp, vp2, B.in + sw);
            vr = ld_rs<T2>(Rw, vp2, RB.rho + sw);
        }
        const bool use2 = kok && ok2;
        o.v[4] = use2 ? va.x : T(0);
        o.v[5] = use2 ? va.y : T(0);
        o.v[6] = use2 ? vr.x : T(0);
        o.v[7] = use2 ? vr.y : T(0);
        return o;
    };
    // the z faces f of the column pair (face f between planes f-1 and f),
    // from the pair's rho on planes f-1 and f; the domain's sides from the
    // slabs; faces beyond a non-periodic side (never used) are zero
    // the per-plane and per-face factors are loaded a step before their use
    // (scalar loads issued at the use stalled the march: 76 -> 68 us per 256^3
    // sweep); cz of face f, clamped into the table
    auto czload = [&](int f) -> T {
        int fw = f;
        if (PZ) fw = f < 0 ? f + nz : (f > nz ? f - nz : f);
        return cz[min(max(fw, 0), nz)];
    };
    auto facez = [&](int f, const In &lo, const In &hi, T c) -> T2 {
        int fw = f;
        if (PZ) fw = f < 0 ? f + nz : (f > nz ? f - nz : f);
        if (fw == 0 || fw == nz) return ld_rs<T2>(Rw, vw, fw == 0 ? RB.bzl : RB.bzh);   // slab: one plane
        if (fw < 0 || fw > nz) return z2();
        T2 v;
        v.x = c * __builtin_amdgcn_rcpf(lo.v[0] + hi.v[0]);
        v.y = c * __builtin_amdgcn_rcpf(lo.v[1] + hi.v[1]);
        return v;
    };
    // the Gauss-Seidel value with the hardware reciprocal of the diagonal (1
    // ulp): the weights themselves come from v_rcp_f32 already, and the
    // division's scaling and fix-up steps cost 2.5% of the sweep
    auto gs_rcp = [&](const T w[6], const T p[6], T rhs) __attribute__((always_inline)) -> T {
        const T tx = dfma(w[0], p[0], w[1] * p[1]), sx = w[0] + w[1];
        const T ty = dfma(w[2], p[2], w[3] * p[3]), sy = w[2] + w[3];
        const T tz = dfma(w[4], p[4], w[5] * p[5]), sz = w[4] + w[5];
        const T num = dfma(L.ih2x, tx, dfma(L.ih2y, ty, dfma(L.ih2z, tz, rhs)));
        const T den = dfma(L.ih2x, sx, dfma(L.ih2y, sy, L.ih2z * sz));
        return num * __builtin_amdgcn_rcpf(den);
    };
    const int kb = z0 - 1, ke = z1;
    const int nsteps = ((ke - kb + 1) + 7) & ~7;
    In wq[4];
    struct Ring {
        T v[16];
    } phr;
#define PH(slot, c) phr.v[2 * (slot) + (c)]
    T2 wz[4];   // z faces of the pair: face kb + m in wz[m & 3]
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        wz[q] = z2();
#pragma unroll
        for (int t = 0; t < NI; ++t) wq[q].v[t] = T(0);
    }
    auto put = [&](auto SC, const T2 &e) __attribute__((always_inline)) {
        constexpr int S = decltype(SC)::value;
        PH(S, 0) = e.x;
        PH(S, 1) = e.y;
    };
    put(std::integral_constant<int, 7>{}, ldphi(kb - 1));
    put(std::integral_constant<int, 0>{}, ldphi(kb));
    put(std::integral_constant<int, 1>{}, ldphi(min(kb + 1, ke + 1)));
    put(std::integral_constant<int, 2>{}, ldphi(min(kb + 2, ke + 1)));
    put(std::integral_constant<int, 3>{}, ldphi(min(kb + 3, ke + 1)));
    put(std::integral_constant<int, 4>{}, ldphi(min(kb + 4, ke + 1)));
    put(std::integral_constant<int, 5>{}, ldphi(min(kb + 5, ke + 1)));
    PH(6, 0) = PH(6, 1) = T(0);
    wq[3] = fetch(kb - 1);   // plane kb-1: its rho forms the z face kb
    wq[0] = fetch(kb);
    wq[1] = fetch(min(kb + 1, ke + 1));
    wz[0] = facez(kb, wq[3], wq[0], czload(kb));
    // factors of the first step: face kb+1, planes kb and kb-1
    T cz_c = czload(kb + 1), cxy_c, cxy_p;
    {
        int kw0;
        zwrap(kb, kw0);
        cxy_c = cxy[kw0];
        zwrap(kb - 1, kw0);
        cxy_p = cxy[kw0];
    }
    __syncthreads();

    auto step = [&](auto SC, int m) __attribute__((always_inline)) {
        constexpr int S = decltype(SC)::value;
        constexpr int C0 = S & 3, CP = (S + 3) & 3, C1 = (S + 1) & 3, CN = (S + 2) & 3;   // m, m-1, m+1, m+2
        constexpr int F0 = S, FM1 = (S + 7) & 7, FM2 = (S + 6) & 7, FP1 = (S + 1) & 7;
        T *const buf = lds + (S & 3) * PS, *const pbuf = lds + ((S + 3) & 3) * PS;
        T *const rbuf = ldr + (S & 3) * PS, *const prbuf = ldr + ((S + 3) & 3) * PS;
        const int k = kb + m;
        int kw;
        const bool kok = zwrap(min(k, ke + 1), kw) && k <= ke;

        // the z faces of plane k's top (face k+1), from planes k and k+1
        const T cz_use = cz_c;
        cz_c = czload(k + 2);
        wz[C1] = facez(k + 1, wq[C0], wq[C1], cz_use);
        wq[CN] = fetch(min(k + 2, ke + 1));
        // (a) publish the old phi and rho of plane k (and the ring-2 rows)
        {
            T2 v, r;
            v.x = kok ? PH(F0, 0) : T(0);
            v.y = kok ? PH(F0, 1) : T(0);
            r.x = kok ? wq[C0].v[0] : T(0);
            r.y = kok ? wq[C0].v[1] : T(0);
            *(T2 *)(buf + lp) = v;
            *(T2 *)(rbuf + lp) = r;
        }
        if (ring) {
            T2 v, r;
            v.x = kok ? wq[C0].v[4] : T(0);
            v.y = kok ? wq[C0].v[5] : T(0);
            r.x = kok ? wq[C0].v[6] : T(0);
            r.y = kok ? wq[C0].v[7] : T(0);
            *(T2 *)(buf + lp2) = v;
            *(T2 *)(rbuf + lp2) = r;
        }
        __syncthreads();
        const int r = (Yw + k) & 1;
        const int km = k - 1;
        const bool do_red = kok, do_black = outrow && km >= z0 && km < z1;
        const T cxy0 = cxy_c, cxym = cxy_p;
        cxy_p = cxy_c;
        {
            int kwn;
            zwrap(min(k + 1, ke + 1), kwn);
            cxy_c = cxy[kwn];
        }
        const T f0a = PH(F0, 0), f0b = PH(F0, 1), m1a = PH(FM1, 0), m1b = PH(FM1, 1);
        const T p1a = PH(FP1, 0), p1b = PH(FP1, 1), m2a = PH(FM2, 0), m2b = PH(FM2, 1);
        T red_v = T(0), blk_v = T(0);
        // x and y faces of column R of a plane: own pair (ra, rb), the LDS plane rb_
        auto faces_xy = [&](auto RC, T c, T ra, T rb, const T *rl, T wt[4]) __attribute__((always_inline)) {
            constexpr int R = decltype(RC)::value;
            const T rR = R ? rb : ra;
            if constexpr (R == 0) {
                wt[0] = c * __builtin_amdgcn_rcpf(rl[lxl] + ra);
                wt[1] = c * __builtin_amdgcn_rcpf(ra + rb);
            } else {
                wt[0] = c * __builtin_amdgcn_rcpf(ra + rb);
                wt[1] = c * __builtin_amdgcn_rcpf(rb + rl[lxr]);
            }
            wt[2] = c * __builtin_amdgcn_rcpf(rl[lp - PXL + R] + rR);
            wt[3] = c * __builtin_amdgcn_rcpf(rR + rl[lp + PXL + R]);
        };
        auto update = [&](auto RC) __attribute__((always_inline)) {
            constexpr int R = decltype(RC)::value;
            const In &cw = wq[C0], &pw = wq[CP];
            T f0R = R ? f0b : f0a;
            if (do_red) {
                T wxy[4];
                faces_xy(RC, cxy0, cw.v[0], cw.v[1], rbuf, wxy);
                const T wt[6] = {wxy[0], wxy[1], wxy[2], wxy[3], R ? wz[C0].y : wz[C0].x, R ? wz[C1].y : wz[C1].x};
                T xm, xp;
                if constexpr (R == 1) {
                    xm = f0a;
                    xp = buf[lxr];
                } else {
                    xm = buf[lxl];
                    xp = f0b;
                }
                const T nb[6] = {xm, xp, buf[lp - PXL + R], buf[lp + PXL + R], R ? m1b : m1a, R ? p1b : p1a};
                red_v = relax<SOR>(gs_rcp(wt, nb, cw.v[2 + R]), f0R, L.omega);
                buf[lp + R] = red_v;
                f0R = red_v;
            }
            if (do_black) {
                T wxy[4];
                faces_xy(RC, cxym, pw.v[0], pw.v[1], prbuf, wxy);
                const T wt[6] = {wxy[0], wxy[1], wxy[2], wxy[3], R ? wz[CP].y : wz[CP].x, R ? wz[C0].y : wz[C0].x};
                T xm, xp;
                if constexpr (R == 1) {
                    xm = m1a;
                    xp = pbuf[lxr];
                } else {
                    xm = pbuf[lxl];
                    xp = m1b;
                }
                const T nb[6] = {xm, xp, pbuf[lp - PXL + R], pbuf[lp + PXL + R], R ? m2b : m2a, f0R};
                blk_v = relax<SOR>(gs_rcp(wt, nb, pw.v[2 + R]), R ? m1b : m1a, L.omega);
                T2 o;
                o.x = R ? m1a : blk_v;
                o.y = R ? blk_v : m1b;
                st_rs<T2, CPOL_NT>(Rp, vw, B.out + (unsigned)km * (unsigned)plane * ES, o);
            }
        };
        if (r) update(std::integral_constant<int, 1>{});
        else update(std::integral_constant<int, 0>{});
        if (do_red) {
            PH(F0, 0) = r ? f0a : red_v;
            PH(F0, 1) = r ? red_v : f0b;
        }
        if (do_black) {
            PH(FM1, 0) = r ? m1a : blk_v;
            PH(FM1, 1) = r ? blk_v : m1b;
        }
        {
            const T2 e = (k + 6 <= ke + 1) ? ldphi(k + 6) : z2();
            PH(FM2, 0) = e.x;
            PH(FM2, 1) = e.y;
        }
    };
    for (int m = 0; m < nsteps; m += 8) {
        step(std::integral_constant<int, 0>{}, m);
        step(std::integral_constant<int, 1>{}, m + 1);
        step(std::integral_constant<int, 2>{}, m + 2);
        step(std::integral_constant<int, 3>{}, m + 3);
        step(std::integral_constant<int, 4>{}, m + 4);
        step(std::integral_constant<int, 5>{}, m + 5);
        step(std::integral_constant<int, 6>{}, m + 6);
        step(std::integral_constant<int, 7>{}, m + 7);
    }
}
#undef PH

// residual of the fine level restricted onto the coarse level in one pass
// (the fine residual is never stored); also clears the coarse iterate.
// Block: 64 coarse cells in x times the 2 (y) x 2 (z) fine offsets.
template <int DIM, typename T>
__global__ void __launch_bounds__(256) k_resid_restrict(Lev<T> F, const T *__restrict__ phi, Lev<T> C)
{
    const int tx = threadIdx.x & 63, b = (threadIdx.x >> 6) & 1, c = threadIdx.x >> 7;
    const int I = blockIdx.x * 64 + tx, J = blockIdx.y, K = blockIdx.z;
    __shared__ T s[4][64];
    T pair = T(0);
    if (I < C.nx && (DIM == 3 || c == 0)) {
        const int j = 2 * J + b, k = (DIM == 3) ? 2 * K + c : 0;
        T r[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const int i = 2 * I + a;
            const long f = i + (long)F.nx * (j + (long)F.ny * k);
            T w[6], p[6];
            stencil<DIM, T>(F, phi, i, j, k, f, w, p);
            r[a] = F.rhs[f] - with_alpha(F, f, apply_value<DIM, T>(F, w, p, phi[f]), phi[f]);
        }
        pair = r[0] + r[1];
    }
    s[b + 2 * c][tx] = pair;
    __syncthreads();
    if (threadIdx.x < 64 && I < C.nx) {
        const long q = I + (long)C.nx * (J + (long)C.ny * K);
        if (DIM == 3)
            C.rhs[q] = T(0.125) * ((s[0][tx] + s[1][tx]) + (s[2][tx] + s[3][tx]));
        else
            C.rhs[q] = T(0.25) * (s[0][tx] + s[1][tx]);
        C.phi[q] = T(0);
    }
}

// ------------------------------------------- residual (+ max norm)
// block max -> one atomic per block (max of non-negative doubles compares
// like their bit patterns, so the result is order independent)
__device__ __forceinline__ void block_max_atomic(double a, unsigned long long *nrm)
{
    a = wave_max(a);
    __shared__ double sm[TPB / 64];
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = sm[0];
        for (int w = 1; w < TPB / 64; ++w) m = fmax(m, sm[w]);
        atomic_max_nonneg(nrm, m);
    }
}

// Grid-stride walk over the x-rows (j,k) of an nx x ny x nz level: each
// wave takes whole rows and its lanes stride x.  The row index is
// wave-uniform, so its division into (j,k) runs once per wave and row on the
// scalar unit instead of a 64-bit divide per cell.  Consecutive blocks take
// consecutive row groups; with the round-robin block-to-XCD dispatch the z
// neighbour rows (a multiple of 8 blocks away for ny a multiple of 32) land
// on the same XCD.  (Giving each XCD one contiguous slab of rows measured
// 16% slower on the 256^3 residual.)
static constexpr int ROW_WAVES = TPB / 64;

template <typename F>
__device__ __forceinline__ void for_rows(int nx, int ny, int nz, F &&f)
{
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nrow = ny * nz;
    for (int row = blockIdx.x * ROW_WAVES + wv; row < nrow; row += gridDim.x * ROW_WAVES) {
        const int k = row / ny, j = row - k * ny;
        const long base = (long)row * nx;
        for (int i = lane; i < nx; i += 64) f(i, j, k, base + i);
    }
}

// grid for for_rows: one wave per row
static inline unsigned row_blocks(int ny, int nz)
{
    return (unsigned)(((long)ny * nz + ROW_WAVES - 1) / ROW_WAVES);
}

// residual array and/or its max norm; grid-stride so the norm costs at most
// one atomic per block of a fixed-size grid
template <int DIM, typename T>
__global__ void __launch_bounds__(TPB) k_residual(Lev<T> L, T *out, unsigned long long *nrm)
{
    double a = 0.0;
    for_rows(L.nx, L.ny, L.nz, [&](int i, int j, int k, long q) {
        T w[6], p[6];
        stencil<DIM, T>(L, L.phi, i, j, k, q, w, p);
        T r = L.rhs[q] - with_alpha(L, q, apply_value<DIM, T>(L, w, p, L.phi[q]), L.phi[q]);
        if (out) out[q] = r;
        a = fmax(a, fabs((double)r));
    });
    if (nrm) block_max_atomic(a, nrm);
}

// FP64 residual of the refinement loop, written as FP32 right side of the
// FP32 V-cycle, with its max norm
template <int DIM>
__global__ void __launch_bounds__(TPB) k_ir_residual(Lev<double> L, float *r32,
                                                    unsigned long long *nrm)
{
    double a = 0.0;
    for_rows(L.nx, L.ny, L.nz, [&](int i, int j, int k, long q) {
        double w[6], p[6];
        stencil<DIM, double>(L, L.phi, i, j, k, q, w, p);
        double r = L.rhs[q] - with_alpha(L, q, apply_value<DIM, double>(L, w, p, L.phi[q]), L.phi[q]);
        r32[q] = (float)r;
        a = fmax(a, fabs(r));
    });
    block_max_atomic(a, nrm);
}

// Refinement update fused with the next FP64 residual: phi_new = phi + e
// (e the FP32 correction) is formed on the fly for the cell and its six
// neighbours, written to phi_out, and r = rhs - A phi_new goes out as the
// FP32 right side of the next V-cycle, with its max norm.  phi_new equals
// what a separate phi += e pass stores, so the iteration is unchanged.
template <int DIM>
__global__ void __launch_bounds__(TPB) k_ir_update_residual(Lev<double> L, const float *__restrict__ e,
                                                           double *__restrict__ phi_out, float *r32,
                                                           unsigned long long *nrm)
{
    double a = 0.0;
    for_rows(L.nx, L.ny, L.nz, [&](int i, int j, int k, long q) {
        double w[6], p[6];
        float pe[6];
        stencil<DIM, double>(L, L.phi, i, j, k, q, w, p);
        nbr_values<DIM>(L, e, i, j, k, q, pe);
#pragma unroll
        for (int s = 0; s < 6; ++s) p[s] = p[s] + (double)pe[s];
        const double pc = L.phi[q] + (double)e[q];
        double r = L.rhs[q] - with_alpha(L, q, apply_value<DIM, double>(L, w, p, pc), pc);
        phi_out[q] = pc;
        r32[q] = (float)r;
        a = fmax(a, fabs(r));
    });
    block_max_atomic(a, nrm);
}

// ---------------------------------------- FP64 residual passes, row march
// The finest-level FP64 residual passes as one wave per (x-row segment,
// z-chunk): each lane owns two neighbouring columns and marches z, keeping
// the plane below, the plane and the plane above in registers, so every
// array is read from HBM once; the y- and x-neighbour values come from the
// neighbouring rows and lanes' loads (L1/L2).  Loads are two-element vectors
// at a uniform plane base plus a lane offset.  Needs 3D and nx = 128 W.
//   MODE 0: max |rhs - A phi|
//   MODE 1: r = rhs - A phi written as FP32 (the refinement's right side)
//   MODE 2: phi' = phi + e (e FP32) formed on the fly for the cell and its
//           neighbours, phi' written to phi_out, r = rhs - A phi' as FP32
// The arithmetic per cell is that of k_residual / k_ir_residual /
// k_ir_update_residual (same operand order), so results are identical.
struct RowMarchGrid {
    int W;         // waves per x-row
    int tz;        // planes per z-chunk
    int total;     // waves in all: ny * W * chunks
    int per_xcd;   // blocks per XCD
};

// B16: the FP32 right side goes out as bfloat16 (r16) instead of r32.
// AL: the level has alpha (+ alpha phi in each residual, as one fma).
template <int MODE, bool B16, bool AL = false>
__global__ void __launch_bounds__(256) k_resid_rows(Lev<double> L, const float *__restrict__ e,
                                                   double *__restrict__ phi_out, float *__restrict__ r32,
                                                   unsigned short *__restrict__ r16, unsigned long long *nrm,
                                                   RowMarchGrid G)
{
    using D2 = double2;
    using F2 = float2;
    constexpr bool HAS_E = MODE == 2;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int bidx = (blockIdx.x & 7) * G.per_xcd + (blockIdx.x >> 3);   // XCD-contiguous
    const int gw = bidx * 4 + wv;
    double a = 0.0;
    if (gw < G.total) {
        const int nx = L.nx, ny = L.ny, nz = L.nz, plane = nx * ny;
        const int seg = gw % G.W, t = gw / G.W;
        const int j = t % ny, chunk = t / ny;
        const int z0 = chunk * G.tz, z1 = min(z0 + G.tz, nz);
        const int x0 = 2 * (lane + 64 * seg);
        // neighbour rows / columns: index in range, plus whether it exists
        const bool okym = j > 0 || L.pery, okyp = j < ny - 1 || L.pery;
        const int jm = j > 0 ? j - 1 : ny - 1, jp = j < ny - 1 ? j + 1 : 0;
        const bool okxm = x0 > 0 || L.perx, okxp = x0 + 2 < nx || L.perx;
        const int xm = x0 > 0 ? x0 - 1 : nx - 1, xp = x0 + 2 < nx ? x0 + 2 : 0;
        const unsigned oc = (unsigned)(x0 + nx * j), oym = (unsigned)(x0 + nx * jm), oyp = (unsigned)(x0 + nx * jp);
        const unsigned oxm = (unsigned)(xm + nx * j), oxp = (unsigned)(xp + nx * j);
        const unsigned obyh = (unsigned)x0;   // byh row offset within its plane slab
        auto kwrap = [&](int k, int &kw) -> bool {
            if (k >= 0 && k < nz) { kw = k; return true; }
            if (L.perz) { kw = k < 0 ? k + nz : k - nz; return true; }
            kw = 0;
            return false;
        };
        struct Z {   // column-pair values of one plane (old iterate, correction, z-low faces)
            double p0, p1, b0, b1;
            float e0, e1;
        };
        auto loadz = [&](int k) -> Z {
            Z z;
            int kw;
            const bool ok = kwrap(k, kw);
            const size_t pk = (size_t)plane * kw;
            const D2 p = ld_off<D2>(L.phi + pk, oc * 8u);
            const D2 bz = ld_off<D2>(L.bz + pk, oc * 8u);
            z.p0 = ok ? p.x : 0.0;
            z.p1 = ok ? p.y : 0.0;
            z.b0 = bz.x;
            z.b1 = bz.y;
            if constexpr (HAS_E) {
                const F2 v = ld_off<F2>(e + pk, oc * 4u);
                z.e0 = ok ? v.x : 0.0f;
                z.e1 = ok ? v.y : 0.0f;
            } else {
                z.e0 = z.e1 = 0.0f;
            }
            return z;
        };
        struct P {   // in-plane inputs of one plane
            double ym0, ym1, yp0, yp1, xm, xp;
            float eym0, eym1, eyp0, eyp1, exm, exp;
            double bx0, bx1, bx2, by0, by1, byn0, byn1, r0, r1, a0, a1;
        };
        auto loadp = [&](int k) -> P {
            P q;
            const size_t pk = (size_t)plane * k;
            const double *ph = L.phi + pk;
            const D2 ym = ld_off<D2>(ph, oym * 8u), yp = ld_off<D2>(ph, oyp * 8u);
            const double vxm = ld_off<double>(ph, oxm * 8u), vxp = ld_off<double>(ph, oxp * 8u);
            q.ym0 = okym ? ym.x : 0.0;
            q.ym1 = okym ? ym.y : 0.0;
            q.yp0 = okyp ? yp.x : 0.0;
            q.yp1 = okyp ? yp.y : 0.0;
            q.xm = okxm ? vxm : 0.0;
            q.xp = okxp ? vxp : 0.0;
            if constexpr (HAS_E) {
                const float *ek = e + pk;
                const F2 eym = ld_off<F2>(ek, oym * 4u), eyp = ld_off<F2>(ek, oyp * 4u);
                const float exm = ld_off<float>(ek, oxm * 4u), exp = ld_off<float>(ek, oxp * 4u);
                q.eym0 = okym ? eym.x : 0.0f;
                q.eym1 = okym ? eym.y : 0.0f;
                q.eyp0 = okyp ? eyp.x : 0.0f;
                q.eyp1 = okyp ? eyp.y : 0.0f;
                q.exm = okxm ? exm : 0.0f;
                q.exp = okxp ? exp : 0.0f;
            } else {
                q.eym0 = q.eym1 = q.eyp0 = q.eyp1 = q.exm = q.exp = 0.0f;
            }
            const D2 bx = ld_off<D2>(L.bx + pk, oc * 8u);
            q.bx0 = bx.x;
            q.bx1 = bx.y;
            q.bx2 = (x0 + 2 < nx) ? ld_off<double>(L.bx + pk, (oc + 2u) * 8u) : L.bxh[j + ny * k];
            const D2 by = ld_off<D2>(L.by + pk, oc * 8u);
            q.by0 = by.x;
            q.by1 = by.y;
            const D2 byn = (j < ny - 1) ? ld_off<D2>(L.by + pk, (oc + nx) * 8u)
                                        : ld_off<D2>(L.byh + (size_t)nx * k, obyh * 8u);
            q.byn0 = byn.x;
            q.byn1 = byn.y;
            const D2 r = ld_off<D2>(L.rhs + pk, oc * 8u);
            q.r0 = r.x;
            q.r1 = r.y;
            if constexpr (AL) {
                const D2 al = ld_off<D2>(L.alpha + pk, oc * 8u);
                q.a0 = al.x;
                q.a1 = al.y;
            } else {
                q.a0 = q.a1 = 0.0;
            }
            return q;
        };
        Z zm = loadz(z0 - 1), zc = loadz(z0), zp = loadz(z0 + 1);
        P pc = loadp(z0);
        for (int k = z0; k < z1; ++k) {
            // next plane's inputs go out first; their latency hides behind this plane's work
            const Z zn = loadz(k + 2);
            const P pn = loadp(min(k + 1, z1 - 1));
            // z-high faces: the next plane's z-low faces, or the bzh slab on top
            double bzn0, bzn1;
            if (k < nz - 1) {
                bzn0 = zp.b0;
                bzn1 = zp.b1;
            } else {
                const D2 t2 = ld_off<D2>(L.bzh, oc * 8u);
                bzn0 = t2.x;
                bzn1 = t2.y;
            }
            double res[2], pcen[2];
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                double w[6], p[6];
                float pe[6];
                w[0] = c ? pc.bx1 : pc.bx0;
                w[1] = c ? pc.bx2 : pc.bx1;
                w[2] = c ? pc.by1 : pc.by0;
                w[3] = c ? pc.byn1 : pc.byn0;
                w[4] = c ? zc.b1 : zc.b0;
                w[5] = c ? bzn1 : bzn0;
                p[0] = c ? zc.p0 : pc.xm;
                p[1] = c ? pc.xp : zc.p1;
                p[2] = c ? pc.ym1 : pc.ym0;
                p[3] = c ? pc.yp1 : pc.yp0;
                p[4] = c ? zm.p1 : zm.p0;
                p[5] = c ? zp.p1 : zp.p0;
                double pcc = c ? zc.p1 : zc.p0;
                if constexpr (HAS_E) {
                    pe[0] = c ? zc.e0 : pc.exm;
                    pe[1] = c ? pc.exp : zc.e1;
                    pe[2] = c ? pc.eym1 : pc.eym0;
                    pe[3] = c ? pc.eyp1 : pc.eyp0;
                    pe[4] = c ? zm.e1 : zm.e0;
                    pe[5] = c ? zp.e1 : zp.e0;
#pragma unroll
                    for (int s = 0; s < 6; ++s) p[s] = p[s] + (double)pe[s];
                    pcc = pcc + (double)(c ? zc.e1 : zc.e0);
                }
                double ap = apply_value<3, double>(L, w, p, pcc);
                if constexpr (AL) ap = fma(c ? pc.a1 : pc.a0, pcc, ap);
                res[c] = (c ? pc.r1 : pc.r0) - ap;
                pcen[c] = pcc;
                a = fmax(a, fabs(res[c]));
            }
            const size_t pk = (size_t)plane * k;
            if constexpr (MODE >= 1 && B16) {
                st_off<unsigned>(r16 + pk, oc * 2u, bf16_pack((float)res[0], (float)res[1]));
            } else if constexpr (MODE >= 1) {
                F2 o;
                o.x = (float)res[0];
                o.y = (float)res[1];
                st_off<F2>(r32 + pk, oc * 4u, o);
            }
            if constexpr (MODE == 2) {
                D2 o;
                o.x = pcen[0];
                o.y = pcen[1];
                st_off<D2>(phi_out + pk, oc * 8u, o);
            }
            zm = zc;
            zc = zp;
            zp = zn;
            pc = pn;
        }
    }
    block_max_atomic(a, nrm);
}

// ------------------------------------- residual + restriction, row march
// The fine residual averaged onto the coarse level (and the coarse iterate
// cleared) as one wave per (coarse x-row segment, coarse z-chunk): a lane
// owns one coarse column, i.e. the fine column pair 2I, 2I+1 of the fine row
// pair 2J, 2J+1, and marches fine planes with the planes below and above in
// registers.  Per-cell arithmetic and the summation order of the eight fine
// residuals are those of k_resid_restrict.  3D, fine nx = 128 W.
// B16: the fine weights and right side come from compact storage K.
// AL: the fine level has alpha (+ alpha phi in each residual, as one fma).
template <typename T, bool B16, bool AL = false>
__global__ void __launch_bounds__(256) k_rr_rows(Lev<T> F, const T *__restrict__ phi, Lev<T> C,
                                                RowMarchGrid G, Compact16 K)
{
    static_assert(!B16 || sizeof(T) == 4, "compact storage widens to FP32");
    static_assert(!(AL && B16), "compact storage carries no alpha");
    using T2 = typename Vec2<T>::type;
    constexpr unsigned ES = sizeof(T);
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int bidx = (blockIdx.x & 7) * G.per_xcd + (blockIdx.x >> 3);
    const int gw = bidx * 4 + wv;
    if (gw >= G.total) return;
    const int nx = F.nx, ny = F.ny, nz = F.nz, plane = nx * ny;
    const int seg = gw % G.W, t = gw / G.W;
    const int J = t % C.ny, chunk = t / C.ny;
    const int K0 = chunk * G.tz, K1 = min(K0 + G.tz, C.nz);
    const int I = lane + 64 * seg, x0 = 2 * I;
    const int j0 = 2 * J, j1 = j0 + 1;
    // rows j0-1 and j1+1 (y-neighbours outside the pair), x-neighbour columns
    const bool okym = j0 > 0 || F.pery, okyp = j1 < ny - 1 || F.pery;
    const int jm = j0 > 0 ? j0 - 1 : ny - 1, jp = j1 < ny - 1 ? j1 + 1 : 0;
    const bool okxm = x0 > 0 || F.perx, okxp = x0 + 2 < nx || F.perx;
    const int xm = x0 > 0 ? x0 - 1 : nx - 1, xp = x0 + 2 < nx ? x0 + 2 : 0;
    const unsigned o0 = (unsigned)(x0 + nx * j0), o1 = o0 + nx;
    const unsigned oym = (unsigned)(x0 + nx * jm), oyp = (unsigned)(x0 + nx * jp);
    const unsigned oxm0 = (unsigned)(xm + nx * j0), oxp0 = (unsigned)(xp + nx * j0);
    auto kwrap = [&](int k, int &kw) -> bool {
        if (k >= 0 && k < nz) { kw = k; return true; }
        if (F.perz) { kw = k < 0 ? k + nz : k - nz; return true; }
        kw = 0;
        return false;
    };
    struct Z {   // own two rows on one plane: old iterate and z-low faces
        T p00, p01, p10, p11, b00, b01, b10, b11;   // [row][column]
    };
    // a column pair of a weight / rhs array at element offset o (base: plane start)
    auto pair = [&](const T *b32, const unsigned short *b16, unsigned o) -> T2 {
        if constexpr (B16) {
            const unsigned u = ld_off<unsigned>(b16, o * 2u);
            T2 v;
            v.x = bf16_lo(u);
            v.y = bf16_hi(u);
            return v;
        } else {
            return ld_off<T2>(b32, o * ES);
        }
    };
    auto one = [&](const T *b32, const unsigned short *b16, unsigned o) -> T {
        if constexpr (B16) return bf16_lo((unsigned)ld_u16(b16, o * 2u));
        else return ld_off<T>(b32, o * ES);
    };
    auto loadz = [&](int k) -> Z {
        Z z;
        int kw;
        const bool ok = kwrap(k, kw);
        const size_t pk = (size_t)plane * kw;
        const T2 a = ld_off<T2>(phi + pk, o0 * ES), c = ld_off<T2>(phi + pk, o1 * ES);
        const T2 ba = pair(F.bz + pk, K.bz + pk, o0), bc = pair(F.bz + pk, K.bz + pk, o1);
        z.p00 = ok ? a.x : T(0);
        z.p01 = ok ? a.y : T(0);
        z.p10 = ok ? c.x : T(0);
        z.p11 = ok ? c.y : T(0);
        z.b00 = ba.x;
        z.b01 = ba.y;
        z.b10 = bc.x;
        z.b11 = bc.y;
        return z;
    };
    // residuals of the four fine cells of plane k (k in range)
    auto plane_res = [&](int k, const Z &zm, const Z &zc, const Z &zp, T r[2][2]) {
        const size_t pk = (size_t)plane * k;
        const T *ph = phi + pk;
        const T2 ym = ld_off<T2>(ph, oym * ES), yp = ld_off<T2>(ph, oyp * ES);
        const T xm0 = ld_off<T>(ph, oxm0 * ES), xp0 = ld_off<T>(ph, oxp0 * ES);
        const T xm1 = ld_off<T>(ph, (oxm0 + nx) * ES), xp1 = ld_off<T>(ph, (oxp0 + nx) * ES);
        const T2 bx0 = pair(F.bx + pk, K.bx + pk, o0), bx1 = pair(F.bx + pk, K.bx + pk, o1);
        T bxe0, bxe1;
        if (x0 + 2 < nx) {
            bxe0 = one(F.bx + pk, K.bx + pk, o0 + 2u);
            bxe1 = one(F.bx + pk, K.bx + pk, o1 + 2u);
        } else if constexpr (B16) {
            bxe0 = bf16_lo((unsigned)K.bxh[j0 + ny * k]);
            bxe1 = bf16_lo((unsigned)K.bxh[j1 + ny * k]);
        } else {
            bxe0 = F.bxh[j0 + ny * k];
            bxe1 = F.bxh[j1 + ny * k];
        }
        const T2 by0 = pair(F.by + pk, K.by + pk, o0), by1 = pair(F.by + pk, K.by + pk, o1);
        const T2 byn = (j1 < ny - 1) ? pair(F.by + pk, K.by + pk, o1 + nx)
                                     : pair(F.byh + (size_t)nx * k, K.byh + (size_t)nx * k, (unsigned)x0);
        const T2 rh0 = pair(F.rhs + pk, K.rhs + pk, o0), rh1 = pair(F.rhs + pk, K.rhs + pk, o1);
        T2 al0, al1;
        if constexpr (AL) {
            al0 = ld_off<T2>(F.alpha + pk, o0 * ES);
            al1 = ld_off<T2>(F.alpha + pk, o1 * ES);
        }
        T bzn00, bzn01, bzn10, bzn11;
        if (k < nz - 1) {
            bzn00 = zp.b00;
            bzn01 = zp.b01;
            bzn10 = zp.b10;
            bzn11 = zp.b11;
        } else {
            const T2 ha = pair(F.bzh, K.bzh, o0), hb = pair(F.bzh, K.bzh, o1);
            bzn00 = ha.x;
            bzn01 = ha.y;
            bzn10 = hb.x;
            bzn11 = hb.y;
        }
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                T w[6], p[6];
                const T2 &bxr = b ? bx1 : bx0;
                const T2 &byr = b ? by1 : by0;
                w[0] = a ? bxr.y : bxr.x;
                w[1] = a ? (b ? bxe1 : bxe0) : bxr.y;
                w[2] = a ? byr.y : byr.x;
                w[3] = b ? (a ? byn.y : byn.x) : (a ? by1.y : by1.x);
                w[4] = b ? (a ? zc.b11 : zc.b10) : (a ? zc.b01 : zc.b00);
                w[5] = b ? (a ? bzn11 : bzn10) : (a ? bzn01 : bzn00);
                const T pr0 = b ? zc.p10 : zc.p00, pr1 = b ? zc.p11 : zc.p01;   // this row
                p[0] = a ? pr0 : (okxm ? (b ? xm1 : xm0) : T(0));
                p[1] = a ? (okxp ? (b ? xp1 : xp0) : T(0)) : pr1;
                p[2] = b ? (a ? zc.p01 : zc.p00) : (okym ? (a ? ym.y : ym.x) : T(0));
                p[3] = b ? (okyp ? (a ? yp.y : yp.x) : T(0)) : (a ? zc.p11 : zc.p10);
                p[4] = b ? (a ? zm.p11 : zm.p10) : (a ? zm.p01 : zm.p00);
                p[5] = b ? (a ? zp.p11 : zp.p10) : (a ? zp.p01 : zp.p00);
                const T pcc = a ? pr1 : pr0;
                const T rh = b ? (a ? rh1.y : rh1.x) : (a ? rh0.y : rh0.x);
                T ap = apply_value<3, T>(F, w, p, pcc);
                if constexpr (AL) ap = dfma(b ? (a ? al1.y : al1.x) : (a ? al0.y : al0.x), pcc, ap);
                r[b][a] = rh - ap;
            }
    };
    Z zm = loadz(2 * K0 - 1), zc = loadz(2 * K0), zp = loadz(2 * K0 + 1);
    const unsigned oc = (unsigned)(I + C.nx * J);
    for (int K = K0; K < K1; ++K) {
        T s[2][2];   // [c][b]: pair sums of fine plane 2K+c, row b
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int k = 2 * K + c;
            const Z zn = loadz(k + 2);
            T r[2][2];
            plane_res(k, zm, zc, zp, r);
            s[c][0] = r[0][0] + r[0][1];
            s[c][1] = r[1][0] + r[1][1];
            zm = zc;
            zc = zp;
            zp = zn;
        }
        const T v = T(0.125) * ((s[0][0] + s[0][1]) + (s[1][0] + s[1][1]));
        const size_t cp = (size_t)C.nx * C.ny * K;
        st_off<T>(C.rhs + cp, oc * ES, v);
        st_off<T>(C.phi + cp, oc * ES, T(0));
    }
}

// k_rr_rows for the rho form of an FP32 finest level (x and y periodic): the
// face weights of the residual formed from rho as in k_gsrb_rows_rho (the
// domain's z sides from the slabs).  Same march and restriction.
__global__ void __launch_bounds__(256) k_rr_rows_rho(Lev<float> F, const float *__restrict__ phi, Lev<float> C,
                                                    RowMarchGrid G, RhoLev R)
{
    using T = float;
    using T2 = float2;
    constexpr unsigned ES = 4u;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int bidx = (blockIdx.x & 7) * G.per_xcd + (blockIdx.x >> 3);
    const int gw = bidx * 4 + wv;
    if (gw >= G.total) return;
    const int nx = F.nx, ny = F.ny, nz = F.nz, plane = nx * ny;
    const int seg = gw % G.W, t = gw / G.W;
    const int J = t % C.ny, chunk = t / C.ny;
    const int K0 = chunk * G.tz, K1 = min(K0 + G.tz, C.nz);
    const int I = lane + 64 * seg, x0 = 2 * I;
    const int j0 = 2 * J, j1 = j0 + 1;
    const int jm = j0 > 0 ? j0 - 1 : ny - 1, jp = j1 < ny - 1 ? j1 + 1 : 0;   // y periodic
    const int xm = x0 > 0 ? x0 - 1 : nx - 1, xp = x0 + 2 < nx ? x0 + 2 : 0;   // x periodic
    const unsigned o0 = (unsigned)(x0 + nx * j0), o1 = o0 + nx;
    const unsigned oym = (unsigned)(x0 + nx * jm), oyp = (unsigned)(x0 + nx * jp);
    const unsigned oxm0 = (unsigned)(xm + nx * j0), oxp0 = (unsigned)(xp + nx * j0);
    auto kwrap = [&](int k, int &kw) -> bool {
        if (k >= 0 && k < nz) { kw = k; return true; }
        if (F.perz) { kw = k < 0 ? k + nz : k - nz; return true; }
        kw = 0;
        return false;
    };
    auto rcp = [](T x) { return __builtin_amdgcn_rcpf(x); };
    struct Z {   // own two rows on one plane: old iterate and rho
        T p00, p01, p10, p11, r00, r01, r10, r11;   // [row][column]
    };
    auto loadz = [&](int k) -> Z {
        Z z;
        int kw;
        const bool ok = kwrap(k, kw);
        const size_t pk = (size_t)plane * kw;
        const T2 a = ld_off<T2>(phi + pk, o0 * ES), c = ld_off<T2>(phi + pk, o1 * ES);
        const T2 ra = ld_off<T2>(R.rho + pk, o0 * ES), rc = ld_off<T2>(R.rho + pk, o1 * ES);
        z.p00 = ok ? a.x : T(0);
        z.p01 = ok ? a.y : T(0);
        z.p10 = ok ? c.x : T(0);
        z.p11 = ok ? c.y : T(0);
        z.r00 = ra.x;
        z.r01 = ra.y;
        z.r10 = rc.x;
        z.r11 = rc.y;
        return z;
    };
    // z faces f of the four cells (between planes f-1 and f); cf: the face's
    // factor (loaded with the plane's neighbours, a plane ahead)
    auto facez = [&](int f, const Z &lo, const Z &hi, T cf, T w[2][2]) {
        if (f == 0 || f == nz) {
            const float *sl = f == 0 ? R.bzl : F.bzh;
            const T2 a = ld_off<T2>(sl, o0 * ES), c = ld_off<T2>(sl, o1 * ES);
            w[0][0] = a.x; w[0][1] = a.y; w[1][0] = c.x; w[1][1] = c.y;
            return;
        }
        w[0][0] = cf * rcp(lo.r00 + hi.r00);
        w[0][1] = cf * rcp(lo.r01 + hi.r01);
        w[1][0] = cf * rcp(lo.r10 + hi.r10);
        w[1][1] = cf * rcp(lo.r11 + hi.r11);
    };
    // the neighbour rows and columns of plane k (phi and rho), its rhs rows and
    // factors: loaded one plane before they are used (issued at the use, their
    // latency stalled the march); a plane past the domain reads plane 0
    struct Nb {
        T2 ym, yp, rym, ryp, rh0, rh1;
        T xm0, xp0, xm1, xp1, rxm0, rxp0, rxm1, rxp1, c, czl, czh;
    };
    auto loadnb = [&](int k) -> Nb {
        Nb n;
        const int kk = (k >= 0 && k < nz) ? k : 0;
        const size_t pk = (size_t)plane * kk;
        const T *ph = phi + pk, *rh = R.rho + pk;
        n.ym = ld_off<T2>(ph, oym * ES);
        n.yp = ld_off<T2>(ph, oyp * ES);
        n.xm0 = ld_off<T>(ph, oxm0 * ES);
        n.xp0 = ld_off<T>(ph, oxp0 * ES);
        n.xm1 = ld_off<T>(ph, (oxm0 + nx) * ES);
        n.xp1 = ld_off<T>(ph, (oxp0 + nx) * ES);
        n.rym = ld_off<T2>(rh, oym * ES);
        n.ryp = ld_off<T2>(rh, oyp * ES);
        n.rxm0 = ld_off<T>(rh, oxm0 * ES);
        n.rxp0 = ld_off<T>(rh, oxp0 * ES);
        n.rxm1 = ld_off<T>(rh, (oxm0 + nx) * ES);
        n.rxp1 = ld_off<T>(rh, (oxp0 + nx) * ES);
        n.rh0 = ld_off<T2>(F.rhs + pk, o0 * ES);
        n.rh1 = ld_off<T2>(F.rhs + pk, o1 * ES);
        n.c = R.cxy[kk];
        n.czl = R.cz[kk];
        n.czh = R.cz[kk + 1];
        return n;
    };
    auto plane_res = [&](int k, const Z &zm, const Z &zc, const Z &zp, const Nb &nb, T r[2][2]) {
        const T2 ym = nb.ym, yp = nb.yp, rym = nb.rym, ryp = nb.ryp, rh0 = nb.rh0, rh1 = nb.rh1;
        const T xm0 = nb.xm0, xp0 = nb.xp0, xm1 = nb.xm1, xp1 = nb.xp1;
        const T rxm0 = nb.rxm0, rxp0 = nb.rxp0, rxm1 = nb.rxm1, rxp1 = nb.rxp1;
        const T c = nb.c;
        T wzl[2][2], wzh[2][2];
        facez(k, zm, zc, nb.czl, wzl);
        facez(k + 1, zc, zp, nb.czh, wzh);
        // rho of the 4 x 4 neighbourhood rows: [row][col] for cols x0-1 .. x0+2
        const T rr[2][4] = {{rxm0, zc.r00, zc.r01, rxp0}, {rxm1, zc.r10, zc.r11, rxp1}};
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                T w[6], p[6];
                const T rc = rr[b][1 + a];
                w[0] = c * rcp(rr[b][a] + rc);
                w[1] = c * rcp(rc + rr[b][2 + a]);
                const T rlo = b ? (a ? zc.r01 : zc.r00) : (a ? rym.y : rym.x);
                const T rhi = b ? (a ? ryp.y : ryp.x) : (a ? zc.r11 : zc.r10);
                w[2] = c * rcp(rlo + rc);
                w[3] = c * rcp(rc + rhi);
                w[4] = wzl[b][a];
                w[5] = wzh[b][a];
                const T pr0 = b ? zc.p10 : zc.p00, pr1 = b ? zc.p11 : zc.p01;   // this row
                p[0] = a ? pr0 : (b ? xm1 : xm0);
                p[1] = a ? (b ? xp1 : xp0) : pr1;
                p[2] = b ? (a ? zc.p01 : zc.p00) : (a ? ym.y : ym.x);
                p[3] = b ? (a ? yp.y : yp.x) : (a ? zc.p11 : zc.p10);
                p[4] = b ? (a ? zm.p11 : zm.p10) : (a ? zm.p01 : zm.p00);
                p[5] = b ? (a ? zp.p11 : zp.p10) : (a ? zp.p01 : zp.p00);
                const T pcc = a ? pr1 : pr0;
                const T rhv = b ? (a ? rh1.y : rh1.x) : (a ? rh0.y : rh0.x);
                r[b][a] = rhv - apply_value<3, T>(F, w, p, pcc);
            }
    };
    Z zm = loadz(2 * K0 - 1), zc = loadz(2 * K0), zp = loadz(2 * K0 + 1);
    Nb nb = loadnb(2 * K0);
    const unsigned oc = (unsigned)(I + C.nx * J);
    for (int K = K0; K < K1; ++K) {
        T s2[2][2];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int k = 2 * K + c;
            const Z zn = loadz(k + 2);
            const Nb nbn = loadnb(k + 1);
            T r[2][2];
            plane_res(k, zm, zc, zp, nb, r);
            nb = nbn;
            s2[c][0] = r[0][0] + r[0][1];
            s2[c][1] = r[1][0] + r[1][1];
            zm = zc;
            zc = zp;
            zp = zn;
        }
        const T v = T(0.125) * ((s2[0][0] + s2[0][1]) + (s2[1][0] + s2[1][1]));
        const size_t cp = (size_t)C.nx * C.ny * K;
        st_off<T>(C.rhs + cp, oc * ES, v);
        st_off<T>(C.phi + cp, oc * ES, T(0));
    }
}

// max |x| over n values (the norm of the solver's right side)
__global__ void __launch_bounds__(TPB) k_absmax(const double *x, long n, unsigned long long *nrm)
{
    double a = 0.0;
    for (long q = (long)blockIdx.x * TPB + threadIdx.x; q < n; q += (long)gridDim.x * TPB)
        a = fmax(a, fabs(x[q]));
    block_max_atomic(a, nrm);
}

// x -= mean and max |x| of the result in one pass (the solvability shift of
// a singular problem and the norm of its right side)
__global__ void __launch_bounds__(TPB) k_sub_mean_absmax(double *x, long n, const double *mean,
                                                         unsigned long long *nrm)
{
    const double m = *mean;
    double a = 0.0;
    for (long q = (long)blockIdx.x * TPB + threadIdx.x; q < n; q += (long)gridDim.x * TPB) {
        const double v = x[q] - m;
        x[q] = v;
        a = fmax(a, fabs(v));
    }
    block_max_atomic(a, nrm);
}

// ------------------------------------------------------- restriction
template <int DIM, typename T>
__global__ void __launch_bounds__(TPB) k_restrict(const T *r, int fnx, int fny, Lev<T> C)
{
    long q = (long)blockIdx.x * TPB + threadIdx.x;
    long nc = (long)C.nx * C.ny * C.nz;
    if (q >= nc) return;
    int I = (int)(q % C.nx), J = (int)((q / C.nx) % C.ny), K = (int)(q / ((long)C.nx * C.ny));
    int i = 2 * I, j = 2 * J;
    auto f = [&](int ii, int jj, int kk) { return r[ii + (long)fnx * (jj + (long)fny * kk)]; };
    T v;
    if (DIM == 3) {
        int k = 2 * K;
        v = ((f(i, j, k) + f(i + 1, j, k)) + (f(i, j + 1, k) + f(i + 1, j + 1, k))) +
            ((f(i, j, k + 1) + f(i + 1, j, k + 1)) + (f(i, j + 1, k + 1) + f(i + 1, j + 1, k + 1)));
        C.rhs[q] = T(0.125) * v;
    } else {
        v = (f(i, j, 0) + f(i + 1, j, 0)) + (f(i, j + 1, 0) + f(i + 1, j + 1, 0));
        C.rhs[q] = T(0.25) * v;
    }
}

// ------------------------------------------------------ prolongation
template <int DIM, typename T>
__global__ void __launch_bounds__(TPB) k_prolong(Lev<T> F, const T *e, int cnx, int cny, int cnz,
                                                 EllBC bc)
{
    for_rows(F.nx, F.ny, F.nz, [&](int i, int j, int k, long q) {
        F.phi[q] = F.phi[q] + prolong_value<DIM, T>(e, cnx, cny, cnz, bc, i, j, k);
    });
}

// The same prolongation with one wave per fine (x-row segment, chunk of pz
// fine planes): the row's coarse rows and the x/y weights are wave-uniform,
// each lane owns the fine column pair 2I, 2I+1 (coarse column I) and reads
// the coarse values at I-1, I, I+1.  The wave marches the chunk's coarse
// planes: the x/y-interpolated correction of a coarse plane is formed once
// and serves the two fine planes it parents and the fine planes on either
// side that take it as their second z point (the row form above formed two
// per fine plane).  The fine pair is loaded and stored as a vector.
// Evaluation order is that of prolong_value.  3D, fine nx = 128 W, fine
// ny, nz twice the coarse extents, pz even.
template <typename T>
__global__ void __launch_bounds__(256) k_prolong_rows(Lev<T> F, const T *__restrict__ e, int cnx, int cny,
                                                     int cnz, EllBC bc, int W, int pz, int total)
{
    using T2 = typename Vec2<T>::type;
    constexpr unsigned ES = sizeof(T);
    const int lane = threadIdx.x & 63;
    const int gw = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (gw >= total) return;
    const int seg = gw % W, t = gw / W;
    const int j = t % F.ny, ch = t / F.ny;
    const int Kb = ch * (pz >> 1), Ke = min(Kb + (pz >> 1), cnz);
    const int I = lane + 64 * seg, x0 = 2 * I;
    // coarse x neighbours of the pair: I-1 for the even fine cell, I+1 for the odd
    int xa0, xa1, xb0, xb1;
    T wa0, wa1, wb0, wb1;
    lin_w<T>(x0, cnx, bc.k[0][0], bc.k[0][1], xa0, xa1, wa0, wa1);
    lin_w<T>(x0 + 1, cnx, bc.k[0][0], bc.k[0][1], xb0, xb1, wb0, wb1);
    int yi0, yi1;
    T wy0, wy1;
    lin_w<T>(j, cny, bc.k[1][0], bc.k[1][1], yi0, yi1, wy0, wy1);
    const int yy[2] = {yi0, yi1};
    // the pair's x/y-interpolated correction on coarse plane K
    auto plane = [&](int K) -> T2 {
        T ya[2], yb[2];
#pragma unroll
        for (int cy = 0; cy < 2; ++cy) {
            const T *row = e + (size_t)cnx * (yy[cy] + (size_t)cny * K);
            const T c0 = ld_off<T>(row, (unsigned)xa0 * ES);   // = coarse I
            ya[cy] = dfma(wa1, ld_off<T>(row, (unsigned)xa1 * ES), wa0 * c0);
            yb[cy] = dfma(wb1, ld_off<T>(row, (unsigned)xb1 * ES), wb0 * c0);
        }
        T2 v;
        v.x = dfma(wy1, ya[1], wy0 * ya[0]);
        v.y = dfma(wy1, yb[1], wy0 * yb[0]);
        return v;
    };
    const unsigned off = (unsigned)(x0 + F.nx * j) * ES;
    const size_t fplane = (size_t)F.nx * F.ny;
    // fine plane k: its parent's correction vp and its second z point's vq
    auto fine = [&](int k, T w0, T w1, const T2 &vp, const T2 &vq) {
        T *fp = F.phi + fplane * k;
        T2 p = ld_off<T2>(fp, off);
        p.x = p.x + dfma(w1, vq.x, w0 * vp.x);
        p.y = p.y + dfma(w1, vq.y, w0 * vp.y);
        st_off<T2>(fp, off, p);
    };
    // the second z point of fine plane 2K is coarse plane K-1 (its periodic
    // image, or K itself at a non-periodic side), that of 2K+1 is K+1
    int zp, zq;
    T w0, w1;
    lin_w<T>(2 * Kb, cnz, bc.k[2][0], bc.k[2][1], zp, zq, w0, w1);
    T2 vc = plane(Kb);
    T2 vm = (zq == Kb) ? vc : plane(zq);
    for (int K = Kb; K < Ke; ++K) {
        int zp0, zq0, zp1, zq1;
        T w00, w01, w10, w11;
        lin_w<T>(2 * K, cnz, bc.k[2][0], bc.k[2][1], zp0, zq0, w00, w01);
        lin_w<T>(2 * K + 1, cnz, bc.k[2][0], bc.k[2][1], zp1, zq1, w10, w11);
        const T2 vn = (zq1 == K) ? vc : plane(zq1);
        fine(2 * K, w00, w01, vc, (zq0 == K) ? vc : vm);
        fine(2 * K + 1, w10, w11, vc, vn);
        vm = vc;
        vc = vn;
    }
}

template <typename T>
__global__ void __launch_bounds__(TPB) k_fill(T *x, long n, T v)
{
    long q = (long)blockIdx.x * TPB + threadIdx.x;
    if (q < n) x[q] = v;
}

// -------------------------------------------- single-workgroup coarse cycle
// The cycle (V, W or F) over levels lc..nlev-1 and the FP64 CG bottom
// solve, all inside one 512-thread workgroup (these levels are latency
// bound; one launch replaces dozens).  When they fit, the levels are
// staged in LDS for the whole cycle; otherwise they stay in global memory
// (L2 resident).  512 threads leave 256 VGPRs per lane, so the inlined
// cycle does not spill.
static constexpr int CTPB = 512;

__device__ __forceinline__ double blk_sum(double v, double *sh)
{
    // fixed-order block sum (deterministic)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int wv = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[wv] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += sh[w];
    __syncthreads();
    return s;
}

__device__ __forceinline__ double blk_max(double v, double *sh)
{
    v = wave_max(v);
    const int wv = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[wv] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s = fmax(s, sh[w]);
    __syncthreads();
    return s;
}

template <int DIM, typename T>
__device__ void blk_gsrb(const Lev<T> &L)
{
    const int hx = (L.nx + 1) >> 1;
    const int nq = hx * L.ny * L.nz;
    for (int color = 0; color < 2; ++color) {
        for (int q = threadIdx.x; q < nq; q += blockDim.x) {
            const int r = q / hx, m = q - r * hx;
            const int k = r / L.ny, j = r - k * L.ny;
            int i = 2 * m + ((j + k + color) & 1);
            if (i >= L.nx) continue;
            const int c = i + L.nx * (j + L.ny * k);
            T w[6], p[6];
            stencil<DIM, T>(L, L.phi, i, j, k, c, w, p);
            L.phi[c] = gs_value<DIM, T>(L, w, p, L.rhs[c], alpha_at(L, c));
        }
        __syncthreads();
    }
}

// A x in FP64 with the level's (T) weights; x, out are FP64 arrays
template <int DIM, typename T>
__device__ void blk_apply64(const Lev<T> &L, const double *x, double *out)
{
    const int n = L.nx * L.ny * L.nz;
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        const int r = q / L.nx, i = q - r * L.nx;
        const int k = r / L.ny, j = r - k * L.ny;
        Lev<double> D;
        D.nx = L.nx; D.ny = L.ny; D.nz = L.nz;
        D.perx = L.perx; D.pery = L.pery; D.perz = L.perz;
        D.ih2x = L.ih2x; D.ih2y = L.ih2y; D.ih2z = L.ih2z;
        T wt[6], pt[6];
        (void)pt;
        // weights from the level, neighbour values from x
        wt[0] = L.bx[q];
        wt[1] = (i < L.nx - 1) ? L.bx[q + 1] : L.bxh[j + L.ny * k];
        wt[2] = L.by[q];
        wt[3] = (j < L.ny - 1) ? L.by[q + L.nx] : L.byh[i + L.nx * k];
        if (DIM == 3) {
            const int sz = L.nx * L.ny;
            wt[4] = L.bz[q];
            wt[5] = (k < L.nz - 1) ? L.bz[q + sz] : L.bzh[i + L.nx * j];
        }
        double w[6], p[6];
        const int sy = L.nx, sz = L.nx * L.ny;
        for (int s = 0; s < 2 * DIM; ++s) w[s] = (double)wt[s];
        p[0] = (i > 0) ? x[q - 1] : (L.perx ? x[q + L.nx - 1] : 0.0);
        p[1] = (i < L.nx - 1) ? x[q + 1] : (L.perx ? x[q - (L.nx - 1)] : 0.0);
        p[2] = (j > 0) ? x[q - sy] : (L.pery ? x[q + (L.ny - 1) * sy] : 0.0);
        p[3] = (j < L.ny - 1) ? x[q + sy] : (L.pery ? x[q - (L.ny - 1) * sy] : 0.0);
        if (DIM == 3) {
            p[4] = (k > 0) ? x[q - sz] : (L.perz ? x[q + (L.nz - 1) * sz] : 0.0);
            p[5] = (k < L.nz - 1) ? x[q + sz] : (L.perz ? x[q - (L.nz - 1) * sz] : 0.0);
        } else {
            w[4] = w[5] = p[4] = p[5] = 0.0;
        }
        const double ap = apply_value<DIM, double>(D, w, p, x[q]);
        out[q] = L.alpha ? fma((double)L.alpha[q], x[q], ap) : ap;
    }
    __syncthreads();
}

template <int DIM, typename T>
__device__ void blk_bottom_cg(const Lev<T> &L, double *cg, double eps, int maxit, int singular,
                              double *sh)
{
    const int n = L.nx * L.ny * L.nz;
    double *x = cg, *r = cg + n, *p = cg + 2 * n, *Ap = cg + 3 * n;
    // rhs (mean removed on singular levels) and initial guess
    double m = 0.0;
    if (singular) {
        double s = 0.0;
        for (int q = threadIdx.x; q < n; q += blockDim.x) s += (double)L.rhs[q];
        m = blk_sum(s, sh) / (double)n;
    }
    for (int q = threadIdx.x; q < n; q += blockDim.x) x[q] = (double)L.phi[q];
    __syncthreads();
    blk_apply64<DIM, T>(L, x, Ap);
    double rs = 0.0;
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        r[q] = ((double)L.rhs[q] - m) - Ap[q];
        rs += r[q];
    }
    if (singular) {
        double rm = blk_sum(rs, sh) / (double)n;
        for (int q = threadIdx.x; q < n; q += blockDim.x) r[q] -= rm;
    }
    __syncthreads();
    double a = 0.0, rr = 0.0;
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        a = fmax(a, fabs(r[q]));
        rr += r[q] * r[q];
        p[q] = r[q];
    }
    double r0 = blk_max(a, sh);
    rr = blk_sum(rr, sh);
    if (r0 > 0.0) {
        for (int it = 0; it < maxit; ++it) {
            blk_apply64<DIM, T>(L, p, Ap);
            double pap = 0.0;
            for (int q = threadIdx.x; q < n; q += blockDim.x) pap += p[q] * Ap[q];
            pap = blk_sum(pap, sh);
            if (!(pap > 0.0)) break;
            double alpha = rr / pap;
            double rsum = 0.0;
            for (int q = threadIdx.x; q < n; q += blockDim.x) {
                x[q] += alpha * p[q];
                r[q] -= alpha * Ap[q];
                rsum += r[q];
            }
            if (singular) {
                double rm = blk_sum(rsum, sh) / (double)n;
                for (int q = threadIdx.x; q < n; q += blockDim.x) r[q] -= rm;
            }
            __syncthreads();
            double am = 0.0, rr1 = 0.0;
            for (int q = threadIdx.x; q < n; q += blockDim.x) {
                am = fmax(am, fabs(r[q]));
                rr1 += r[q] * r[q];
            }
            am = blk_max(am, sh);
            rr1 = blk_sum(rr1, sh);
            if (am <= eps * r0) break;
            double beta = rr1 / rr;
            rr = rr1;
            for (int q = threadIdx.x; q < n; q += blockDim.x) p[q] = r[q] + beta * p[q];
            __syncthreads();
        }
    }
    for (int q = threadIdx.x; q < n; q += blockDim.x) L.phi[q] = (T)x[q];
    __syncthreads();
}

// BiCGStab (van der Vorst, SIAM J. Sci. Stat. Comput. 13 (1992) 631-644),
// FBoxLib's default bottom solver (mg_bottom_solver = 1, mg.tex:242-243), in
// FP64 on A x = rhs (- mean): the start, mean projection and stopping rule
// of blk_bottom_cg; s = r - alpha v overwrites r.  cg holds 6 n doubles.
template <int DIM, typename T>
__device__ void blk_bottom_bicgstab(const Lev<T> &L, double *cg, double eps, int maxit, int singular,
                                    double *sh)
{
    const int n = L.nx * L.ny * L.nz;
    double *x = cg, *r = cg + n, *rh = cg + 2 * n, *p = cg + 3 * n, *v = cg + 4 * n, *t = cg + 5 * n;
    double m = 0.0;
    if (singular) {
        double s = 0.0;
        for (int q = threadIdx.x; q < n; q += blockDim.x) s += (double)L.rhs[q];
        m = blk_sum(s, sh) / (double)n;
    }
    for (int q = threadIdx.x; q < n; q += blockDim.x) x[q] = (double)L.phi[q];
    __syncthreads();
    blk_apply64<DIM, T>(L, x, v);
    double rs = 0.0;
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        r[q] = ((double)L.rhs[q] - m) - v[q];
        rs += r[q];
    }
    if (singular) {
        double rm = blk_sum(rs, sh) / (double)n;
        for (int q = threadIdx.x; q < n; q += blockDim.x) r[q] -= rm;
    }
    __syncthreads();
    double a = 0.0;
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        a = fmax(a, fabs(r[q]));
        rh[q] = r[q];
    }
    const double r0 = blk_max(a, sh);
    // r -= c u, then the mean projection on singular levels; returns max |r|
    auto axpy_r = [&](double c, const double *u) {
        double rsum = 0.0;
        for (int q = threadIdx.x; q < n; q += blockDim.x) {
            r[q] -= c * u[q];
            rsum += r[q];
        }
        if (singular) {
            double rm = blk_sum(rsum, sh) / (double)n;
            for (int q = threadIdx.x; q < n; q += blockDim.x) r[q] -= rm;
        }
        __syncthreads();
        double am = 0.0;
        for (int q = threadIdx.x; q < n; q += blockDim.x) am = fmax(am, fabs(r[q]));
        return blk_max(am, sh);
    };
    auto dot = [&](const double *u, const double *w) {
        double d = 0.0;
        for (int q = threadIdx.x; q < n; q += blockDim.x) d += u[q] * w[q];
        return blk_sum(d, sh);
    };
    if (r0 > 0.0) {
        double rho_old = 1.0, al = 1.0, om = 1.0;
        for (int it = 0; it < maxit; ++it) {
            const double rho = dot(rh, r);
            if (rho == 0.0) break;
            if (it == 0) {
                for (int q = threadIdx.x; q < n; q += blockDim.x) p[q] = r[q];
            } else {
                const double beta = (rho / rho_old) * (al / om);
                for (int q = threadIdx.x; q < n; q += blockDim.x) p[q] = r[q] + beta * (p[q] - om * v[q]);
            }
            __syncthreads();
            blk_apply64<DIM, T>(L, p, v);
            const double rv = dot(rh, v);
            if (rv == 0.0) break;
            al = rho / rv;
            if (axpy_r(al, v) <= eps * r0) {
                for (int q = threadIdx.x; q < n; q += blockDim.x) x[q] += al * p[q];
                break;
            }
            blk_apply64<DIM, T>(L, r, t);
            const double tt = dot(t, t);
            if (!(tt > 0.0)) {
                for (int q = threadIdx.x; q < n; q += blockDim.x) x[q] += al * p[q];
                break;
            }
            om = dot(t, r) / tt;
            for (int q = threadIdx.x; q < n; q += blockDim.x) x[q] += al * p[q] + om * r[q];
            rho_old = rho;
            if (axpy_r(om, t) <= eps * r0 || om == 0.0) break;
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < n; q += blockDim.x) L.phi[q] = (T)x[q];
    __syncthreads();
}

// sums over the 64 lanes of a wave in a fixed order (butterfly), so every
// lane holds the same, deterministic result
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// The bottom solve of a level of at most 64 cells inside one wave: nub
// red-black sweeps, the FP64 CG of blk_bottom_cg or BiCGStab of
// blk_bottom_bicgstab (mean projected out on singular problems; neither for
// smoothing only), nub sweeps.  Lane q owns cell q; neighbour values come
// by lane shuffles (zero outside a non-periodic side), so the wave needs no
// barrier.  Same updates and stopping rule as the block form; only the order
// of the CG's sums differs.
template <int DIM, typename T>
__device__ void wave_bottom(const Lev<T> &L, const CoarseOpts &o)
{
    const int q = threadIdx.x & 63;
    const int nx = L.nx, ny = L.ny, nz = L.nz, n = nx * ny * nz, plane = nx * ny;
    const bool live = q < n;
    const int qc = live ? q : 0;
    const int i = qc % nx, j = (qc / nx) % ny, k = qc / plane;
    // neighbour lanes (-1: outside a non-periodic side) and face weights
    int nb[6];
    nb[0] = i > 0 ? qc - 1 : (L.perx ? qc + nx - 1 : -1);
    nb[1] = i < nx - 1 ? qc + 1 : (L.perx ? qc - (nx - 1) : -1);
    nb[2] = j > 0 ? qc - nx : (L.pery ? qc + (ny - 1) * nx : -1);
    nb[3] = j < ny - 1 ? qc + nx : (L.pery ? qc - (ny - 1) * nx : -1);
    if (DIM == 3) {
        nb[4] = k > 0 ? qc - plane : (L.perz ? qc + (nz - 1) * plane : -1);
        nb[5] = k < nz - 1 ? qc + plane : (L.perz ? qc - (nz - 1) * plane : -1);
    } else {
        nb[4] = nb[5] = -1;
    }
    T w[6];
    w[0] = L.bx[qc];
    w[1] = (i < nx - 1) ? L.bx[qc + 1] : L.bxh[j + ny * k];
    w[2] = L.by[qc];
    w[3] = (j < ny - 1) ? L.by[qc + nx] : L.byh[i + nx * k];
    if (DIM == 3) {
        w[4] = L.bz[qc];
        w[5] = (k < nz - 1) ? L.bz[qc + plane] : L.bzh[i + nx * j];
    } else {
        w[4] = w[5] = T(0);
    }
    const int colour = (i + j + k) & 1;
    T phi = live ? L.phi[qc] : T(0);
    const T rhs = live ? L.rhs[qc] : T(0);
    const T acell = alpha_at(L, qc);
    auto gather = [&](auto v, auto *p) {   // neighbour values of v (all lanes shuffle)
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            const auto t = __shfl(v, nb[s] < 0 ? q : nb[s], 64);
            p[s] = nb[s] < 0 ? decltype(v)(0) : t;
        }
    };
    auto sweeps = [&](int count) {
        for (int it = 0; it < count; ++it)
            for (int c = 0; c < 2; ++c) {
                T p[6];
                gather(phi, p);
                const T v = gs_value<DIM, T>(L, w, p, rhs, acell);
                if (live && colour == c) phi = v;
            }
    };
    sweeps(o.nub);
    // ---- CG in FP64 on A x = rhs (- mean)
    Lev<double> D;
    D.nx = nx; D.ny = ny; D.nz = nz;
    D.perx = L.perx; D.pery = L.pery; D.perz = L.perz;
    D.ih2x = L.ih2x; D.ih2y = L.ih2y; D.ih2z = L.ih2z;
    double wd[6];
#pragma unroll
    for (int s = 0; s < 6; ++s) wd[s] = (double)w[s];
    auto apply = [&](double x) -> double {
        double p[6];
        gather(x, p);
        if (!live) return 0.0;
        const double ap = apply_value<DIM, double>(D, wd, p, x);
        return L.alpha ? fma((double)acell, x, ap) : ap;
    };
    const double nd = (double)n;
    const double m = o.singular ? wave_sum(live ? (double)rhs : 0.0) / nd : 0.0;
    double x = live ? (double)phi : 0.0;
    double r = live ? ((double)rhs - m) - apply(x) : 0.0;
    if (o.singular) {
        const double rm = wave_sum(r) / nd;
        if (live) r -= rm;
    }
    const double r0 = wave_max(fabs(r));
    if (r0 > 0.0 && o.bottom == MAC_BOTTOM_BICGSTAB) {
        // the BiCGStab of blk_bottom_bicgstab, one cell per lane
        const double rh = r;
        double p = 0.0, v = 0.0, rho_old = 1.0, al = 1.0, om = 1.0;
        auto project = [&]() {
            if (o.singular) {
                const double rm = wave_sum(r) / nd;
                if (live) r -= rm;
            }
        };
        for (int it = 0; it < o.bottom_maxit; ++it) {
            const double rho = wave_sum(rh * r);
            if (rho == 0.0) break;
            if (it == 0) p = r;
            else p = r + ((rho / rho_old) * (al / om)) * (p - om * v);
            v = apply(p);
            const double rv = wave_sum(rh * v);
            if (rv == 0.0) break;
            al = rho / rv;
            r -= al * v;
            project();
            if (wave_max(fabs(r)) <= o.bottom_eps * r0) {
                x += al * p;
                break;
            }
            const double t = apply(r);
            const double tt = wave_sum(t * t);
            if (!(tt > 0.0)) {
                x += al * p;
                break;
            }
            om = wave_sum(t * r) / tt;
            x += al * p + om * r;
            r -= om * t;
            project();
            rho_old = rho;
            if (wave_max(fabs(r)) <= o.bottom_eps * r0 || om == 0.0) break;
        }
    } else if (r0 > 0.0 && o.bottom == MAC_BOTTOM_CG) {
        double p = r;
        double rr = wave_sum(r * r);
        for (int it = 0; it < o.bottom_maxit; ++it) {
            const double Ap = apply(p);
            const double pap = wave_sum(p * Ap);
            if (!(pap > 0.0)) break;
            const double alpha = rr / pap;
            x += alpha * p;
            r -= alpha * Ap;
            if (o.singular) {
                const double rm = wave_sum(r) / nd;
                if (live) r -= rm;
            }
            const double am = wave_max(fabs(r));
            const double rr1 = wave_sum(r * r);
            if (am <= o.bottom_eps * r0) break;
            const double beta = rr1 / rr;
            rr = rr1;
            p = r + beta * p;
        }
    }
    phi = live ? (T)x : T(0);
    sweeps(o.nub);
    if (live) L.phi[qc] = phi;
}

template <int DIM, typename T>
__device__ void blk_resid_restrict(const Lev<T> &F, const Lev<T> &C)
{
    const int nc = C.nx * C.ny * C.nz;
    for (int q = threadIdx.x; q < nc; q += blockDim.x) {
        const int rq = q / C.nx, I = q - rq * C.nx;
        const int K = rq / C.ny, J = rq - K * C.ny;
        T rv[2][2][2];
        for (int c = 0; c < (DIM == 3 ? 2 : 1); ++c)
            for (int b = 0; b < 2; ++b)
                for (int a = 0; a < 2; ++a) {
                    int i = 2 * I + a, j = 2 * J + b, k = (DIM == 3) ? 2 * K + c : 0;
                    const int f = i + F.nx * (j + F.ny * k);
                    T w[6], p[6];
                    stencil<DIM, T>(F, F.phi, i, j, k, f, w, p);
                    rv[c][b][a] = F.rhs[f] - with_alpha(F, f, apply_value<DIM, T>(F, w, p, F.phi[f]), F.phi[f]);
                }
        T v;
        if (DIM == 3) {
            v = ((rv[0][0][0] + rv[0][0][1]) + (rv[0][1][0] + rv[0][1][1])) +
                ((rv[1][0][0] + rv[1][0][1]) + (rv[1][1][0] + rv[1][1][1]));
            C.rhs[q] = T(0.125) * v;
        } else {
            v = (rv[0][0][0] + rv[0][0][1]) + (rv[0][1][0] + rv[0][1][1]);
            C.rhs[q] = T(0.25) * v;
        }
        C.phi[q] = T(0);
    }
    __syncthreads();
}

template <int DIM, typename T>
__device__ void blk_prolong(const Lev<T> &F, const Lev<T> &C, const EllBC &bc)
{
    const int n = F.nx * F.ny * F.nz;
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        const int r = q / F.nx, i = q - r * F.nx;
        const int k = r / F.ny, j = r - k * F.ny;
        F.phi[q] = F.phi[q] + prolong_value<DIM, T>(C.phi, C.nx, C.ny, C.nz, bc, i, j, k);
    }
    __syncthreads();
}

// LDS image of the coarse levels: per level phi, rhs, bx, by, (bz), then the
// high-face slabs, each rounded to 16 bytes; after the last level the four
// FP64 CG vectors.  The host sizes the allocation with the same function.
__host__ __device__ inline int lds_round(int bytes) { return (bytes + 15) & ~15; }

template <int DIM, typename T>
__host__ __device__ inline int lds_level_bytes(int nx, int ny, int nz)
{
    const int n = nx * ny * nz, s = (int)sizeof(T);
    int b = 4 * lds_round(n * s) + lds_round(ny * nz * s) + lds_round(nx * nz * s);
    if (DIM == 3) b += lds_round(n * s) + lds_round(nx * ny * s);
    return b;
}

template <int DIM, typename T>
__device__ Lev<T> lds_view(const Lev<T> &G, char *p)
{
    Lev<T> S = G;
    const int n = G.nx * G.ny * G.nz;
    auto carve = [&](int cnt) {
        T *r = (T *)p;
        p += lds_round(cnt * (int)sizeof(T));
        return r;
    };
    S.phi = carve(n);
    S.rhs = carve(n);
    S.bx = carve(n);
    S.by = carve(n);
    S.bxh = carve(G.ny * G.nz);
    S.byh = carve(G.nx * G.nz);
    if (DIM == 3) {
        S.bz = carve(n);
        S.bzh = carve(G.nx * G.ny);
    }
    S.res = S.phi2 = nullptr;
    return S;
}

// Level l of the coarse cycle: the global level itself, or its LDS image.
template <int DIM, typename T, bool LDS>
__device__ Lev<T> coarse_level(const Lev<T> *levs, int lc, int l, char *smem)
{
    if (!LDS) return levs[l];
    char *p = smem;
    for (int m = lc; m < l; ++m) p += lds_level_bytes<DIM, T>(levs[m].nx, levs[m].ny, levs[m].nz);
    return lds_view<DIM, T>(levs[l], p);
}

// the bottom level: one wave for at most 64 cells (one cell per lane in
// registers, no block barriers), else nub sweeps, the block CG or BiCGStab
// (none for smoothing only), nub sweeps
template <int DIM, typename T>
__device__ __forceinline__ void coarse_bottom(const Lev<T> &L, const CoarseOpts &o, double *cg, double *sh)
{
    if (L.nx * L.ny * L.nz <= 64) {
        if (threadIdx.x < 64) wave_bottom<DIM, T>(L, o);
        __syncthreads();
    } else {
        for (int s = 0; s < o.nub; ++s) blk_gsrb<DIM, T>(L);
        if (o.bottom == MAC_BOTTOM_CG)
            blk_bottom_cg<DIM, T>(L, cg, o.bottom_eps, o.bottom_maxit, o.singular, sh);
        else if (o.bottom == MAC_BOTTOM_BICGSTAB)
            blk_bottom_bicgstab<DIM, T>(L, cg, o.bottom_eps, o.bottom_maxit, o.singular, sh);
        for (int s = 0; s < o.nub; ++s) blk_gsrb<DIM, T>(L);
    }
}

// VC: a V-cycle (straight down and up); else a W or F cycle
template <int DIM, typename T, bool LDS, bool VC>
__global__ void __launch_bounds__(CTPB) k_coarse_cycle(const Lev<T> *__restrict__ levs, int lc,
                                                       int nlev, CoarseOpts o, double *cg)
{
    extern __shared__ __align__(16) char smem[];
    __shared__ double sh[CTPB / 64];
    const int bot = nlev - 1;
    if (LDS) {
        // stage the weights of every coarse level, and phi/rhs of the first
        // (deeper phi/rhs are written by the restriction before use).  One
        // pass per level: a thread issues every load of its cell and slab
        // entries before it stores any, so each level costs one round of
        // global latency rather than one per array
        for (int l = lc; l <= bot; ++l) {
            const Lev<T> G = levs[l];
            const Lev<T> S = coarse_level<DIM, T, true>(levs, lc, l, smem);
            const int n = G.nx * G.ny * G.nz, nsx = G.ny * G.nz, nsy = G.nx * G.nz;
            const int nsz = (DIM == 3) ? G.nx * G.ny : 0;
            const int m = max(max(n, nsx), max(nsy, nsz));
            for (int q = threadIdx.x; q < m; q += blockDim.x) {
                const bool c = q < n, hx = q < nsx, hy = q < nsy, hz = q < nsz;
                T vx = T(0), vy = T(0), vz = T(0), vp = T(0), vr = T(0), wx = T(0), wy = T(0), wz = T(0);
                if (c) {
                    vx = G.bx[q];
                    vy = G.by[q];
                    if (DIM == 3) vz = G.bz[q];
                    if (l == lc) {
                        vp = G.phi[q];
                        vr = G.rhs[q];
                    }
                }
                if (hx) wx = G.bxh[q];
                if (hy) wy = G.byh[q];
                if (hz) wz = G.bzh[q];
                if (c) {
                    S.bx[q] = vx;
                    S.by[q] = vy;
                    if (DIM == 3) S.bz[q] = vz;
                    if (l == lc) {
                        S.phi[q] = vp;
                        S.rhs[q] = vr;
                    }
                }
                if (hx) S.bxh[q] = wx;
                if (hy) S.byh[q] = wy;
                if (hz) S.bzh[q] = wz;
            }
        }
        char *p = smem;
        for (int m = lc; m <= bot; ++m) p += lds_level_bytes<DIM, T>(levs[m].nx, levs[m].ny, levs[m].nz);
        cg = (double *)p;
        __syncthreads();
    }
    if constexpr (VC) {
        for (int l = lc; l < bot; ++l) {
            const Lev<T> L = coarse_level<DIM, T, LDS>(levs, lc, l, smem);
            for (int s = 0; s < o.nu1; ++s) blk_gsrb<DIM, T>(L);
            blk_resid_restrict<DIM, T>(L, coarse_level<DIM, T, LDS>(levs, lc, l + 1, smem));
        }
        coarse_bottom<DIM, T>(coarse_level<DIM, T, LDS>(levs, lc, bot, smem), o, cg, sh);
        for (int l = bot - 1; l >= lc; --l) {
            const Lev<T> L = coarse_level<DIM, T, LDS>(levs, lc, l, smem);
            blk_prolong<DIM, T>(L, coarse_level<DIM, T, LDS>(levs, lc, l + 1, smem), o.bc);
            for (int s = 0; s < o.nu2; ++s) blk_gsrb<DIM, T>(L);
        }
    } else {
        // W and F cycles, without recursion.  Going down, each level owes a
        // second visit of the next coarser level (bit l of `again`; bit l of
        // `again_w`: that visit is a W, else a V); going up, a level that
        // owes one turns back down before its prolongation.  The first visit
        // has the shape of the visiting level.  The masks are the same in
        // every thread, so the block stays in step.
        unsigned again = 0u, again_w = 0u;
        bool w = o.cycle == MAC_CYCLE_W;   // shape of the current visit (W, or F/V below)
        bool v = false;                     // the current visit is a V
        int l = lc;
        for (;;) {
            for (; l < bot; ++l) {
                const Lev<T> L = coarse_level<DIM, T, LDS>(levs, lc, l, smem);
                for (int s = 0; s < o.nu1; ++s) blk_gsrb<DIM, T>(L);
                blk_resid_restrict<DIM, T>(L, coarse_level<DIM, T, LDS>(levs, lc, l + 1, smem));
                if (!v) {
                    again |= 1u << l;
                    if (w) again_w |= 1u << l;
                    else again_w &= ~(1u << l);
                }
            }
            coarse_bottom<DIM, T>(coarse_level<DIM, T, LDS>(levs, lc, bot, smem), o, cg, sh);
            bool down = false;
            while (l > lc) {
                --l;
                if (again & (1u << l)) {   // second visit of level l + 1
                    again &= ~(1u << l);
                    w = (again_w >> l) & 1u;
                    v = !w;
                    ++l;
                    down = true;
                    break;
                }
                const Lev<T> L = coarse_level<DIM, T, LDS>(levs, lc, l, smem);
                blk_prolong<DIM, T>(L, coarse_level<DIM, T, LDS>(levs, lc, l + 1, smem), o.bc);
                for (int s = 0; s < o.nu2; ++s) blk_gsrb<DIM, T>(L);
            }
            if (!down) break;
        }
    }
    if (LDS) {
        const Lev<T> G = levs[lc];
        const Lev<T> S = coarse_level<DIM, T, true>(levs, lc, lc, smem);
        const int n = G.nx * G.ny * G.nz;
        for (int q = threadIdx.x; q < n; q += blockDim.x) G.phi[q] = S.phi[q];
    }
}

// ------------------------------------------------------ face correction
// Valid faces of every box: U = (beta0 U - w (phi_hi - phi_lo)/dx) / beta0.
// One thread per (i, j) column of a box marching k through a chunk of ZCH
// planes: each cell corrects its three low faces, and the cells of the box's
// last column, row and plane its high faces too.  phi of the plane below is carried, so each
// phi value is read once from HBM (the x and y neighbours come from the cache
// lines the neighbouring lanes read), and the thread also copies phi into the
// caller's phi box.  32-bit indexing.
template <int DIM>
__global__ void __launch_bounds__(TPB) k_correct_valid(CorrectArgs A)
{
    const int b = blockIdx.y;
    const BoxDev B = A.boxes[b];
    const int ex = B.hi[0] - B.lo[0] + 1, ey = B.hi[1] - B.lo[1] + 1;
    const int ez = (DIM == 3) ? B.hi[2] - B.lo[2] + 1 : 1;
    const int q = blockIdx.x * TPB + threadIdx.x;
    const int kz0 = (DIM == 3) ? (int)blockIdx.z * ZCH : 0, kz1 = (DIM == 3) ? min(kz0 + ZCH, ez) : 1;
    if (q >= ex * ey || kz0 >= kz1) return;
    const int li = q % ex, lj = q / ex;
    const int i = B.lo[0] + li, j = B.lo[1] + lj;
    const int I = i - A.dlo[0], J = j - A.dlo[1];
    const int nx = A.n[0], ny = A.n[1];
    const int st[3] = {1, nx, nx * ny};   // level strides
    const int rd = DIM - 1;
    Col cu[3], cc[3];
    double *u[3];
    const double *cart[3];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        cu[d] = col_of<DIM>(B, A.ng_u, d, li, lj);
        cc[d] = col_of<DIM>(B, A.ng_cart, d, li, lj);
        u[d] = A.umac[b * DIM + d];
        cart[d] = A.cart ? A.cart[b * DIM + d] : nullptr;
    }
    const Col cp = col_of<DIM>(B, A.ng_phi, -1, li, lj);
    double *pout = A.phi_out ? A.phi_out[b] : nullptr;
    const double *wl[3] = {A.bx, A.by, A.bz};
    // one face: its position in the column (sd: 0 low, 1 high face of the
    // cell), its velocity as read, effective weight and the phi values on both
    // sides; pb: the planar beta0 of plane kk.  The corrected velocity is stored
    auto correct = [&](int d, int sd, int kk, const B0 &pb, double t, double w, double phi_hi, double phi_lo) {
        const int o = cu[d].at(kk, d == 0 ? sd : 0, d == 1 ? sd : 0, d == 2 ? sd : 0);
        double b0 = 1.0;
        bool has = false;
        if (A.beta0) {
            b0 = (d == rd) ? (sd ? pb.hi : pb.lo) : pb.t;
            has = true;
        } else if (A.cart) {
            b0 = cart[d][cc[d].at(kk, d == 0 ? sd : 0, d == 1 ? sd : 0, d == 2 ? sd : 0)];
            has = true;
        }
        if (has) t = t * b0;
        const double gphi = (phi_hi - phi_lo) / A.dx[d];
        t = t - w * gphi;
        if (has) t = t / b0;
        u[d][o] = t;
    };
    auto face_u = [&](int d, int sd, int kk) {
        return u[d][cu[d].at(kk, d == 0 ? sd : 0, d == 1 ? sd : 0, d == 2 ? sd : 0)];
    };
    // phi below the cell's low face in direction d (periodic image, or 0
    // outside a non-periodic side, whose folded weight carries the BC)
    auto below = [&](int d, int Id, int c) -> double {
        if (Id > 0) return A.phi[c - st[d]];
        return (A.bc.k[d][0] == ELL_PER) ? A.phi[c + (A.n[d] - 1) * st[d]] : 0.0;
    };
    const int K0 = (DIM == 3) ? B.lo[2] - A.dlo[2] : 0;
    // what one plane's low faces need: their velocities and weights, phi of
    // the cell and of its x and y low neighbours.  The next plane's are
    // loaded before this plane's faces are stored, so the march never waits
    // on its own stores.
    struct CIn {
        double u[3], w[3], pc, px, py;
    };
    auto load = [&](int kk) -> CIn {
        CIn o;
        const int c = I + nx * (J + ny * (K0 + kk));
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            o.u[d] = d < DIM ? face_u(d, 0, kk) : 0.0;
            o.w[d] = d < DIM ? wl[d][c] : 0.0;
        }
        o.pc = A.phi[c];
        o.px = below(0, I, c);
        o.py = below(1, J, c);
        return o;
    };
    double pz_lo = 0.0;
    if (DIM == 3) pz_lo = below(2, K0 + kz0, I + nx * (J + ny * (K0 + kz0)));
    auto b0load = [&](int kk) -> B0 {
        return planar_b0(A.beta0, A.beta0e, (rd == 2 ? B.lo[2] + kk : j) - A.dlo[rd]);
    };
    CIn cur = load(kz0);
    B0 bcur = b0load(kz0);
    for (int kk = kz0; kk < kz1; ++kk) {
        const CIn nxt = load(min(kk + 1, kz1 - 1));
        const B0 bnxt = b0load(min(kk + 1, kz1 - 1));
        const int K = K0 + kk;
        const int c = I + nx * (J + ny * K);
        const int Ix[3] = {I, J, K};
        const double pc = cur.pc;
        // low faces
        correct(0, 0, kk, bcur, cur.u[0], cur.w[0], pc, cur.px);
        correct(1, 0, kk, bcur, cur.u[1], cur.w[1], pc, cur.py);
        if (DIM == 3) correct(2, 0, kk, bcur, cur.u[2], cur.w[2], pc, pz_lo);
        // the box's high faces
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const bool last = d == 0 ? li == ex - 1 : (d == 1 ? lj == ey - 1 : kk == ez - 1);
            if (!last) continue;
            const int nd = A.n[d];
            const double t = face_u(d, 1, kk);
            if (Ix[d] + 1 < nd) {
                correct(d, 1, kk, bcur, t, wl[d][c + st[d]], A.phi[c + st[d]], pc);
            } else if (A.bc.k[d][0] == ELL_PER) {   // the periodic image of face 0
                const int c0 = c - (nd - 1) * st[d];
                correct(d, 1, kk, bcur, t, wl[d][c0], A.phi[c0], pc);
            } else {                                 // the domain's high face
                const double w = d == 0 ? A.bxh[J + ny * K] : (d == 1 ? A.byh[I + nx * K] : A.bzh[I + nx * J]);
                correct(d, 1, kk, bcur, t, w, 0.0, pc);
            }
        }
        if (pout) pout[cp.at(kk)] = pc;
        pz_lo = pc;
        cur = nxt;
        bcur = bnxt;
    }
}

// ghost faces that lie inside the (periodic) domain take the value of the
// owning box's valid face (multifab_fill_boundary semantics)
template <int DIM>
__global__ void __launch_bounds__(TPB) k_fill_ghost(CorrectArgs A)
{
    const int b = blockIdx.y, d = blockIdx.z;
    const BoxDev B = A.boxes[b];
    const int ng = A.ng_u;
    const int ex = B.hi[0] - B.lo[0] + 1 + 2 * ng + (d == 0);
    const int ey = B.hi[1] - B.lo[1] + 1 + 2 * ng + (d == 1);
    const int ez = (DIM == 3) ? B.hi[2] - B.lo[2] + 1 + 2 * ng + (d == 2) : 1;
    // the thread's ghost face: the grown box minus the valid faces is the
    // z-slabs (full x-y), then the y-slabs of the inner z range, then the
    // x-slabs of the inner y-z range, each ng thick on both sides
    const int ngz = (DIM == 3) ? ng : 0;
    const int iy = ey - 2 * ng, iz = ez - 2 * ngz;
    const int nA = 2 * ngz * ex * ey, nB = 2 * ng * ex * iz, nC = 2 * ng * iy * iz;
    int q = blockIdx.x * TPB + threadIdx.x;
    if (q >= nA + nB + nC) return;
    int x, y, z;
    if (q < nA) {
        const int sl = q / (ex * ey), r = q - sl * ex * ey;
        z = sl < ngz ? sl : ez - 2 * ngz + sl;
        y = r / ex;
        x = r - y * ex;
    } else if ((q -= nA) < nB) {
        const int sl = q / (ex * iz), r = q - sl * ex * iz;
        y = sl < ng ? sl : ey - 2 * ng + sl;
        z = ngz + r / ex;
        x = r - (r / ex) * ex;
    } else {
        q -= nB;
        const int sl = q / (iy * iz), r = q - sl * iy * iz;
        x = sl < ng ? sl : ex - 2 * ng + sl;
        y = ng + r % iy;
        z = ngz + r / iy;
    }
    int g[3] = {B.lo[0] - ng + x, B.lo[1] - ng + y, (DIM == 3) ? B.lo[2] - ng + z : 0};
    // inside this box's valid region?  then nothing to do (not reached)
    bool valid = true;
    for (int e = 0; e < DIM; ++e) {
        int hi = B.hi[e] + (e == d);
        if (g[e] < B.lo[e] || g[e] > hi) valid = false;
    }
    if (valid) return;
    // map to the domain (periodic wrap), give up outside non-periodic sides
    int G[3];
    for (int e = 0; e < DIM; ++e) {
        int x = g[e] - A.dlo[e];
        int ne = A.n[e];
        int top = ne - 1 + (e == d);   // last valid index (faces include n)
        if (x < 0 || x > top) {
            if (A.bc.k[e][0] != ELL_PER) return;
            x = ((x % ne) + ne) % ne;
        }
        G[e] = x;
    }
    if (DIM == 2) G[2] = 0;
    // owner: box whose cell contains the face's cell index (n -> n-1)
    int cell[3];
    for (int e = 0; e < 3; ++e) cell[e] = G[e];
    if (cell[d] == A.n[d]) cell[d] -= 1;
    int mi = cell[0] / A.mapg[0], mj = cell[1] / A.mapg[1], mk = (DIM == 3) ? cell[2] / A.mapg[2] : 0;
    int ob = A.boxmap[mi + A.mapn[0] * (mj + A.mapn[1] * mk)];
    if (ob < 0) {
        // a level whose boxes leave holes (the fine level of a two-level
        // projection): the face may end the box on its low side, else no
        // box holds it
        if (cell[d] == 0) return;
        cell[d] -= 1;
        mi = cell[0] / A.mapg[0]; mj = cell[1] / A.mapg[1]; mk = (DIM == 3) ? cell[2] / A.mapg[2] : 0;
        ob = A.boxmap[mi + A.mapn[0] * (mj + A.mapn[1] * mk)];
        if (ob < 0) return;
    }
    const BoxDev OB = A.boxes[ob];
    int gi = G[0] + A.dlo[0], gj = G[1] + A.dlo[1], gk = (DIM == 3) ? G[2] + A.dlo[2] : 0;
    double v = A.umac[ob * DIM + d][boxoff(OB, ng, d, DIM, gi, gj, gk)];
    A.umac[b * DIM + d][boxoff(B, ng, d, DIM, g[0], g[1], g[2])] = v;
}

// physical edge-velocity BCs on the first ghost layer of one side
// (Exec/SCIENCE/flame/multifab_physbc_edgevel.f90:211-321)
template <int DIM>
__global__ void __launch_bounds__(TPB) k_physbc(CorrectArgs A, int dir, int side, int phys)
{
    const int b = blockIdx.y, comp = blockIdx.z;   // comp = face direction
    const BoxDev B = A.boxes[b];
    if (side == 0 && B.lo[dir] != A.dlo[dir]) return;
    if (side == 1 && B.hi[dir] != A.dlo[dir] + A.n[dir] - 1) return;
    const int ng = A.ng_u;
    if (ng < 1) return;
    // extents of the face array of direction comp (with ghosts)
    int ext[3];
    for (int e = 0; e < 3; ++e) ext[e] = (e < DIM) ? B.hi[e] - B.lo[e] + 1 + 2 * ng + (e == comp) : 1;
    // transverse plane of the array across 'dir'
    int t0 = (dir == 0) ? 1 : 0, t1 = (dir == 2) ? 1 : 2;
    long nt = (long)ext[t0] * (DIM == 3 ? ext[t1] : 1);
    long q = (long)blockIdx.x * TPB + threadIdx.x;
    if (q >= nt) return;
    int a0 = (int)(q % ext[t0]), a1 = (DIM == 3) ? (int)(q / ext[t0]) : 0;
    bool normal = (comp == dir);
    int lo = B.lo[dir], hi = B.hi[dir];
    // index along dir of the target and the source (or -1 for none)
    int tgt, src = -1000000;
    double sign = 1.0, setv = 0.0;
    bool set_zero = false;
    if (side == 0) {
        if (phys == MAC_SLIP_WALL || phys == MAC_NO_SLIP_WALL || phys == MAC_INLET) {
            if (normal) return;   // normal ghost untouched (INLET_VEL is problem data)
            tgt = lo - 1;
            set_zero = true;
        } else if (phys == MAC_OUTLET) {
            tgt = lo - 1;
            src = lo;
        } else if (phys == MAC_SYMMETRY) {
            tgt = lo - 1;
            if (normal) { src = lo + 1; sign = -1.0; }
            else src = lo;
        } else {
            return;
        }
    } else {
        if (phys == MAC_SLIP_WALL || phys == MAC_NO_SLIP_WALL || phys == MAC_INLET) {
            if (normal) return;
            tgt = hi + 1;
            set_zero = true;
        } else if (phys == MAC_OUTLET) {
            if (normal) { tgt = hi + 2; src = hi + 1; }
            else { tgt = hi + 1; src = hi; }
        } else if (phys == MAC_SYMMETRY) {
            if (normal) { tgt = hi + 2; src = hi; sign = -1.0; }
            else { tgt = hi + 1; src = hi; }
        } else {
            return;
        }
    }
    (void)setv;
    int gt[3], gs[3];
    gt[t0] = gs[t0] = B.lo[t0] - ng + a0;
    if (DIM == 3) gt[t1] = gs[t1] = B.lo[t1] - ng + a1;
    else gt[2] = gs[2] = 0;
    gt[dir] = tgt;
    gs[dir] = src;
    double *u = A.umac[b * DIM + comp];
    long ot = boxoff(B, ng, comp, DIM, gt[0], gt[1], gt[2]);
    if (set_zero) u[ot] = 0.0;
    else u[ot] = sign * u[boxoff(B, ng, comp, DIM, gs[0], gs[1], gs[2])];
}

// ------------------------------- whole-domain (level) API kernels
// natural-layout faces -> folded level weights, rh = macrhs - D(U), phi copy
template <int DIM>
__global__ void __launch_bounds__(TPB) k_level_setup(LevelIO A, Lev<double> L, double *part)
{
    long q = (long)blockIdx.x * TPB + threadIdx.x;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    long n = (long)nx * ny * nz;
    double rh = 0.0;
    if (q < n) {
        int i = (int)(q % nx), j = (int)((q / nx) % ny), k = (int)(q / ((long)nx * ny));
        long fx = i + (long)(nx + 1) * (j + (long)ny * k);
        long fy = i + (long)nx * (j + (long)(ny + 1) * k);
        long fz = q;
        double D = (A.ux[fx + 1] - A.ux[fx]) / A.dx[0] + (A.uy[fy + nx] - A.uy[fy]) / A.dx[1];
        if (DIM == 3) D = D + (A.uz[fz + (long)nx * ny] - A.uz[fz]) / A.dx[2];
        rh = A.macrhs ? -(D - A.macrhs[q]) : -D;
        L.rhs[q] = rh;
        L.phi[q] = A.phi[q];
        L.bx[q] = (i == 0) ? fold_bc(A.bx[fx], A.bc.k[0][0]) : A.bx[fx];
        L.by[q] = (j == 0) ? fold_bc(A.by[fy], A.bc.k[1][0]) : A.by[fy];
        if (i == nx - 1) L.bxh[j + (long)ny * k] = fold_bc(A.bx[fx + 1], A.bc.k[0][1]);
        if (j == ny - 1) L.byh[i + (long)nx * k] = fold_bc(A.by[fy + nx], A.bc.k[1][1]);
        if (DIM == 3) {
            L.bz[q] = (k == 0) ? fold_bc(A.bz[fz], A.bc.k[2][0]) : A.bz[fz];
            if (k == nz - 1) L.bzh[i + (long)nx * j] = fold_bc(A.bz[fz + (long)nx * ny], A.bc.k[2][1]);
        }
    }
    double s = rh, c = 0.0;
    wave_comp_sum(s, c);
    __shared__ double ss[TPB / 64], sc[TPB / 64];
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { ss[wv] = s; sc[wv] = c; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double S = ss[0], C = sc[0];
        for (int w = 1; w < TPB / 64; ++w) comp_add(S, C, ss[w], sc[w]);
        part[2 * blockIdx.x] = S;
        part[2 * blockIdx.x + 1] = C;
    }
}

// gradient of a cell field across every face of direction d, with the
// folded weight w:  returns w and (v_hi - v_lo)/dx (periodic, Neumann w=0,
// Dirichlet neighbour 0)
template <int DIM>
__device__ __forceinline__ void face_grad(const Lev<double> &L, const double *v, int d, int i, int j,
                                          int k, double dxd, int per, double &w, double &g)
{
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int n[3] = {nx, ny, nz};
    int f = (d == 0) ? i : (d == 1 ? j : k);
    int nd = n[d];
    const long s[3] = {1, nx, (long)nx * ny};
    int cc[3] = {i, j, k};
    if (per && f == nd) f = 0;
    cc[d] = f;
    double hi, lo;
    if (f < nd) {
        long cu = cc[0] + s[1] * cc[1] + s[2] * cc[2];
        w = (d == 0 ? L.bx : (d == 1 ? L.by : L.bz))[cu];
        hi = v[cu];
        lo = (f > 0) ? v[cu - s[d]] : (per ? v[cu + (long)(nd - 1) * s[d]] : 0.0);
    } else {
        long cl = cc[0] + s[1] * cc[1] + s[2] * cc[2] - s[d];
        if (d == 0) w = L.bxh[j + (long)ny * k];
        else if (d == 1) w = L.byh[i + (long)nx * k];
        else w = L.bzh[i + (long)nx * j];
        hi = 0.0;
        lo = v[cl];
    }
    g = (hi - lo) / dxd;
}

// U -= w G phi on every face (natural layout)
template <int DIM>
__global__ void __launch_bounds__(TPB) k_level_correct(LevelIO A, Lev<double> L, int d)
{
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int ex = nx + (d == 0), ey = ny + (d == 1), ez = nz + (d == 2);
    long q = (long)blockIdx.x * TPB + threadIdx.x;
    if (q >= (long)ex * ey * ez) return;
    int i = (int)(q % ex), j = (int)((q / ex) % ey), k = (int)(q / ((long)ex * ey));
    double w, g;
    face_grad<DIM>(L, L.phi, d, i, j, k, A.dx[d], A.bc.k[d][0] == ELL_PER, w, g);
    double *u = (d == 0) ? A.ux : (d == 1 ? A.uy : A.uz);
    u[q] = u[q] - w * g;
}

// weight of the domain's high face in the row of cell (i,j,k) along d
__device__ __forceinline__ double whi_row(const Lev<double> &L, int d, int i, int j, int k)
{
    if (d == 0) return L.bxh[j + (long)L.ny * k];
    if (d == 1) return L.byh[i + (long)L.nx * k];
    return L.bzh[i + (long)L.nx * j];
}

// adjoint right side: a = g_phi - G^T(w g_U)  (cells, natural-layout g_U)
template <int DIM>
__global__ void __launch_bounds__(TPB) k_adj_rhs(AdjIO A, Lev<double> L, double *part)
{
    long q = (long)blockIdx.x * TPB + threadIdx.x;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    long n = (long)nx * ny * nz;
    double a = 0.0;
    if (q < n) {
        int i = (int)(q % nx), j = (int)((q / nx) % ny), k = (int)(q / ((long)nx * ny));
        // G^T(w gU) at cell c = sum over faces of dG_f/dphi_c * w_f gU_f
        // face f = lo face of c: G_f = (phi_c - phi_lo)/dx  -> +1/dx
        // face f = hi face of c: G_f = (phi_hi - phi_c)/dx  -> -1/dx
        double acc = 0.0;
        for (int d = 0; d < DIM; ++d) {
            const double *gu = (d == 0) ? A.gux : (d == 1 ? A.guy : A.guz);
            long flo, fhi;
            if (d == 0) { flo = i + (long)(nx + 1) * (j + (long)ny * k); fhi = flo + 1; }
            else if (d == 1) { flo = i + (long)nx * (j + (long)(ny + 1) * k); fhi = flo + nx; }
            else { flo = q; fhi = q + (long)nx * ny; }
            double wlo = (d == 0 ? L.bx : (d == 1 ? L.by : L.bz))[q];
            double whi;
            int ii = (d == 0) ? i : (d == 1 ? j : k), nd = (d == 0) ? nx : (d == 1 ? ny : nz);
            if (ii < nd - 1) whi = (d == 0 ? L.bx : (d == 1 ? L.by : L.bz))[q + (d == 0 ? 1 : (d == 1 ? (long)nx : (long)nx * ny))];
            else if (d == 0) whi = L.bxh[j + (long)ny * k];
            else if (d == 1) whi = L.byh[i + (long)nx * k];
            else whi = L.bzh[i + (long)nx * j];
            double glo = gu[flo], ghi = gu[fhi];
            double t = wlo * glo - whi * ghi;
            // periodic rows: face n carries the same wrapped gradient as face 0,
            // so it also touches cell 0 (as its high side) and face 0 also
            // touches cell n-1 (as its low side)
            if (A.bc.k[d][0] == ELL_PER) {
                long row_f0 = flo - (long)ii * (d == 0 ? 1 : (d == 1 ? (long)nx : (long)nx * ny));
                long row_fn = row_f0 + (long)nd * (d == 0 ? 1 : (d == 1 ? (long)nx : (long)nx * ny));
                long row_c0 = q - (long)ii * (d == 0 ? 1 : (d == 1 ? (long)nx : (long)nx * ny));
                if (ii == 0) t += whi_row(L, d, i, j, k) * gu[row_fn];
                if (ii == nd - 1) t -= (d == 0 ? L.bx : (d == 1 ? L.by : L.bz))[row_c0] * gu[row_f0];
            }
            acc += t / A.dx[d];
        }
        a = A.gphi[q] - acc;
        A.out_rhs[q] = a;
    }
    double s = a, c = 0.0;
    wave_comp_sum(s, c);
    __shared__ double ss[TPB / 64], sc[TPB / 64];
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { ss[wv] = s; sc[wv] = c; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double S = ss[0], C = sc[0];
        for (int w = 1; w < TPB / 64; ++w) comp_add(S, C, ss[w], sc[w]);
        part[2 * blockIdx.x] = S;
        part[2 * blockIdx.x + 1] = C;
    }
}

// per-face gradients:  dL/dU = gU - D^T lambda,  dL/dbeta = -G(phi) (G(lambda) + gU)
// (interior and periodic faces; Neumann faces carry no beta dependence;
// Dirichlet faces use the folded one-sided gradients)
template <int DIM>
__global__ void __launch_bounds__(TPB) k_adj_faces(AdjIO A, Lev<double> L, const double *phi,
                                                   const double *lam, int d)
{
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int ex = nx + (d == 0), ey = ny + (d == 1), ez = nz + (d == 2);
    long q = (long)blockIdx.x * TPB + threadIdx.x;
    if (q >= (long)ex * ey * ez) return;
    int i = (int)(q % ex), j = (int)((q / ex) % ey), k = (int)(q / ((long)ex * ey));
    int per = A.bc.k[d][0] == ELL_PER;
    double w, gp, gl;
    face_grad<DIM>(L, phi, d, i, j, k, A.dx[d], per, w, gp);
    face_grad<DIM>(L, lam, d, i, j, k, A.dx[d], per, w, gl);
    const double *gu = (d == 0) ? A.gux : (d == 1 ? A.guy : A.guz);
    double *ou = (d == 0) ? A.out_ux : (d == 1 ? A.out_uy : A.out_uz);
    double *ob = (d == 0) ? A.out_bx : (d == 1 ? A.out_by : A.out_bz);
    // D^T lambda at face f = (lambda_lo - lambda_hi)/dx over the cells that
    // exist (rh = macrhs - D U, so dL/dU through rh is -D^T lambda)
    int n[3] = {nx, ny, nz};
    int f = (d == 0) ? i : (d == 1 ? j : k);
    const long s[3] = {1, nx, (long)nx * ny};
    int cc[3] = {i, j, k};
    double lhi = 0.0, llo = 0.0;
    if (f < n[d]) { cc[d] = f; lhi = lam[cc[0] + s[1] * cc[1] + s[2] * cc[2]]; }
    if (f > 0) { cc[d] = f - 1; llo = lam[cc[0] + s[1] * cc[1] + s[2] * cc[2]]; }
    double dt = (llo - lhi) / A.dx[d];
    ou[q] = gu[q] - dt;
    // dL/dw_f = G(phi)_f ((D^T lambda)_f - gU_f); on interior and Dirichlet
    // faces D^T lambda = -G(lambda), on the duplicated periodic faces it is
    // not.  The folded Dirichlet weight is 2 beta (chain factor 2); Neumann
    // faces do not depend on beta.
    (void)gl;
    int kind = (f == 0) ? A.bc.k[d][0] : ((f == n[d]) ? A.bc.k[d][1] : ELL_PER);
    double fac = (kind == ELL_NEU) ? 0.0 : (kind == ELL_DIR ? 2.0 : 1.0);
    ob[q] = fac * gp * (dt - gu[q]);
}

// --------------------------------------------------------- launchers
static inline unsigned nblk(long n) { return (unsigned)((n + TPB - 1) / TPB); }

#define DISPATCH_DIM(dim, ...)            \
    do {                                  \
        if ((dim) == 3) {                 \
            constexpr int D_ = 3;         \
            __VA_ARGS__;                  \
        } else {                          \
            constexpr int D_ = 2;         \
            __VA_ARGS__;                  \
        }                                 \
    } while (0)

static unsigned zchunks(int dim, int maxez) { return dim == 3 ? (unsigned)((maxez + ZCH - 1) / ZCH) : 1u; }

void launch_setup(const SetupArgs &A, int maxcols, int maxez, hipStream_t s)
{
    dim3 g(nblk(maxcols), A.nboxes, zchunks(A.dim, maxez));
    DISPATCH_DIM(A.dim, hipLaunchKernelGGL(k_setup<D_>, g, dim3(TPB), 0, s, A));
}

long setup_nparts(int dim, int maxcols, int maxez, int nboxes)
{
    return (long)nblk(maxcols) * nboxes * zchunks(dim, maxez);
}

void launch_mean(const double *part, long nparts, double ncell, double *mean, hipStream_t s)
{
    hipLaunchKernelGGL(k_mean_finalize, dim3(1), dim3(1024), 0, s, part, nparts, ncell, mean);
}

template <typename T>
void launch_sub_mean(T *x, long n, const double *mean, hipStream_t s)
{
    hipLaunchKernelGGL(k_sub_mean<T>, dim3(nblk(n)), dim3(TPB), 0, s, x, n, mean);
}

template <typename T>
long launch_partial_sum(const T *x, long n, double *part, hipStream_t s)
{
    hipLaunchKernelGGL(k_partial_sum<T>, dim3(nblk(n)), dim3(TPB), 0, s, x, n, part);
    return nblk(n);
}

template <typename T>
void launch_coarsen(int dim, const Lev<T> &F, const Lev<T> &C, hipStream_t s)
{
    long n = (long)C.nx * C.ny * C.nz;
    DISPATCH_DIM(dim, hipLaunchKernelGGL((k_coarsen<D_, T>), dim3(nblk(n)), dim3(TPB), 0, s, F, C));
}

void launch_weights_to_f32(int dim, const Lev<double> &S, const Lev<float> &D, hipStream_t s)
{
    long n = (long)S.nx * S.ny * S.nz;
    DISPATCH_DIM(dim, hipLaunchKernelGGL(k_weights_to_f32<D_>, dim3(nblk(n)), dim3(TPB), 0, s, S, D));
}

template <typename A, typename B>
void launch_convert(const A *in, B *out, long n, hipStream_t s)
{
    hipLaunchKernelGGL((k_convert<A, B>), dim3(nblk(n)), dim3(TPB), 0, s, in, out, n);
}
template void launch_convert<double, float>(const double *, float *, long, hipStream_t);
template void launch_convert<float, double>(const float *, double *, long, hipStream_t);

template <typename T>
void launch_gsrb(int dim, const Lev<T> &L, int color, hipStream_t s)
{
    long nq = (long)((L.nx + 1) >> 1) * L.ny * L.nz;
    if (L.omega != T(1))
        DISPATCH_DIM(dim, hipLaunchKernelGGL((k_gsrb<D_, T, true>), dim3(nblk(nq)), dim3(TPB), 0, s, L, color));
    else
        DISPATCH_DIM(dim, hipLaunchKernelGGL((k_gsrb<D_, T>), dim3(nblk(nq)), dim3(TPB), 0, s, L, color));
}

template <typename T, int EY, int PD>
static void launch_deep(const Lev<T> &L, const T *pin, T *pout, int mode, hipStream_t s)
{
    FusedGrid G;
    G.ntx = (L.nx + 62 - 1) / 62;
    G.nty = (L.ny + (EY - 2) - 1) / (EY - 2);
    const int tiles = G.ntx * G.nty;
    int zc = (2048 + tiles - 1) / tiles;
    // small levels are latency bound: their march is kept short (2 planes and
    // more chunks) even though each chunk re-reads its two ring planes
    const long cells = (long)L.nx * L.ny * L.nz;
    const int min_tz = cells <= (1L << 18) ? 2 : 8;
    zc = std::max(1, std::min(zc, std::max(1, L.nz / min_tz)));
    G.tz = (L.nz + zc - 1) / zc;
    zc = (L.nz + G.tz - 1) / G.tz;
    G.total = tiles * zc;
    G.per_xcd = (G.total + 7) / 8;
    dim3 g(8 * G.per_xcd), b(64 * EY);
    auto go = [&](auto AC, auto SC) {
        constexpr bool AL = decltype(AC)::value, SOR = decltype(SC)::value;
        if (mode == SWEEP_ZERO)
            hipLaunchKernelGGL((k_gsrb_deep<T, 1, EY, PD, AL, SOR>), g, b, 0, s, L, pin, pout, G);
        else
            hipLaunchKernelGGL((k_gsrb_deep<T, 0, EY, PD, AL, SOR>), g, b, 0, s, L, pin, pout, G);
    };
    using Y = std::true_type;
    using N = std::false_type;
    const bool sor = L.omega != T(1);
    if (L.alpha) sor ? go(Y{}, Y{}) : go(Y{}, N{});
    else sor ? go(N{}, Y{}) : go(N{}, N{});
}

template <typename T, int EY, int NTH>
static bool rows_fit(const Lev<T> &L)
{
    if (L.nx != 128 && L.nx != 256) return false;   // W = 1, 2 (compile-time row width)
    if (L.nz < 8) return false;                      // the march wraps at most one period
    const int W = L.nx / 128;
    if (64 * W * EY > NTH) return false;
    // byte offsets within a level's allocation (11 cell-sized parts at most) are 32-bit
    if (11.0 * L.nx * L.ny * L.nz * sizeof(T) >= 4294967296.0) return false;
    return sizeof(T) * 4 * (size_t)(EY + 2) * (L.nx + 4) <= 96 * 1024;
}

// offsets of a row sweep's arrays from the lowest address of each group;
// false if one does not fit 32 bits
template <typename T>
static bool row_bufs(const Lev<T> &L, const T *pin, T *pout, const Compact16 *K, RowBufs &B)
{
    const long nc = (long)L.nx * L.ny * L.nz;
    auto group = [](const char *const *p, const long *bytes, int n, const char *&base, unsigned *off) {
        base = p[0];
        for (int a = 1; a < n; ++a) base = std::min(base, p[a]);
        for (int a = 0; a < n; ++a) {
            const long o = p[a] - base;
            if (o < 0 || (double)o + (double)bytes[a] >= 4294967296.0) return false;
            off[a] = (unsigned)o;
        }
        return true;
    };
    const char *pp[2] = {(const char *)pin, (const char *)pout};
    const long pb[2] = {nc * (long)sizeof(T), nc * (long)sizeof(T)};
    unsigned po[2];
    if (!group(pp, pb, 2, B.bp, po)) return false;
    B.in = po[0];
    B.out = po[1];
    const char *wp[8];
    long es;
    if (K) {
        const unsigned short *k[7] = {K->bx, K->by, K->bz, K->rhs, K->bxh, K->byh, K->bzh};
        for (int a = 0; a < 7; ++a) wp[a] = (const char *)k[a];
        es = 2;
    } else {
        const T *w[7] = {L.bx, L.by, L.bz, L.rhs, L.bxh, L.byh, L.bzh};
        for (int a = 0; a < 7; ++a) wp[a] = (const char *)w[a];
        es = sizeof(T);
    }
    const long wb[8] = {nc * es, nc * es, nc * es, nc * es, (long)L.ny * L.nz * es, (long)L.nx * L.nz * es,
                        (long)L.nx * L.ny * es, nc * es};
    // alpha joins the weights' group (it lies in the same allocation)
    const int nw = (L.alpha && !K) ? 8 : 7;
    if (nw == 8) wp[7] = (const char *)L.alpha;
    unsigned wo[8];
    if (!group(wp, wb, nw, B.bw, wo)) return false;
    B.bx = wo[0]; B.by = wo[1]; B.bz = wo[2]; B.rhs = wo[3];
    B.bxh = wo[4]; B.byh = wo[5]; B.bzh = wo[6];
    B.alpha = nw == 8 ? wo[7] : 0u;
    return true;
}

// Tiling of a full-row sweep: y-tiles of EY - 2 written rows, and the
// z-chunk count.  Time ~ rounds of resident blocks x (march steps + a block's
// start-up, about 8 steps: the ring prologue and the LDS clear); ties go to
// more blocks (better balance across CUs).  Measured at 256^3 FP64: 4 chunks
// (one round, 72 steps) 173 us against 12 chunks (three rounds of 24 steps)
// 188 us, which a model without the start-up term preferred.
static RowSweepGrid row_grid(int ny, int nz, int EY, int W, int slots)
{
    RowSweepGrid G;
    G.W = W;
    G.nty = (ny + (EY - 2) - 1) / (EY - 2);
    int best_zc = 1;
    long best = -1, best_blocks = 0;
    for (int zc = 1; zc <= std::max(1, nz / 4); ++zc) {
        const int tz = (nz + zc - 1) / zc;
        const int nzc = (nz + tz - 1) / tz;
        const long blocks = (long)G.nty * nzc;
        const long rounds = (blocks + slots - 1) / slots;
        const int steps = ((tz + 2) + 7) & ~7;   // the march runs in groups of 8 planes
        const long cost = rounds * (steps + 8);
        if (best < 0 || cost < best || (cost == best && blocks > best_blocks)) {
            best = cost;
            best_zc = zc;
            best_blocks = blocks;
        }
    }
    G.tz = (nz + best_zc - 1) / best_zc;
    const int nzc = (nz + G.tz - 1) / G.tz;
    G.total = G.nty * nzc;
    G.per_xcd = (G.total + 7) / 8;
    return G;
}

// resident blocks of a kernel on the device (CUs x blocks per CU), 0 on error
static int resident_slots(const void *kern, int threads, size_t lds)
{
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, threads, lds) != hipSuccess || per < 1) return 0;
    return cus * per;
}

template <typename T, int EY, int NTH, bool B16, int W, bool AL = false, int PDW = 2, bool SOR = false>
static bool launch_rows_w(const Lev<T> &L, const RowBufs &RB, int mode, hipStream_t s)
{
    const size_t lds = sizeof(T) * 4 * (size_t)(EY + 2) * (L.nx + 4);
    const int threads = 64 * W * EY;
    static int slots_cache[2] = {0, 0};   // [mode] resident blocks per device
    int &slots = slots_cache[mode == SWEEP_ZERO ? 1 : 0];
    if (slots == 0) {
        int dev = 0, cus = 0, per = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return false;
        hipError_t e = (mode == SWEEP_ZERO)
                           ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                                 &per, k_gsrb_rows<T, 1, EY, NTH, B16, W, AL, PDW, SOR>, threads, lds)
                           : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                                 &per, k_gsrb_rows<T, 0, EY, NTH, B16, W, AL, PDW, SOR>, threads, lds);
        if (e != hipSuccess || per < 1) return false;
        slots = cus * per;
    }
    const RowSweepGrid G = row_grid(L.ny, L.nz, EY, W, slots);
    dim3 g(8 * G.per_xcd), b(threads);
    if (mode == SWEEP_ZERO)
        hipLaunchKernelGGL((k_gsrb_rows<T, 1, EY, NTH, B16, W, AL, PDW, SOR>), g, b, lds, s, L, RB, G);
    else
        hipLaunchKernelGGL((k_gsrb_rows<T, 0, EY, NTH, B16, W, AL, PDW, SOR>), g, b, lds, s, L, RB, G);
    return true;
}

// Full-row sweep: usable when nx = 128 W (W = 1, 2) and a block of EY rows
// (64 W EY threads) fits in NTH.  The z-chunk count is chosen so the number
// of blocks fills whole rounds of the resident block slots (CUs x blocks per
// CU) with the least z-ring overhead.  PDW: the kernel's prefetch distance of
// the weights (with alpha it stays 2).
template <typename T, int EY, int NTH, bool B16, int PDW = 2>
static bool launch_rows(const Lev<T> &L, const T *pin, T *pout, int mode, hipStream_t s, const Compact16 *K)
{
    RowBufs RB;
    if (!rows_fit<T, EY, NTH>(L) || !row_bufs<T>(L, pin, pout, B16 ? K : nullptr, RB)) return false;
    const bool sor = L.omega != T(1);
    if constexpr (!B16) {
        if (L.alpha) {
            if (L.nx == 128)
                return sor ? launch_rows_w<T, EY, NTH, false, 1, true, 2, true>(L, RB, mode, s)
                           : launch_rows_w<T, EY, NTH, false, 1, true>(L, RB, mode, s);
            return sor ? launch_rows_w<T, EY, NTH, false, 2, true, 2, true>(L, RB, mode, s)
                       : launch_rows_w<T, EY, NTH, false, 2, true>(L, RB, mode, s);
        }
    } else {
        if (L.alpha) return false;
    }
    if (L.nx == 128)
        return sor ? launch_rows_w<T, EY, NTH, B16, 1, false, PDW, true>(L, RB, mode, s)
                   : launch_rows_w<T, EY, NTH, B16, 1, false, PDW>(L, RB, mode, s);
    return sor ? launch_rows_w<T, EY, NTH, B16, 2, false, PDW, true>(L, RB, mode, s)
               : launch_rows_w<T, EY, NTH, B16, 2, false, PDW>(L, RB, mode, s);
}

// The rho form of the FP32 finest-level sweep (k_gsrb_rows_rho): 6 rows of
// 768 threads at nx = 256, 12 rows of 768 at nx = 128; phi and rho planes
// in LDS (8 planes), two blocks per CU.
template <int EY, int NTH, int W, bool SOR, bool PZ>
static bool launch_rows_rho_w(const Lev<float> &L, const RowBufs &RB, const RhoBufs &HB, int mode, hipStream_t s)
{
    const size_t lds = sizeof(float) * 8 * (size_t)(EY + 2) * (L.nx + 4);
    const int threads = 64 * W * EY;
    static int slots_cache[2] = {0, 0};   // [mode] resident blocks per device
    int &slots = slots_cache[mode == SWEEP_ZERO ? 1 : 0];
    if (slots == 0) {
        slots = (mode == SWEEP_ZERO) ? resident_slots((const void *)k_gsrb_rows_rho<1, EY, NTH, W, SOR, PZ>, threads, lds)
                                     : resident_slots((const void *)k_gsrb_rows_rho<0, EY, NTH, W, SOR, PZ>, threads, lds);
        if (slots == 0) return false;
    }
    const RowSweepGrid G = row_grid(L.ny, L.nz, EY, W, slots);
    dim3 g(8 * G.per_xcd), b(threads);
    if (mode == SWEEP_ZERO)
        hipLaunchKernelGGL((k_gsrb_rows_rho<1, EY, NTH, W, SOR, PZ>), g, b, lds, s, L, RB, HB, G);
    else
        hipLaunchKernelGGL((k_gsrb_rows_rho<0, EY, NTH, W, SOR, PZ>), g, b, lds, s, L, RB, HB, G);
    return true;
}

bool rho_level_ok(int dim, const Lev<float> &L)
{
    if (dim != 3 || !L.perx || !L.pery || L.alpha) return false;
    if (L.nx == 256) return rows_fit<float, 6, 768>(L);
    if (L.nx == 128) return rows_fit<float, 12, 768>(L);
    return false;
}

// byte offsets of the rho form's arrays in the level's allocation (the
// weights' group of the row sweep); false if one does not fit 32 bits
static bool rho_bufs(const Lev<float> &L, const RhoLev &R, const RowBufs &RB, RhoBufs &HB)
{
    const float *p[5] = {R.rho, R.cxy, R.cz, R.bzl, L.bzh};
    unsigned o[5];
    for (int a = 0; a < 5; ++a) {
        const long d = (const char *)p[a] - RB.bw;
        if (!p[a] || d < 0 || (double)d + 4.0 * L.nx * L.ny * L.nz >= 4294967296.0) return false;
        o[a] = (unsigned)d;
    }
    HB.rho = o[0]; HB.cxy = o[1]; HB.cz = o[2]; HB.bzl = o[3]; HB.bzh = o[4];
    return true;
}

bool launch_gsrb_rho(const Lev<float> &L, const RhoLev &R, const float *pin, float *pout, int mode, hipStream_t s)
{
    RowBufs RB;
    RhoBufs HB;
    if (!rho_level_ok(3, L) || !row_bufs<float>(L, pin, pout, nullptr, RB) || !rho_bufs(L, R, RB, HB)) return false;
    const bool sor = L.omega != 1.0f;
    auto go = [&](auto SC, auto PC) {
        constexpr bool SR = decltype(SC)::value, PZ = decltype(PC)::value;
        if (L.nx == 256) return launch_rows_rho_w<6, 768, 2, SR, PZ>(L, RB, HB, mode, s);
        return launch_rows_rho_w<12, 768, 1, SR, PZ>(L, RB, HB, mode, s);
    };
    using T_ = std::true_type;
    using F_ = std::false_type;
    if (sor) return L.perz ? go(T_{}, T_{}) : go(T_{}, F_{});
    return L.perz ? go(F_{}, T_{}) : go(F_{}, F_{});
}

bool compact_level_ok(int dim, const Lev<float> &L)
{
    return dim == 3 && rows_fit<float, 6, 768>(L);
}

// One fused red-black sweep of a 3D level (pin -> pout); false where no fused
// form applies (the caller then sweeps per colour).
//   - nx = 128 or 256: the full-row sweep.  FP32 at nx = 256 and the compact
//     (bfloat16) form run 6 rows in blocks of 768 threads with 80 VGPRs (the
//     FP32 form with the weights fetched one plane ahead), two blocks per CU:
//     a 256^3 level gives 64 y-tiles x 8 z-chunks, 512 blocks on 256 CUs.
//     8 rows of 1024 threads (88 VGPRs, one block per CU) leave 43 x 5 = 215
//     blocks and idle CUs: 111 us against 106.5 us per FP32 sweep.  FP32 at
//     nx = 128 runs 16 rows of 1024 threads, and 8 rows of 1024 with alpha.
//     FP64 runs 6 rows of 768 threads, and 4 rows of 512 with alpha (the
//     alpha ring does not fit the 168 VGPRs a 768-thread block leaves a lane).
//   - other sizes: one column per thread (k_gsrb_deep, two planes of prefetch).
template <typename T>
bool launch_gsrb_fused(const Lev<T> &L, const T *pin, T *pout, int mode, hipStream_t s, const Compact16 *K)
{
    if constexpr (sizeof(T) == 4) {
        if (K) {
            if (L.alpha) return false;   // compact storage carries no alpha
            return launch_rows<T, 6, 768, true>(L, pin, pout, mode, s, K);
        }
    } else {
        if (K) return false;
    }
    bool done;
    if constexpr (sizeof(T) == 8)
        done = L.alpha ? launch_rows<T, 4, 512, false>(L, pin, pout, mode, s, nullptr)
                       : launch_rows<T, 6, 768, false>(L, pin, pout, mode, s, nullptr);
    else if (L.nx == 128)
        done = launch_rows<T, 16, 1024, false>(L, pin, pout, mode, s, nullptr);
    else if (L.alpha)
        done = launch_rows<T, 8, 1024, false>(L, pin, pout, mode, s, nullptr);
    else
        done = launch_rows<T, 6, 768, false, 1>(L, pin, pout, mode, s, nullptr);
    if (!done) launch_deep<T, 8, 2>(L, pin, pout, mode, s);
    return true;
}

template <typename T>
bool launch_resid_restrict(int dim, const Lev<T> &F, const T *phi, const Lev<T> &C, hipStream_t s,
                           const Compact16 *K, const RhoLev *Rh)
{
    if (F.alpha && K) return false;
    if (Rh && (K || sizeof(T) != 4 || !rho_level_ok(dim, *(const Lev<float> *)&F))) return false;
    if (dim == 3 && F.nx % 128 == 0 && C.nz >= 1) {
        RowMarchGrid G;
        G.W = F.nx / 128;
        // coarse planes per chunk: at most 16, fewer on small levels so that
        // there are some 4096 waves (16 per CU) to hide the march's latency
        const long segs = (long)C.ny * G.W;
        G.tz = (int)std::max(1L, std::min(16L, segs * C.nz / 4096));
        G.tz = std::min(G.tz, C.nz);
        const int chunks = (C.nz + G.tz - 1) / G.tz;
        G.total = C.ny * G.W * chunks;
        const int blocks = (G.total + 3) / 4;
        G.per_xcd = (blocks + 7) / 8;
        if (Rh) {
            if constexpr (sizeof(T) == 4)
                hipLaunchKernelGGL(k_rr_rows_rho, dim3(8 * G.per_xcd), dim3(256), 0, s, F, phi, C, G, *Rh);
            else
                return false;
        } else if (K) {
            if constexpr (sizeof(T) == 4)
                hipLaunchKernelGGL((k_rr_rows<T, true>), dim3(8 * G.per_xcd), dim3(256), 0, s, F, phi, C, G, *K);
            else
                return false;
        } else if (F.alpha) {
            const Compact16 K0 = {};
            hipLaunchKernelGGL((k_rr_rows<T, false, true>), dim3(8 * G.per_xcd), dim3(256), 0, s, F, phi, C, G,
                               K0);
        } else {
            const Compact16 K0 = {};
            hipLaunchKernelGGL((k_rr_rows<T, false>), dim3(8 * G.per_xcd), dim3(256), 0, s, F, phi, C, G, K0);
        }
        return true;
    }
    if (K) return false;
    dim3 g((C.nx + 63) / 64, C.ny, C.nz);
    if (dim == 3)
        hipLaunchKernelGGL((k_resid_restrict<3, T>), g, dim3(256), 0, s, F, phi, C);
    else
        hipLaunchKernelGGL((k_resid_restrict<2, T>), g, dim3(128), 0, s, F, phi, C);
    return true;
}

template bool launch_gsrb_fused<double>(const Lev<double> &, const double *, double *, int, hipStream_t,
                                        const Compact16 *);
template bool launch_gsrb_fused<float>(const Lev<float> &, const float *, float *, int, hipStream_t,
                                       const Compact16 *);
template bool launch_resid_restrict<double>(int, const Lev<double> &, const double *,
                                            const Lev<double> &, hipStream_t, const Compact16 *, const RhoLev *);
template bool launch_resid_restrict<float>(int, const Lev<float> &, const float *,
                                           const Lev<float> &, hipStream_t, const Compact16 *, const RhoLev *);

void launch_compact_level0(const Lev<double> &F, const Compact16 &K, const Lev<float> &C, hipStream_t s)
{
    const long n = (long)C.nx * C.ny * C.nz;
    hipLaunchKernelGGL(k_compact_level0, dim3(nblk(n)), dim3(TPB), 0, s, F, K, C);
}

void launch_rho_prep(const Lev<double> &F, const RhoLev &R, const Lev<float> &F32, const Lev<float> &C,
                     const double *b0, const double *b0e, int alt, hipStream_t s)
{
    hipLaunchKernelGGL(k_rho_coef, dim3(nblk(F.nz + 1)), dim3(TPB), 0, s, b0, b0e, alt, F.nz, R);
    const long n = (long)C.nx * C.ny * C.nz;
    hipLaunchKernelGGL(k_rho_level0, dim3(nblk(n)), dim3(TPB), 0, s, F, R, F32.bzh, C);
}

void launch_to_bf16(const float *in, unsigned short *out, long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_to_bf16, dim3(nblk(n)), dim3(TPB), 0, s, in, out, n);
}

static constexpr unsigned NORM_GRID = 2048;   // 256 CUs x 8 blocks

// Row-march form of the FP64 residual passes where it applies (3D, nx a
// multiple of 128); false otherwise (the row-per-wave kernels then run).
// The z-chunk length is 32 planes, fewer on small levels.
template <int MODE>
static bool launch_resid_rows(const Lev<double> &L, const float *e, double *phi_out, float *r32,
                              unsigned long long *nrm, hipStream_t s, unsigned short *r16 = nullptr)
{
    if (L.nx % 128 != 0 || L.nz < 2) return false;
    RowMarchGrid G;
    G.W = L.nx / 128;
    // planes per chunk: 32, fewer on small levels (some 4096 waves in all)
    const long segs = (long)L.ny * G.W;
    G.tz = (int)std::max(2L, std::min(32L, segs * L.nz / 4096));
    G.tz = std::min(G.tz, L.nz);
    const int chunks = (L.nz + G.tz - 1) / G.tz;
    G.total = L.ny * G.W * chunks;
    const int blocks = (G.total + 3) / 4;
    G.per_xcd = (blocks + 7) / 8;
    if (L.alpha) {
        if (r16) return false;   // compact storage carries no alpha
        hipLaunchKernelGGL((k_resid_rows<MODE, false, true>), dim3(8 * G.per_xcd), dim3(256), 0, s, L, e,
                           phi_out, r32, r16, nrm, G);
    } else if (r16)
        hipLaunchKernelGGL((k_resid_rows<MODE, true>), dim3(8 * G.per_xcd), dim3(256), 0, s, L, e, phi_out,
                           r32, r16, nrm, G);
    else
        hipLaunchKernelGGL((k_resid_rows<MODE, false>), dim3(8 * G.per_xcd), dim3(256), 0, s, L, e, phi_out,
                           r32, r16, nrm, G);
    return true;
}

template <typename T>
void launch_residual(int dim, const Lev<T> &L, T *out, unsigned long long *nrm, hipStream_t s)
{
    if constexpr (sizeof(T) == 8) {
        if (dim == 3 && out == nullptr && nrm != nullptr &&
            launch_resid_rows<0>(L, nullptr, nullptr, nullptr, nrm, s))
            return;
    }
    const unsigned rb = row_blocks(L.ny, L.nz);
    unsigned g = nrm ? std::min(rb, NORM_GRID) : rb;
    DISPATCH_DIM(dim, hipLaunchKernelGGL((k_residual<D_, T>), dim3(g), dim3(TPB), 0, s, L, out, nrm));
}

void launch_ir_residual(int dim, const Lev<double> &L, float *r32, unsigned long long *nrm,
                        hipStream_t s, unsigned short *r16)
{
    if (dim == 3 && launch_resid_rows<1>(L, nullptr, nullptr, r32, nrm, s, r16)) return;
    unsigned g = std::min(row_blocks(L.ny, L.nz), NORM_GRID);
    DISPATCH_DIM(dim, hipLaunchKernelGGL(k_ir_residual<D_>, dim3(g), dim3(TPB), 0, s, L, r32, nrm));
}

void launch_ir_update_residual(int dim, const Lev<double> &L, const float *e, double *phi_out, float *r32,
                               unsigned long long *nrm, hipStream_t s, unsigned short *r16)
{
    if (dim == 3 && launch_resid_rows<2>(L, e, phi_out, r32, nrm, s, r16)) return;
    unsigned g = std::min(row_blocks(L.ny, L.nz), NORM_GRID);
    DISPATCH_DIM(dim, hipLaunchKernelGGL(k_ir_update_residual<D_>, dim3(g), dim3(TPB), 0, s, L, e, phi_out,
                                         r32, nrm));
}

__global__ void k_publish_norm(unsigned long long *src, unsigned long long *dst, unsigned long long seq,
                               unsigned long long *keep, const int *flag)
{
    if (threadIdx.x == 0) {
        const unsigned long long v = __hip_atomic_load(src, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(keep, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(dst, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        if (flag)
            __hip_atomic_store(dst + 2, (unsigned long long)(unsigned)*flag, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(dst + 1, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(src, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

void launch_publish_norm(unsigned long long *src, unsigned long long *dst, unsigned long long seq,
                         hipStream_t s, unsigned long long *keep, const int *flag)
{
    hipLaunchKernelGGL(k_publish_norm, dim3(1), dim3(64), 0, s, src, dst, seq, keep, flag);
}

void launch_sub_mean_absmax(double *x, long n, const double *mean, unsigned long long *nrm, hipStream_t s)
{
    unsigned g = std::min(nblk(n), NORM_GRID);
    hipLaunchKernelGGL(k_sub_mean_absmax, dim3(g), dim3(TPB), 0, s, x, n, mean, nrm);
}

void launch_absmax(const double *x, long n, unsigned long long *nrm, hipStream_t s)
{
    unsigned g = std::min(nblk(n), NORM_GRID);
    hipLaunchKernelGGL(k_absmax, dim3(g), dim3(TPB), 0, s, x, n, nrm);
}

template <typename T>
void launch_restrict(int dim, const T *r, int fnx, int fny, const Lev<T> &C, hipStream_t s)
{
    long n = (long)C.nx * C.ny * C.nz;
    DISPATCH_DIM(dim, hipLaunchKernelGGL((k_restrict<D_, T>), dim3(nblk(n)), dim3(TPB), 0, s, r, fnx, fny, C));
}

template <typename T>
void launch_prolong(int dim, const Lev<T> &F, const Lev<T> &C, const EllBC &bc, hipStream_t s)
{
    if (dim == 3 && F.nx % 128 == 0 && F.ny == 2 * C.ny && F.nz == 2 * C.nz) {
        // fine planes per wave: up to 16 while some 8192 waves remain
        const int W = F.nx / 128;
        const long rows = (long)W * F.ny;
        const int pz = (int)std::max(2L, std::min(16L, (rows * F.nz / 8192) & ~1L));
        const int nch = (F.nz + pz - 1) / pz;
        const int total = (int)(rows * nch);
        hipLaunchKernelGGL((k_prolong_rows<T>), dim3((total + 3) / 4), dim3(256), 0, s, F, (const T *)C.phi,
                           C.nx, C.ny, C.nz, bc, W, pz, total);
        return;
    }
    DISPATCH_DIM(dim, hipLaunchKernelGGL((k_prolong<D_, T>), dim3(row_blocks(F.ny, F.nz)), dim3(TPB), 0, s,
                                         F, (const T *)C.phi, C.nx, C.ny, C.nz, bc));
}

template <typename T>
void launch_fill(T *x, long n, T v, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill<T>, dim3(nblk(n)), dim3(TPB), 0, s, x, n, v);
}

// LDS bytes the coarse cycle needs to keep levels lc..nlev-1 and the bottom
// CG vectors on chip
template <typename T>
static long coarse_lds_bytes(int dim, const Lev<T> *levs_host, int lc, int nlev)
{
    long b = 0;
    for (int l = lc; l < nlev; ++l) {
        const Lev<T> &L = levs_host[l];
        b += dim == 3 ? lds_level_bytes<3, T>(L.nx, L.ny, L.nz) : lds_level_bytes<2, T>(L.nx, L.ny, 1);
    }
    const Lev<T> &B = levs_host[nlev - 1];
    return b + (long)BOTTOM_VECS * B.nx * B.ny * B.nz * (long)sizeof(double);
}

static constexpr long COARSE_LDS_MAX = 64 * 1024;

template <typename T>
void launch_coarse_cycle(int dim, const Lev<T> *levs_dev, const Lev<T> *levs_host, int lc, int nlev,
                         const CoarseOpts &o, double *cg, hipStream_t s)
{
    const long lds = coarse_lds_bytes<T>(dim, levs_host, lc, nlev);
    const bool in_lds = lds <= COARSE_LDS_MAX;
    const unsigned sh = in_lds ? (unsigned)lds : 0u;
    if (o.cycle == MAC_CYCLE_V) {
        if (in_lds)
            DISPATCH_DIM(dim, hipLaunchKernelGGL((k_coarse_cycle<D_, T, true, true>), dim3(1), dim3(CTPB), sh, s,
                                                 levs_dev, lc, nlev, o, cg));
        else
            DISPATCH_DIM(dim, hipLaunchKernelGGL((k_coarse_cycle<D_, T, false, true>), dim3(1), dim3(CTPB), sh,
                                                 s, levs_dev, lc, nlev, o, cg));
    } else {
        if (in_lds)
            DISPATCH_DIM(dim, hipLaunchKernelGGL((k_coarse_cycle<D_, T, true, false>), dim3(1), dim3(CTPB), sh,
                                                 s, levs_dev, lc, nlev, o, cg));
        else
            DISPATCH_DIM(dim, hipLaunchKernelGGL((k_coarse_cycle<D_, T, false, false>), dim3(1), dim3(CTPB), sh,
                                                 s, levs_dev, lc, nlev, o, cg));
    }
}

void launch_correct(const CorrectArgs &A, int maxcols, int maxez, int maxghost, hipStream_t s)
{
    DISPATCH_DIM(A.dim, hipLaunchKernelGGL(k_correct_valid<D_>,
                                           dim3(nblk(maxcols), A.nboxes, zchunks(A.dim, maxez)), dim3(TPB), 0,
                                           s, A));
    launch_fill_ghosts(A, maxghost, true, true, s);
}

void launch_fill_ghosts(const CorrectArgs &A, int maxghost, bool fill, bool phys, hipStream_t s)
{
    if (A.ng_u <= 0) return;
    if (fill)
        DISPATCH_DIM(A.dim, hipLaunchKernelGGL(k_fill_ghost<D_>, dim3(nblk(maxghost), A.nboxes, A.dim),
                                               dim3(TPB), 0, s, A));
    if (!phys) return;
    for (int d = 0; d < A.dim; ++d)
        for (int side = 0; side < 2; ++side) {
            int pc = A.phys[d][side];
            if (pc == MAC_PERIODIC || pc == MAC_INTERIOR) continue;
            DISPATCH_DIM(A.dim, hipLaunchKernelGGL(k_physbc<D_>, dim3(nblk(A.maxplane), A.nboxes, A.dim),
                                                   dim3(TPB), 0, s, A, d, side, pc));
        }
}

void launch_level_setup(int dim, const LevelIO &A, const Lev<double> &L, double *part, hipStream_t s)
{
    long n = (long)L.nx * L.ny * L.nz;
    DISPATCH_DIM(dim, hipLaunchKernelGGL(k_level_setup<D_>, dim3(nblk(n)), dim3(TPB), 0, s, A, L, part));
}

void launch_level_correct(int dim, const LevelIO &A, const Lev<double> &L, hipStream_t s)
{
    for (int d = 0; d < dim; ++d) {
        long n = (long)(L.nx + (d == 0)) * (L.ny + (d == 1)) * (L.nz + (d == 2));
        DISPATCH_DIM(dim, hipLaunchKernelGGL(k_level_correct<D_>, dim3(nblk(n)), dim3(TPB), 0, s, A, L, d));
    }
}

void launch_adj_rhs(int dim, const AdjIO &A, const Lev<double> &L, double *part, hipStream_t s)
{
    long n = (long)L.nx * L.ny * L.nz;
    DISPATCH_DIM(dim, hipLaunchKernelGGL(k_adj_rhs<D_>, dim3(nblk(n)), dim3(TPB), 0, s, A, L, part));
}

void launch_adj_faces(int dim, const AdjIO &A, const Lev<double> &L, const double *phi,
                      const double *lam, hipStream_t s)
{
    for (int d = 0; d < dim; ++d) {
        long n = (long)(L.nx + (d == 0)) * (L.ny + (d == 1)) * (L.nz + (d == 2));
        DISPATCH_DIM(dim, hipLaunchKernelGGL(k_adj_faces<D_>, dim3(nblk(n)), dim3(TPB), 0, s, A, L, phi,
                                             lam, d));
    }
}

// explicit instantiations
template void launch_sub_mean<double>(double *, long, const double *, hipStream_t);
template long launch_partial_sum<double>(const double *, long, double *, hipStream_t);
template void launch_coarsen<double>(int, const Lev<double> &, const Lev<double> &, hipStream_t);
template void launch_coarsen<float>(int, const Lev<float> &, const Lev<float> &, hipStream_t);
template void launch_gsrb<double>(int, const Lev<double> &, int, hipStream_t);
template void launch_gsrb<float>(int, const Lev<float> &, int, hipStream_t);
template void launch_residual<double>(int, const Lev<double> &, double *, unsigned long long *, hipStream_t);
template void launch_residual<float>(int, const Lev<float> &, float *, unsigned long long *, hipStream_t);
template void launch_restrict<double>(int, const double *, int, int, const Lev<double> &, hipStream_t);
template void launch_restrict<float>(int, const float *, int, int, const Lev<float> &, hipStream_t);
template void launch_prolong<double>(int, const Lev<double> &, const Lev<double> &, const EllBC &, hipStream_t);
template void launch_prolong<float>(int, const Lev<float> &, const Lev<float> &, const EllBC &, hipStream_t);
template void launch_fill<double>(double *, long, double, hipStream_t);
template void launch_fill<float>(float *, long, float, hipStream_t);
template void launch_coarse_cycle<double>(int, const Lev<double> *, const Lev<double> *, int, int, const CoarseOpts &, double *, hipStream_t);
template void launch_coarse_cycle<float>(int, const Lev<float> *, const Lev<float> *, int, int, const CoarseOpts &, double *, hipStream_t);

}  // namespace mac
