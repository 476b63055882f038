// mac_internal.h -- device data structures and stencil helpers shared by the
// gfx950 kernels of libmaestro_mac.so.
//
// Level storage (one buffer per quantity per MG level, C order, x fastest,
// no ghost cells):
//   cells            n = nx*ny*nz (nz = 1 in 2D)
//   face weights     bx[c], by[c], bz[c] hold the LOW face of cell c in x/y/z;
//                    the domain's high faces live in slabs
//                    bxh[j + ny*k], byh[i + nx*k], bzh[i + nx*j].
// Rows of nx values start on 256-byte boundaries for power-of-two nx, which
// keeps every wavefront load of a row coalesced.
//
// Boundary conditions are folded into the stored face weights so that the
// 7-point operator never branches on BC type:
//   periodic face   weight beta, neighbour = periodic image
//   Neumann face    weight 0                (no flux; press_comp walls,
//                                            Source/define_bc_tower.f90:338)
//   Dirichlet face  weight 2*beta, neighbour value 0   (phi = 0 on the face,
//                                            OUTLET, define_bc_tower.f90:371)
// so (A phi)_c = sum_d ih2_d sum_s w_s (phi_c - p_s).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mac {

enum { ELL_PER = -1, ELL_DIR = 1, ELL_NEU = 2 };

template <typename T>
struct Lev {
    int nx, ny, nz;
    int perx, pery, perz;
    T ih2x, ih2y, ih2z;      // 1/h^2 per direction (ih2z = 0 in 2D)
    T *phi, *phi2, *rhs, *res;
    T *bx, *by, *bz;
    T *bxh, *byh, *bzh;
    // cell coefficient of the Helmholtz operator (alpha - D beta G); null for
    // the projection (alpha = 0).  Only the generic kernels (per-colour
    // sweep, residual, coarsening, coarse cycle) read it.
    T *alpha;
    // over-relaxation of this level's red-black sweeps: a cell moves to
    // old + omega (gs - old); 1 (set when the level is built) keeps
    // Gauss-Seidel, and only then do the sweeps store the Gauss-Seidel value
    // itself (the single-workgroup coarse cycle never reads it)
    T omega;
};

struct BoxDev {
    int lo[3];
    int hi[3];
};

// Level 0 of the FP32 hierarchy in compact storage: the face weights (same
// layout as Lev::bx.. including the high-face slabs) and the right side as
// bfloat16 bit patterns.  The iterate stays FP32 and all arithmetic is FP32;
// a bfloat16 value widens to FP32 exactly (its bits in the high half).
struct Compact16 {
    unsigned short *bx, *by, *bz;
    unsigned short *bxh, *byh, *bzh;
    unsigned short *rhs;
};

// FP32 -> bfloat16, round to nearest even (finite inputs)
__host__ __device__ __forceinline__ unsigned short f32_to_bf16(float f)
{
    const unsigned u = __builtin_bit_cast(unsigned, f);
    return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
// the two bfloat16 values packed in a 32-bit word (low half first)
__device__ __forceinline__ float bf16_lo(unsigned u) { return __uint_as_float(u << 16); }
__device__ __forceinline__ float bf16_hi(unsigned u) { return __uint_as_float(u & 0xffff0000u); }
__device__ __forceinline__ unsigned bf16_pack(float lo, float hi)
{
    return (unsigned)f32_to_bf16(lo) | ((unsigned)f32_to_bf16(hi) << 16);
}

// ------------------------------------------------------------ math helpers
__device__ __forceinline__ double dfma(double a, double b, double c) { return __fma_rn(a, b, c); }
__device__ __forceinline__ float dfma(float a, float b, float c) { return __fmaf_rn(a, b, c); }
// the new value of a cell from its Gauss-Seidel value gs and its old value:
// gs itself, or over-relaxed old + omega (gs - old) as one fma (red-black SOR
// as a multigrid smoother: I. Yavneh, SIAM J. Sci. Comput. 17 (1996) 180-192)
template <bool SOR, typename T>
__device__ __forceinline__ T relax(T gs, T old, T omega)
{
    if constexpr (SOR) return dfma(omega, gs - old, old);
    else return gs;
}

// Neighbour values and face weights of cell (i,j,k), index c.
// Order: x-, x+, y-, y+, z-, z+.
template <int DIM, typename T>
__device__ __forceinline__ void stencil(const Lev<T> &L, const T *__restrict__ phi, int i, int j,
                                        int k, long c, T w[6], T p[6])
{
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const long sx = 1, sy = nx, sz = (long)nx * ny;
    w[0] = L.bx[c];
    w[1] = (i < nx - 1) ? L.bx[c + sx] : L.bxh[j + (long)ny * k];
    w[2] = L.by[c];
    w[3] = (j < ny - 1) ? L.by[c + sy] : L.byh[i + (long)nx * k];
    p[0] = (i > 0) ? phi[c - sx] : (L.perx ? phi[c + (nx - 1) * sx] : T(0));
    p[1] = (i < nx - 1) ? phi[c + sx] : (L.perx ? phi[c - (nx - 1) * sx] : T(0));
    p[2] = (j > 0) ? phi[c - sy] : (L.pery ? phi[c + (long)(ny - 1) * sy] : T(0));
    p[3] = (j < ny - 1) ? phi[c + sy] : (L.pery ? phi[c - (long)(ny - 1) * sy] : T(0));
    if (DIM == 3) {
        w[4] = L.bz[c];
        w[5] = (k < nz - 1) ? L.bz[c + sz] : L.bzh[i + (long)nx * j];
        p[4] = (k > 0) ? phi[c - sz] : (L.perz ? phi[c + (long)(nz - 1) * sz] : T(0));
        p[5] = (k < nz - 1) ? phi[c + sz] : (L.perz ? phi[c - (long)(nz - 1) * sz] : T(0));
    } else {
        w[4] = w[5] = p[4] = p[5] = T(0);
    }
}

// The six neighbour values of cell (i,j,k) of an array x laid out like
// level L, with the same periodic wrap and zero outside as stencil()
template <int DIM, typename S, typename T>
__device__ __forceinline__ void nbr_values(const Lev<S> &L, const T *__restrict__ x, int i, int j, int k,
                                           long c, T p[6])
{
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const long sy = nx, sz = (long)nx * ny;
    p[0] = (i > 0) ? x[c - 1] : (L.perx ? x[c + (nx - 1)] : T(0));
    p[1] = (i < nx - 1) ? x[c + 1] : (L.perx ? x[c - (nx - 1)] : T(0));
    p[2] = (j > 0) ? x[c - sy] : (L.pery ? x[c + (long)(ny - 1) * sy] : T(0));
    p[3] = (j < ny - 1) ? x[c + sy] : (L.pery ? x[c - (long)(ny - 1) * sy] : T(0));
    if (DIM == 3) {
        p[4] = (k > 0) ? x[c - sz] : (L.perz ? x[c + (long)(nz - 1) * sz] : T(0));
        p[5] = (k < nz - 1) ? x[c + sz] : (L.perz ? x[c - (long)(nz - 1) * sz] : T(0));
    } else {
        p[4] = p[5] = T(0);
    }
}

// Gauss-Seidel point value (fixed evaluation order, shared with the oracle);
// a = the cell's alpha, added to the diagonal last.  The default -0.0 adds nothing (x + -0.0 == x for every x), so the
// compiler drops it from the kernels without alpha.
template <int DIM, typename T>
__device__ __forceinline__ T gs_value(const Lev<T> &L, const T w[6], const T p[6], T rhs, T a = T(-0.0))
{
    T tx = dfma(w[0], p[0], w[1] * p[1]), sx = w[0] + w[1];
    T ty = dfma(w[2], p[2], w[3] * p[3]), sy = w[2] + w[3];
    T num, den;
    if (DIM == 3) {
        T tz = dfma(w[4], p[4], w[5] * p[5]), sz = w[4] + w[5];
        num = dfma(L.ih2x, tx, dfma(L.ih2y, ty, dfma(L.ih2z, tz, rhs)));
        den = dfma(L.ih2x, sx, dfma(L.ih2y, sy, L.ih2z * sz));
    } else {
        num = dfma(L.ih2x, tx, dfma(L.ih2y, ty, rhs));
        den = dfma(L.ih2x, sx, L.ih2y * sy);
    }
    return num / (den + a);
}
// alpha of cell c, or -0.0 (no contribution) on a level without alpha
template <typename T>
__device__ __forceinline__ T alpha_at(const Lev<T> &L, long c)
{
    return L.alpha ? L.alpha[c] : T(-0.0);
}
// A phi at cell c from its stencil part ap: + alpha phi as one fma
template <typename T>
__device__ __forceinline__ T with_alpha(const Lev<T> &L, long c, T ap, T pc)
{
    return L.alpha ? dfma(L.alpha[c], pc, ap) : ap;
}

// A phi at a cell (fixed evaluation order, shared with the oracle)
template <int DIM, typename T>
__device__ __forceinline__ T apply_value(const Lev<T> &L, const T w[6], const T p[6], T pc)
{
    T ax = dfma(w[0], pc - p[0], w[1] * (pc - p[1]));
    T ay = dfma(w[2], pc - p[2], w[3] * (pc - p[3]));
    if (DIM == 3) {
        T az = dfma(w[4], pc - p[4], w[5] * (pc - p[5]));
        return dfma(L.ih2x, ax, dfma(L.ih2y, ay, L.ih2z * az));
    }
    return dfma(L.ih2x, ax, L.ih2y * ay);
}

// 1-D linear prolongation weights for fine index fi on a coarse line of nc
// cells (parent 3/4, neighbour 1/4; out of the domain: periodic image,
// parent (Neumann) or parent with -1/4 (Dirichlet)).
template <typename T>
__device__ __forceinline__ void lin_w(int fi, int nc, int kind_lo, int kind_hi, int &P, int &Q,
                                      T &wp, T &wq)
{
    int p = fi >> 1;
    int q = (fi & 1) ? p + 1 : p - 1;
    P = p;
    wp = T(0.75);
    wq = T(0.25);
    if (q < 0 || q >= nc) {
        int kind = (q < 0) ? kind_lo : kind_hi;
        if (kind == ELL_PER) q = (q + nc) % nc;
        else if (kind == ELL_DIR) { q = p; wq = T(-0.25); }
        else q = p;
    }
    Q = q;
}

struct EllBC {
    int k[3][2];   // elliptic kind per direction and side
};

// Linear cell-centred interpolation of the coarse correction e at fine cell
// (i,j,k) (fixed evaluation order, shared with the oracle).
template <int DIM, typename T>
__device__ __forceinline__ T prolong_value(const T *__restrict__ e, int cnx, int cny, int cnz,
                                           const EllBC &bc, int i, int j, int k)
{
    int xi0, xi1, yi0, yi1;
    T wx0, wx1, wy0, wy1;
    lin_w<T>(i, cnx, bc.k[0][0], bc.k[0][1], xi0, xi1, wx0, wx1);
    lin_w<T>(j, cny, bc.k[1][0], bc.k[1][1], yi0, yi1, wy0, wy1);
    if (DIM == 3) {
        int zi0, zi1;
        T wz0, wz1;
        lin_w<T>(k, cnz, bc.k[2][0], bc.k[2][1], zi0, zi1, wz0, wz1);
        T vz[2];
        int zz[2] = {zi0, zi1};
        int yy[2] = {yi0, yi1};
#pragma unroll
        for (int cz = 0; cz < 2; ++cz) {
            T vy[2];
#pragma unroll
            for (int cy = 0; cy < 2; ++cy) {
                const T *row = e + (long)cnx * (yy[cy] + (long)cny * zz[cz]);
                vy[cy] = dfma(wx1, row[xi1], wx0 * row[xi0]);
            }
            vz[cz] = dfma(wy1, vy[1], wy0 * vy[0]);
        }
        return dfma(wz1, vz[1], wz0 * vz[0]);
    } else {
        const T *r0 = e + (long)cnx * yi0;
        const T *r1 = e + (long)cnx * yi1;
        T v0 = dfma(wx1, r0[xi1], wx0 * r0[xi0]);
        T v1 = dfma(wx1, r1[xi1], wx0 * r1[xi0]);
        return dfma(wy1, v1, wy0 * v0);
    }
}

// ---------------------------------------------------- reductions (wave64)
__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// error-free sum of two doubles (Knuth TwoSum)
__device__ __forceinline__ void two_sum(double a, double b, double &s, double &e)
{
    s = a + b;
    double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// compensated pair (s, c) combine
__device__ __forceinline__ void comp_add(double &s, double &c, double s2, double c2)
{
    double t, e;
    two_sum(s, s2, t, e);
    s = t;
    c = c + (c2 + e);
}

__device__ __forceinline__ void wave_comp_sum(double &s, double &c)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        double s2 = __shfl_xor(s, o, 64);
        double c2 = __shfl_xor(c, o, 64);
        // combine in a lane-order-independent way: lower lane first
        if ((threadIdx.x & o) == 0) comp_add(s, c, s2, c2);
        else {
            double ts = s2, tc = c2;
            comp_add(ts, tc, s, c);
            s = ts;
            c = tc;
        }
    }
}

__device__ __forceinline__ void atomic_max_nonneg(unsigned long long *addr, double v)
{
    atomicMax(addr, (unsigned long long)__double_as_longlong(v));
}

}  // namespace mac
