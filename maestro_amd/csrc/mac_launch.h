// mac_launch.h -- kernel argument blocks and host launchers of
// libmaestro_mac.so (internal; the public ABI is include/maestro_mac.h).
#pragma once
#include "mac_internal.h"
#include "../../include/maestro_mac.h"

namespace mac {

// fused setup from the caller's boxes (see k_setup)
struct SetupArgs {
    int dim, nboxes;
    const BoxDev *boxes;
    int dlo[3], n[3];
    double dx[3];
    EllBC bc;
    int ng_u, ng_phi, ng_rho, ng_rhs, ng_cart;
    double *const *umac;
    double *const *phi;
    const double *const *rho;
    const double *const *macrhs;
    const double *beta0, *beta0e;
    const double *const *cart;
    int alt;
    Lev<double> L;
    double *part;
    int *err;
    float *rho32;   // optional: rho of each cell in FP32 (the FP32 solver's rho form)
};

struct CorrectArgs {
    int dim, nboxes;
    const BoxDev *boxes;
    int dlo[3], n[3];
    double dx[3];
    EllBC bc;
    int phys[3][2];
    int ng_u, ng_cart;
    double *const *umac;
    const double *beta0, *beta0e;
    const double *const *cart;
    const double *phi;
    const double *bx, *by, *bz, *bxh, *byh, *bzh;
    const int *boxmap;
    int mapg[3], mapn[3];
    long maxplane;
    double *const *phi_out;   // optional: the caller's per-box phi (valid cells written)
    int ng_phi;
};

// tiling of the one-column-per-thread fused 3D sweep (k_gsrb_deep)
struct FusedGrid {
    int ntx, nty, tz, total, per_xcd;
};

enum { SWEEP_PLAIN = 0, SWEEP_ZERO = 1 };

// FP64 vectors of the bottom solver's scratch (BiCGStab: x r r^ p v t)
static constexpr int BOTTOM_VECS = 6;

struct CoarseOpts {
    int cycle;               // MAC_CYCLE_F, _W or _V
    int bottom;              // MAC_BOTTOM_SMOOTH, _BICGSTAB or _CG
    int nu1, nu2, nub;
    double bottom_eps;
    int bottom_maxit;
    int singular;
    EllBC bc;
};

// whole-domain (natural layout) forward arguments
struct LevelIO {
    double *ux, *uy, *uz;
    const double *bx, *by, *bz;
    const double *macrhs;
    const double *phi;
    double dx[3];
    EllBC bc;
};

struct AdjIO {
    const double *gphi, *gux, *guy, *guz;
    double *out_rhs;
    double *out_ux, *out_uy, *out_uz, *out_bx, *out_by, *out_bz;
    double dx[3];
    EllBC bc;
};

// spherical base-state geometry: radial bins of width dr (centres
// (i + 1/2) dr, i < nr), the star's centre, the domain's low corner, the cell
// size and the radial interpolation type (1 constant, 2 linear, 3 quadratic)
struct SphrGeom {
    double dr;
    int nr;
    double center[3], prob_lo[3], dx[3];
    int interp;
};

// make_macrhs over the caller's boxes (see k_make_macrhs)
struct MacrhsArgs {
    int dim, nboxes;
    const BoxDev *boxes;
    int dlo[3];
    double *const *rhs;
    int ng_rhs;
    const double *const *S;
    int ng_S;
    const double *const *dg;
    int ng_dg;
    const double *const *dp;
    int ng_dp;
    double *const *dchi;
    int ng_dchi;
    const double *Sbar, *beta0, *gamma1bar, *p0, *rho0;   // 1-D, device
    double dt, dpdt;
    int dpdt_on, predictor, cutoff_coord;
    double cutoff_density;
    SphrGeom geo;
};

// make_s0mac over the caller's boxes (face arrays with one ghost layer)
struct S0macArgs {
    int nboxes;
    const BoxDev *boxes;
    int dlo[3], n[3];
    const double *s0;   // radial profile, device
    double *cells;      // whole-domain scratch: the profile on the cells
    double *const *out; // [box*3 + d]
    SphrGeom geo;
};

// mac_solve's setup over the caller's boxes (see k_solve_setup)
struct SolveArgs {
    int dim, nboxes;
    const BoxDev *boxes;
    int dlo[3], n[3];
    EllBC bc;
    const double *const *rh;
    int ng_rh;
    const double *const *phi;
    int ng_phi;
    const double *const *alpha;   // null: alpha = 0
    int ng_alpha;
    const double *const *beta;    // [box*dim + d]
    int ng_beta;
    Lev<double> L;                // level 0 (L.alpha set when alpha is given)
    double *part;
    int *err;
};

void launch_solve_setup(const SolveArgs &A, int maxcells, hipStream_t s);
// ---- two-level projection (mac_ml.hip) ------------------------------------
// region side classes: periodic (the region spans a periodic direction), on a
// physical side of the domain, or coarse-fine
enum { ML_SIDE_PER = 0, ML_SIDE_PHYS = 1, ML_SIDE_CF = 2 };
struct MlGeom {
    int dim;
    int nc[3];               // coarse cells (1 beyond dim); the domain starts at 0
    int rlo[3], nr[3];       // refined rectangle: lowest fine cell, fine extents
    double ihc[3], ihf[3];   // 1/h of the coarse and the fine level
    int ckind[3][2];         // elliptic kind of each physical side (ELL_*)
    int side[3][2];          // ML_SIDE_* of each side of the region
    // the fine level's boxes: per coarse cell of the region (nr/2 per
    // direction, x fastest), 1 where a fine box covers it; null: all
    const unsigned char *cov;
};
// whole-level arrays the composite operator reads
struct MlState {
    MlGeom G;
    const double *bc[3], *bf[3];   // face coefficients (beta0^p / rho form)
    const double *rhc, *rhf;       // right sides
    const double *phc, *phf;       // iterate
};
// caller's boxes of one level -> whole-level arrays (see k_ml_gather)
struct MlGather {
    const BoxDev *boxes;
    int org[3], n[3];        // the level array's lowest cell and extents
    int ng_u, ng_rho, ng_rhs, ng_phi;
    double *const *umac;     // [box*dim + d]
    const double *const *rho;
    const double *const *macrhs;   // null: none
    double *const *phi;
    const double *b0, *b0e;  // the level's planar beta0 over its whole domain; null: none
    int alt;
    double *u[3], *beta[3], *rhs, *phi_out;
    int *err;
    // the fine level: the region's covered-cell mask (see MlGeom), so a box
    // also writes its high face where no box follows; null on the coarse level
    const unsigned char *cov;
};
struct MlScatter {
    const BoxDev *boxes;
    int org[3], n[3];
    int ng_u, ng_phi;
    double *const *umac;
    double *const *phi;
    const double *u[3];
    const double *phi_in;
};
struct MlGrown {
    MlGeom G;
    const BoxDev *boxes;
    int ng_u;
    double *const *umac;     // the fine level's boxes
    const double *uc[3];     // final coarse faces (whole level)
};
void launch_ml_gather(int dim, const MlGather &A, int nboxes, int maxcells, hipStream_t s);
void launch_ml_edge_restrict(const MlGeom &G, double *const uc[3], const double *const uf[3], hipStream_t s);
void launch_ml_div(int dim, const int n[3], const double dx[3], const double *const u[3], const double *macrhs,
                   double *rh, hipStream_t s);
void launch_ml_cell_restrict(const MlGeom &G, double *xc, const double *xf, hipStream_t s);
void launch_ml_comp_vec(const MlGeom &G, const double *rhc, const double *rhf, double *out, hipStream_t s);
void launch_ml_resid(const MlState &S, double *rc, double *rf, unsigned long long *nrm, bool fine_only,
                     hipStream_t s);
void launch_ml_prolong(const MlGeom &G, const double *e, double *phc, double *phf, hipStream_t s);
// the fine correction plan's folded weights on the region's coarse-fine
// sides (level 0 of its hierarchy) scaled by f
void launch_ml_cf_weights(const MlGeom &G, const Lev<double> &L, double f, hipStream_t s);
// the fine correction plan's level 0 over a region with holes (cells no
// fine box covers): alpha 0 on the boxes' cells and a large value on the
// holes (their corrections stay zero), the faces between a box cell and a
// hole scaled by f
void launch_ml_holes(const MlGeom &G, const Lev<double> &L, double f, hipStream_t s);
void launch_ml_axpy(long n, const double *x, double *y, hipStream_t s);
void launch_ml_mkumac(const MlState &S, double *const uc[3], double *const uf[3], hipStream_t s);
void launch_ml_unscale(int dim, const int n[3], int org_r, double *const u[3], const double *b0, const double *b0e,
                       hipStream_t s);
void launch_ml_scatter(int dim, const MlScatter &A, int nboxes, int maxfaces, hipStream_t s);
void launch_ml_grown(int dim, const MlGrown &A, int nboxes, int maxgrown, hipStream_t s);


void launch_cell_to_edge(const double *c, int nr, double *e, hipStream_t s);
void launch_make_macrhs(const MacrhsArgs &A, int maxcells, bool sph, hipStream_t s);
void launch_make_s0mac(const S0macArgs &A, int maxfaces, hipStream_t s);

// maxcols: the largest x-y cross-section (ex * ey) of a box, maxez the
// largest z extent; one thread per column and chunk of planes of each box
void launch_setup(const SetupArgs &A, int maxcols, int maxez, hipStream_t s);
// partial sums written by launch_setup (and by any other kernel on the same
// grid); per-cell kernels pass dim 2 and their cell count as maxcols
long setup_nparts(int dim, int maxcols, int maxez, int nboxes);
void launch_mean(const double *part, long nparts, double ncell, double *mean, hipStream_t s);
template <typename T> void launch_sub_mean(T *x, long n, const double *mean, hipStream_t s);
template <typename T> long launch_partial_sum(const T *x, long n, double *part, hipStream_t s);
template <typename T> void launch_coarsen(int dim, const Lev<T> &F, const Lev<T> &C, hipStream_t s);
void launch_weights_to_f32(int dim, const Lev<double> &S, const Lev<float> &D, hipStream_t s);
template <typename T> void launch_gsrb(int dim, const Lev<T> &L, int color, hipStream_t s);
// one full red-black sweep of a 3D level, pin -> pout; mode SWEEP_*.  K
// (FP32 only) takes the weights and right side from compact storage; the
// call then returns false if the level does not fit the full-row kernel
// (compact_level_ok) and launches nothing.  A level with alpha in compact
// storage also returns false.
template <typename T>
bool launch_gsrb_fused(const Lev<T> &L, const T *pin, T *pout, int mode, hipStream_t s,
                       const Compact16 *K = nullptr);
// fine residual (of phi) restricted to C.rhs; C.phi cleared.  K as above;
// Rh: FP32 level in the rho form (the weights formed from rho)
struct RhoLev;
template <typename T>
bool launch_resid_restrict(int dim, const Lev<T> &F, const T *phi, const Lev<T> &C, hipStream_t s,
                           const Compact16 *K = nullptr, const RhoLev *Rh = nullptr);
// FP32 finest level of a plane-parallel projection in the rho form: the
// cell density and the factors the rho sweep forms its face weights from
// (parts of the level's allocation): cxy[k] = 2 beta0(k)^p per plane, cz[f]
// = 2 beta0_edge(f)^p per z face, bzl the folded low z faces (the level's
// bzh the high ones)
struct RhoLev {
    float *rho, *cxy, *cz, *bzl;
};
// whether a level takes the rho-form sweep (3D, x and y periodic, nx = 128
// or 256, no alpha)
bool rho_level_ok(int dim, const Lev<float> &L);
// the rho form's face factors (from the planar beta0 and beta0_edge), its
// z slabs and the FP32 level 1 C, from the FP64 level 0 F (F32: the FP32
// level 0, whose bzh slab is written)
void launch_rho_prep(const Lev<double> &F, const RhoLev &R, const Lev<float> &F32, const Lev<float> &C,
                     const double *b0, const double *b0e, int alt, hipStream_t s);
// one fused red-black sweep (pin -> pout) with the weights formed from rho;
// false (nothing launched) where it does not apply
bool launch_gsrb_rho(const Lev<float> &L, const RhoLev &R, const float *pin, float *pout, int mode, hipStream_t s);
// whether every level-0 kernel of the mixed path has a compact-storage form
// for this level (3D, nx = 128 or 256 and the full-row sweep fits)
bool compact_level_ok(int dim, const Lev<float> &L);
// FP32 -> bfloat16 copy of n values
void launch_to_bf16(const float *in, unsigned short *out, long n, hipStream_t s);
// compact level 0 (bfloat16 of the FP32 rounding of F's weights) and the FP32
// weights of level 1 (C), from the FP64 level 0 F in one pass (3D)
void launch_compact_level0(const Lev<double> &F, const Compact16 &K, const Lev<float> &C, hipStream_t s);
template <typename T>
void launch_residual(int dim, const Lev<T> &L, T *out, unsigned long long *nrm, hipStream_t s);
// r16 (compact level 0) replaces r32 as the output of the FP32 right side
void launch_ir_update_residual(int dim, const Lev<double> &L, const float *e, double *phi_out, float *r32,
                               unsigned long long *nrm, hipStream_t s, unsigned short *r16 = nullptr);
template <typename A, typename B> void launch_convert(const A *in, B *out, long n, hipStream_t s);
void launch_absmax(const double *x, long n, unsigned long long *nrm, hipStream_t s);
// x -= *mean with max |x - mean| into nrm, one pass
void launch_sub_mean_absmax(double *x, long n, const double *mean, unsigned long long *nrm, hipStream_t s);
// one thread: *keep = dst[0] = *src (and dst[2] = *flag where flag is
// given), then dst[1] = seq (system scope, dst mapped host memory), then
// *src = 0
void launch_publish_norm(unsigned long long *src, unsigned long long *dst, unsigned long long seq,
                         hipStream_t s, unsigned long long *keep, const int *flag = nullptr);
void launch_ir_residual(int dim, const Lev<double> &L, float *r32, unsigned long long *nrm,
                        hipStream_t s, unsigned short *r16 = nullptr);
template <typename T>
void launch_restrict(int dim, const T *r, int fnx, int fny, const Lev<T> &C, hipStream_t s);
template <typename T>
void launch_prolong(int dim, const Lev<T> &F, const Lev<T> &C, const EllBC &bc, hipStream_t s);
template <typename T> void launch_fill(T *x, long n, T v, hipStream_t s);
template <typename T>
void launch_coarse_cycle(int dim, const Lev<T> *levs_dev, const Lev<T> *levs_host, int lc, int nlev,
                         const CoarseOpts &o, double *cg, hipStream_t s);
void launch_correct(const CorrectArgs &A, int maxcols, int maxez, int maxghost, hipStream_t s);
// the ghost faces of the caller's boxes: box and periodic images (fill),
// then the physical edge-velocity rules of every non-periodic side whose
// phys code is not MAC_INTERIOR (phys)
void launch_fill_ghosts(const CorrectArgs &A, int maxghost, bool fill, bool phys, hipStream_t s);
void launch_level_setup(int dim, const LevelIO &A, const Lev<double> &L, double *part, hipStream_t s);
void launch_level_correct(int dim, const LevelIO &A, const Lev<double> &L, hipStream_t s);
void launch_adj_rhs(int dim, const AdjIO &A, const Lev<double> &L, double *part, hipStream_t s);
void launch_adj_faces(int dim, const AdjIO &A, const Lev<double> &L, const double *phi,
                      const double *lam, hipStream_t s);

}  // namespace mac
