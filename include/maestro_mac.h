/*
 * maestro_mac.h -- C ABI of the MI355X (gfx950) MAC-projection library
 * (libmaestro_mac.so).
 *
 * This library stands in for MAESTRO's MAC projection (one level, or two
 * through mac_ml_project):
 *   macproject(mla, umac, phi, rho, dx, the_bc_tower
 *              [, macrhs][, beta0_1d, beta0_1d_edge][, beta0_cart_edge])
 *     Source/macproject.f90:51-76
 * together with the FBoxLib cell-centred multigrid solve that
 *   mac_multigrid(mla, rh, phi, fine_flx, alpha, beta, dx, bc_tower, bc_comp,
 *                 stencil_order, rel_eps, abs_eps)
 *     Source/mac_multigrid.f90:18-19,56-67
 * hands it to.  The Fortran module maestro_amd/fortran/macproject_hip.f90
 * binds these entry points through iso_c_binding under the reference's own
 * name and argument list; maestro_amd/mac.py binds them with ctypes.
 *
 * Data layout (per box, as FBoxLib's fab storage, architecture.tex:617-644):
 * every array is Fortran column-major over global 0-based indices,
 *   cell array  a(lo1-ng:hi1+ng, lo2-ng:hi2+ng[, lo3-ng:hi3+ng])
 *   face array of direction d: the same, with hi_d+1+ng in direction d.
 * The pointer handed in for a box addresses element (lo-ng, ...).  Arrays
 * of box pointers are ordered box-major: umac[b*dim + d].
 *
 * mac_project is single level (the boxes must tile the domain exactly);
 * mac_ml_project takes two levels.
 * Return codes: MAC_OK, or one of MAC_ERR_* below.
 */
#ifndef MAESTRO_MAC_H
#define MAESTRO_MAC_H

#ifdef __cplusplus
extern "C" {
#endif

/* physical boundary codes (FBoxLib bc_module) */
#define MAC_PERIODIC (-1)
#define MAC_INTERIOR 0
#define MAC_INLET 11
#define MAC_OUTLET 12
#define MAC_SYMMETRY 13
#define MAC_SLIP_WALL 14
#define MAC_NO_SLIP_WALL 15

/* return codes */
#define MAC_OK 0
#define MAC_ERR_INVALID 1        /* invalid argument or option combination */
#define MAC_ERR_LAYOUT 2         /* boxes do not tile the domain, bad ghosts */
#define MAC_ERR_NONFINITE 3      /* NaN/Inf in rho, umac or the right side */
#define MAC_ERR_NOT_CONVERGED 4  /* max_iter reached; stats are filled */
#define MAC_ERR_HIP 5            /* a HIP runtime call failed */
#define MAC_ERR_NOMEM 6
#define MAC_ERR_TOO_LARGE 7      /* mac_project_noalloc: the call needs more than its fixed scratch */

/* cycle shapes, numbered as mg_cycle_type (Source/_parameters:376-377): a V
 * visits the next coarser level once, a W twice (each visit a W), an F first
 * as an F and then as a V; a second visit starts from the first one's iterate */
#define MAC_CYCLE_F 1
#define MAC_CYCLE_W 2
#define MAC_CYCLE_V 3

/* solvers of the bottom level, numbered as mg_bottom_solver (Docs/mg/mg.tex:
 * 239-252), each between mac_opts.nub red-black sweeps: none (smoothing
 * only), BiCGStab (FBoxLib's default) or CG, both FP64 to bottom_eps.  The
 * hierarchy always coarsens the whole domain down to two cells per direction,
 * as FBoxLib's agglomerating bottom solver (mg_bottom_solver = 4) does.     */
#define MAC_BOTTOM_SMOOTH 0
#define MAC_BOTTOM_BICGSTAB 1
#define MAC_BOTTOM_CG 2

/* precision of the V-cycle */
#define MAC_PREC_FP64 0          /* plain FP64 V-cycle (the reference) */
#define MAC_PREC_MIXED 1         /* FP32 V-cycle inside FP64 refinement */
#define MAC_PREC_MIXED_BF16 2    /* as MAC_PREC_MIXED, with the finest level's
                                    face weights and right side stored as
                                    bfloat16 (iterate and arithmetic FP32) where
                                    the 3D full-row kernels apply (nx a multiple
                                    of 128, nx <= 256); FP32 storage elsewhere */

typedef struct mac_layout {
    int dim;                /* 2 or 3 */
    int nboxes;
    const int *box_lo;      /* [nboxes*3] inclusive lower cell index (z unused in 2D) */
    const int *box_hi;      /* [nboxes*3] inclusive upper cell index */
    int domain_lo[3];       /* problem domain, inclusive */
    int domain_hi[3];
    double dx[3];
    int phys_bc[3][2];      /* [direction][lo, hi] physical BC code */
    int ng_umac;            /* ghost widths of the caller's arrays */
    int ng_phi;
    int ng_rho;             /* >= 1: rho ghosts feed the boundary coefficients */
    int ng_macrhs;
    int ng_beta0_cart;
} mac_layout;

typedef struct mac_opts {
    double eps_rel;          /* relative residual target; default 1e-10
                                (min(eps_mac*10^(nlevs-1), eps_mac_max),
                                Source/macproject.f90:217, mg_eps.f90:22-28) */
    double eps_abs;          /* absolute target; <= 0 disables (:209)       */
    int max_iter;            /* V-cycles (or refinement steps); default 200 */
    int nu1, nu2;            /* GSRB sweeps down / up; default 2, 2         */
    int nub;                 /* sweeps around the bottom solve; default 10
                                (mg_bottom_nu, Source/_parameters:392)       */
    double bottom_eps;       /* bottom solver's relative tolerance; default 1e-3
                                (Source/mac_multigrid.f90:61)                */
    int bottom_max_iter;     /* default 1000 */
    int max_levels;          /* 0: coarsen down to 2 cells per direction     */
    int use_alt_energy_fix;  /* 1: beta = beta0^2/rho (reference default,
                                Source/_parameters:552-554); 0: beta0/rho    */
    int precision;           /* MAC_PREC_FP64, MAC_PREC_MIXED or
                                MAC_PREC_MIXED_BF16                          */
    int inner_cycles;        /* FP32 V-cycles per refinement step (mixed);
                                default 0: 3 with FP32 storage, 2 with
                                bfloat16 storage                             */
    int device_ptrs;         /* 1: every array pointer is a device pointer;
                                0: host pointers (copied in and out)         */
    int verbose;
    int cycle_type;          /* MAC_CYCLE_F, _W or _V; 0 = V (mg_cycle_type,
                                Source/_parameters:376-377, passed as
                                cycle_type in Source/mac_multigrid.f90:60)   */
    int bottom_solver;       /* MAC_BOTTOM_SMOOTH, _BICGSTAB or _CG; default
                                CG (mg_bottom_solver, Source/_parameters:386,
                                Source/mac_multigrid.f90:62)                 */
    double omega;            /* over-relaxation of the red-black sweeps on the
                                levels above the single-workgroup coarse
                                cycle (a cell moves to old + omega (gs - old)),
                                0 < omega < 2; default 0: 1, FBoxLib's GSRB.
                                1.25 takes the 256^3 plane-parallel solve from
                                9 to 6 V-cycles to the same relative residual,
                                with phi 6.5e-9 from the exact discrete
                                solution against 5.1e-10.  No reference
                                counterpart (the reference's smoother is
                                GS_RB).                                      */
} mac_opts;

typedef struct mac_stats {
    int iters;               /* V-cycles (FP64) or refinement steps (mixed) */
    int inner_cycles;        /* FP32 V-cycles in total (mixed)             */
    int converged;
    int nlevels;
    double rhs_norm;         /* max |rh| after mean removal               */
    double res0;             /* max |rh - A phi0|  ("error0")             */
    double res;              /* final max |rh - A phi|                     */
    double t_h2d_ms, t_setup_ms, t_solve_ms, t_correct_ms, t_d2h_ms, t_total_ms;
} mac_stats;

typedef struct mac_plan mac_plan;

/* Fill *o with the defaults listed above. */
int mac_default_opts(mac_opts *o);

/* Build a plan: validates the layout, allocates the device MG hierarchy and
 * scratch.  One plan serves any number of calls on the same layout (the two
 * MAC projections of a time step, Source/advance.f90:466,937).  The plan
 * owns one HIP stream; all work of a call runs on it and the call returns
 * after it completes. */
int mac_plan_create(const mac_layout *layout, const mac_opts *opts, mac_plan **plan);
int mac_plan_destroy(mac_plan *plan);

/* Change solver options of an existing plan (tolerances, precision,
 * device_ptrs, ...).  Layout-dependent fields are ignored. */
int mac_plan_set_opts(mac_plan *plan, const mac_opts *opts);

/* macproject, Source/macproject.f90:51-328 (single level):
 *   umac  [nboxes*dim]  in/out edge velocities (ghosts filled on return:
 *                       box/periodic images, then physical edge BCs)
 *   phi   [nboxes]      in/out, warm start (ng_phi ghosts, not read)
 *   rho   [nboxes]      cell density with >= 1 filled ghost layer
 *   macrhs[nboxes]      optional (NULL): rh = macrhs - D(beta0 U)
 *   beta0_1d[nr], beta0_1d_edge[nr+1]   optional planar beta0 along the
 *                       last direction (nr = domain cells in it)
 *   beta0_cart_edge[nboxes*dim]  optional face beta0 (spherical); at most
 *                       one of the two beta0 forms may be given (:99-109)
 *   stats               optional output */
int mac_project(mac_plan *plan, double *const *umac, double *const *phi,
                const double *const *rho, const double *const *macrhs,
                const double *beta0_1d, const double *beta0_1d_edge,
                const double *const *beta0_cart_edge, mac_stats *stats);

/* mac_project that allocates nothing.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory during the call (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, results (bit for bit) and return codes as mac_project, plus:
 *   - opts.device_ptrs must be 1, and the plan's solver storage for its
 *     precision must exist: the FP64 levels do from mac_plan_create on, the
 *     FP32 levels (and the compact finest level of MAC_PREC_MIXED_BF16) after
 *     one mac_project call in that precision; otherwise MAC_ERR_INVALID;
 *   - the call's own device buffers (the pointer table, the partial sums of
 *     the solvability mean, the mean, an error flag) live in a fixed 2 MiB
 *     per-GPU array of the library: nboxes * (2*dim + 3) pointers plus 16
 *     bytes per partial-sum block (about nboxes * ceil(box cross-section /
 *     256) * ceil(box z extent / 8) blocks) must fit, else MAC_ERR_TOO_LARGE
 *     and nothing runs (65536 boxes' worth of blocks at 64^3 per box, so
 *     512^3 in 64^3 boxes fits);
 *   - calls that overlap on one GPU share that array and race. */
int mac_project_noalloc(mac_plan *plan, double *const *umac, double *const *phi,
                        const double *const *rho, const double *const *macrhs,
                        const double *beta0_1d, const double *beta0_1d_edge,
                        const double *const *beta0_cart_edge, mac_stats *stats);

/* ---- whole-domain device entry points (used by the autograd wrapper) ----
 * Arrays are single device buffers over the whole domain, C order with x
 * fastest: cells (nz,ny,nx); x faces (nz,ny,nx+1); y faces (nz,ny+1,nx);
 * z faces (nz+1,ny,nx).  beta_* are the solver face coefficients (already
 * including beta0).  They require opts.device_ptrs = 1. */

/* Forward: rh = macrhs - D(U); solve (-D beta G) phi = rh; U -= beta G phi.
 * Same operator, boundary handling and stopping rule as mac_project. */
int mac_project_level(mac_plan *plan, double *ux, double *uy, double *uz,
                      double *phi, const double *beta_x, const double *beta_y,
                      const double *beta_z, const double *macrhs, mac_stats *stats);

/* Adjoint of mac_project_level for L(phi, U_proj): given the forward phi
 * (mean-removed where singular), dL/dphi (g_phi, cells) and dL/dU_proj
 * (g_ux.., faces), solve A lambda = g_phi - G^T(beta g_U) with the same
 * self-adjoint operator and return dL/dU (out_ux..) and dL/dbeta
 * (out_bx..) per face, and lambda (cells). */
int mac_adjoint_level(mac_plan *plan, const double *phi, const double *beta_x,
                      const double *beta_y, const double *beta_z,
                      const double *g_phi, const double *g_ux, const double *g_uy,
                      const double *g_uz, double *lambda, double *out_ux,
                      double *out_uy, double *out_uz, double *out_bx,
                      double *out_by, double *out_bz, mac_stats *stats);

/* ---- producers of the projection's inputs --------------------------------
 * Per-box arrays as for mac_project (the plan's layout; pointers follow
 * opts.device_ptrs, 1-D arrays included).  Each is one pass on the plan's
 * stream and returns when it completes. */

/* cell_to_edge (Source/cell_to_edge.f90:14-68, one level, planar): the
 * radial edge values s0_edge[0..nr] of the bin values s0_cell[0..nr-1] --
 * edges 0 and nr copy the end bins, edges 1 and nr-1 average their two bins,
 * the others take the fourth-order average 7/12 (s_r + s_r-1) - 1/12 (s_r+1
 * + s_r-2) limited to the range of s_r-1, s_r.  nr >= 2. */
int mac_cell_to_edge(mac_plan *plan, const double *s0_cell, int nr, double *s0_edge);

/* Spherical base-state geometry (Source/geometry.f90:140-156,
 * initialize.f90:276-291): nr radial bins of width dr with centres
 * (i + 1/2) dr; the star's centre and the domain's low corner in physical
 * units; interp = s0_interp_type of put_1d_array_on_cart_3d_sphr
 * (fill_3d_data.f90:406-531): 1 piecewise constant, 2 piecewise linear,
 * 3 quadratic (the reference default, _parameters:632). */
typedef struct mac_sphr {
    double dr;
    int nr;
    double center[3];
    double prob_lo[3];
    int interp;
} mac_sphr;

/* Inputs of make_macrhs (Source/make_macrhs.f90:24-342).  Per-box cell
 * arrays with their ghost widths; 1-D arrays indexed by the radial
 * coordinate (the last direction's cell index when planar, the radial bin
 * when spherical).  The dp/dt term is on when dpdt_factor > 0 (planar) or
 * >= 0 (spherical, as :306 of the reference); it then needs delta_p_term,
 * delta_chi, gamma1bar, p0 and, spherical, rho0. */
typedef struct mac_macrhs_in {
    const double *const *S_cc;               int ng_S;
    const double *const *delta_gamma1_term;  int ng_dg;
    const double *const *delta_p_term;       int ng_dp;     /* may be NULL if the term is off */
    double *const *delta_chi;                int ng_dchi;   /* in/out; NULL if the term is off */
    const double *Sbar, *beta0, *gamma1bar, *p0, *rho0;
    double dt, dpdt_factor;
    int base_cutoff_coord;          /* planar: the term applies below this radial index */
    double base_cutoff_density;     /* spherical: ... where rho0 exceeds this density   */
    int is_predictor;               /* delta_chi starts from zero (whole arrays)       */
} mac_macrhs_in;

/* make_macrhs: macrhs = beta0 (S_cc - Sbar + delta_gamma1_term) per valid
 * cell, plus beta0 delta_chi with delta_chi += dpdt_factor delta_p_term /
 * (dt gamma1bar p0) where the dp/dt term applies.  sph NULL: planar (2D/3D);
 * otherwise spherical (3D), the 1-D profiles put on the cells by sph->interp.
 * macrhs uses the layout's ng_macrhs. */
int mac_make_macrhs(mac_plan *plan, double *const *macrhs, const mac_macrhs_in *in,
                    const mac_sphr *sph);

/* make_s0mac, s0mac_interp_type 1 (Source/fill_3d_data.f90:942-1075): the
 * radial profile s0[0..sph->nr-1] on the cell centres, cells outside the
 * domain by first-order extrapolation, then every face of the per-box face
 * arrays s0mac[box*3 + d] (one ghost layer, the layout's ng_beta0_cart
 * must be 1) the mean of its two cells.  3D. */
int mac_make_s0mac(mac_plan *plan, const double *s0, const mac_sphr *sph, double *const *s0mac);

/* ---- solver-level entry point ---------------------------------------------
 * mac_multigrid(mla, rh, phi, fine_flx, alpha, beta, dx, the_bc_tower,
 *               bc_comp, stencil_order, rel_solver_eps, abs_solver_eps)
 *   Source/mac_multigrid.f90:18-71
 * solves (alpha - D beta G) phi = rh with caller-built coefficients: alpha
 * = 0 from macproject, alpha = rho and beta = (dt/2) k/c_p on faces from the
 * thermal diffusion solve (Source/thermal_conduct.f90:316-377, bc_comp =
 * dm + rhoh_comp).  Same V-cycle, bottom solver and precision options as
 * mac_project; the solves with alpha run the generic (per-colour) kernels.
 * Single level: fine_flx is not produced (the single-level callers discard
 * it, thermal_conduct.f90:366-368). */

/* elliptic boundary kinds (FBoxLib bc_module's BC_PER, BC_DIR, BC_NEU) */
#define MAC_ELL_PER (-1)
#define MAC_ELL_DIR 1
#define MAC_ELL_NEU 2

/* component classes of ell_bc_level_build (Source/define_bc_tower.f90:296-402) */
#define MAC_COMP_PRESS 0    /* press_comp: OUTLET Dirichlet                    */
#define MAC_COMP_SCALAR 1   /* rho, rho h, rho X, temp, pi, tracers: INLET
                               Dirichlet                                       */
#define MAC_COMP_EXTRAP 2   /* foextrap_comp, hoextrap_comp: Neumann           */

/* The elliptic kind of a physical boundary code for a component class:
 * MAC_ELL_PER for MAC_PERIODIC, else MAC_ELL_DIR or MAC_ELL_NEU as
 * ell_bc_level_build assigns them; 0 for a code it does not list. */
int mac_ell_bc(int phys_bc, int comp);

typedef struct mac_solve_in {
    const double *const *rh;     /* [nboxes] right side                          */
    int ng_rh;
    double *const *phi;          /* [nboxes] in: warm start (valid cells) and the
                                    Dirichlet values (the ghost cells beyond each
                                    MAC_ELL_DIR side hold the value on that face;
                                    ng_phi >= 1 then); out: the solution on the
                                    valid cells (ghost cells untouched)           */
    int ng_phi;
    const double *const *alpha;  /* [nboxes] cell coefficient; NULL: alpha = 0   */
    int ng_alpha;
    const double *const *beta;   /* [nboxes*dim] face coefficients, box-major    */
    int ng_beta;
    int ell_bc[3][2];            /* MAC_ELL_* per direction and side; periodic
                                    exactly where the layout is periodic          */
    double eps_rel;              /* stop at max|r| <= eps_rel * max|r0| ...       */
    double eps_abs;              /* ... or max|r| <= eps_abs (<= 0: off)          */
} mac_solve_in;

/* Solve on the plan's layout.  The problem is singular (the mean of the
 * right side is removed, phi defined up to a constant) when no side is
 * Dirichlet and alpha is NULL or zero everywhere. */
int mac_solve(mac_plan *plan, const mac_solve_in *in, mac_stats *stats);

/* ---- two-level (AMR) projection ------------------------------------------
 * macproject(mla, umac, phi, rho, dx, the_bc_tower[, macrhs]
 *            [, beta0_1d, beta0_1d_edge]) with nlevs = 2, refinement ratio 2
 *   Source/macproject.f90:51-328 (the nlevs > 1 branches: :141-143 via
 *   divumac :357-361 and :401-406, :217, :256-275, mkumac :829-833)
 * layouts[0]: the coarse level, boxes tiling its domain (which starts at
 * 0); layouts[1]: the fine level, domain the coarse one refined by 2, same
 * physical BCs and ghost conventions, disjoint boxes that each cover whole
 * coarse cells (any union: an L, separate patches).  The composite problem (FBoxLib ml_cc_solve,
 * absent from the reference tree; Docs/mg/mg.tex:9-19,124-156): fine cells
 * next to a coarse-fine face see the ghost value 8/15 C + 2/3 phi_0 - 1/5
 * phi_1 (C: the coarse value across the face moved to the ghost's
 * transverse position with linear slopes), coarse cells there take the mean
 * of the fine fluxes.  Solved by fast adaptive composite iteration with the
 * two levels' V-cycles to opts.eps_rel relative (MAESTRO passes
 * min(eps_mac * 10^(nlevs-1), eps_mac_max) = 1e-9 for two levels, :217).  Host
 * pointers (opts.device_ptrs = 0) are copied through one device allocation
 * per call; beta0_cart_edge (spherical) is not taken.  Per level l: umac[l][box*dim + d], phi[l][box], rho[l][box]
 * (one filled ghost layer), macrhs[l][box] or macrhs NULL, beta0_1d[l]
 * and beta0_1d_edge[l] over level l's whole radial extent or both NULL.
 * Returned: both levels' umac (ghost faces filled: the coarse level as by
 * mac_project, the fine level from its boxes, the coarse faces beyond the
 * refined region and the physical rules) and phi (covered coarse cells the
 * mean of their fine cells); stats.iters = composite steps. */
typedef struct mac_ml_plan mac_ml_plan;
int mac_ml_plan_create(int nlevs, const mac_layout *layouts, const mac_opts *opts, mac_ml_plan **plan);
int mac_ml_plan_destroy(mac_ml_plan *plan);
int mac_ml_plan_set_opts(mac_ml_plan *plan, const mac_opts *opts);
int mac_ml_project(mac_ml_plan *plan, double *const *const *umac, double *const *const *phi,
                   const double *const *const *rho, const double *const *const *macrhs,
                   const double *const *beta0_1d, const double *const *beta0_1d_edge, mac_stats *stats);

/* ---- diagnostics (no counterpart in the reference) ----------------------
 * Read or overwrite an internal FP64 array of MG level `level` (0 = finest)
 * through host memory, and run single solver operators on it, so tests can
 * compare each kernel with a CPU restatement.  `what`: MAC_ARR_* below;
 * sizes are the level's cell count, or the slab sizes ny*nz, nx*nz, nx*ny
 * for the high-face weights.  `op`: MAC_OP_* below, repeated `count` times. */
#define MAC_ARR_PHI 0
#define MAC_ARR_RHS 1
#define MAC_ARR_BX 2
#define MAC_ARR_BY 3
#define MAC_ARR_BZ 4
#define MAC_ARR_BXH 5
#define MAC_ARR_BYH 6
#define MAC_ARR_BZH 7
#define MAC_ARR_RES 8
#define MAC_ARR_F32 0x10     /* or-ed in (mac_debug_get only): the FP32 hierarchy's
                                copy of the array, widened to FP64            */
#define MAC_ARR_BF16 0x20    /* or-ed in (mac_debug_get only, level 0): the
                                compact bfloat16 copy of a weight array or the
                                right side, widened to FP64                    */
#define MAC_OP_GSRB 0        /* one red-black sweep on `level`             */
#define MAC_OP_RESIDUAL 1    /* res = rhs - A phi on `level`              */
#define MAC_OP_RESTRICT 2    /* rhs(level+1) = average of res(level)      */
#define MAC_OP_PROLONG 3     /* phi(level) += P phi(level+1)              */
#define MAC_OP_COARSEN 4     /* weights(level+1) from weights(level)      */
#define MAC_OP_VCYCLE 5      /* one cycle (of opts.cycle_type) from `level` (0 only) */
#define MAC_OP_RESID_RESTRICT 6  /* rhs(level+1) = average of (rhs - A phi)(level),
                                    phi(level+1) = 0, one fused pass       */
#define MAC_OP_PROLONG_SWEEP 7   /* phi(level) += P phi(level+1), then one
                                    sweep, as the V-cycle's first upward sweep */
#define MAC_OP_GSRB32 8          /* `count` sweeps in FP32 with the kernels of the
                                    mixed path: weights, rhs and phi of the level
                                    rounded to FP32 (weights and rhs further to
                                    bfloat16 on level 0 of a MAC_PREC_MIXED_BF16
                                    plan where it stores them so), the result
                                    widened into phi */
#define MAC_OP_GSRB32_RHO 9      /* `count` FP32 sweeps of level 0 in the rho form
                                    (face weights formed from the FP32 density):
                                    rho and its factors as the last MAC_PREC_MIXED
                                    plane-parallel mac_project on this plan left
                                    them, rhs and phi of level 0 rounded to FP32;
                                    MAC_ERR_INVALID where the form does not apply */
int mac_debug_level_dims(mac_plan *plan, int level, int dims[3]);
int mac_debug_get(mac_plan *plan, int level, int what, double *out);
int mac_debug_set(mac_plan *plan, int level, int what, const double *in);
int mac_debug_op(mac_plan *plan, int level, int op, int count);

/* Kernel timing on the plan's stream (no reference counterpart): when
 * enabled, launches of the finest-level kernels are bracketed by HIP
 * events; mac_plan_profile_read returns the number of timed launches and
 * their summed duration for one class.  Enabling resets the counters. */
#define MAC_PROF_GSRB64 0     /* finest-level smoothing launches reading phi, FP64 */
#define MAC_PROF_GSRB32 1     /* finest-level smoothing launches reading phi, FP32 */
#define MAC_PROF_RESID64 2    /* finest-level FP64 residual/norm launches */
#define MAC_PROF_SETUP 3      /* fused setup from the caller's boxes */
#define MAC_PROF_CORRECT 4    /* face correction + ghost fill */
#define MAC_PROF_COARSE 5     /* single-workgroup coarse-level cycle */
#define MAC_PROF_GSRB64_ZERO 6   /* finest-level first sweeps from a zero iterate, FP64 */
#define MAC_PROF_GSRB32_ZERO 7   /* finest-level first sweeps from a zero iterate, FP32 */
#define MAC_PROF_RESTRICT64 8    /* finest-level residual + restriction, FP64 */
#define MAC_PROF_RESTRICT32 9    /* finest-level residual + restriction, FP32 */
#define MAC_PROF_PROLONG64 10    /* finest-level prolongation, FP64 */
#define MAC_PROF_PROLONG32 11    /* finest-level prolongation, FP32 */
#define MAC_PROF_IR_UPDATE 12    /* phi64 += e32 of the refinement */
#define MAC_PROF_NCLASS 13
int mac_plan_profile(mac_plan *plan, int enable);
int mac_plan_profile_read(mac_plan *plan, int cls, long *count, double *total_ms);

/* Library version string. */
const char *mac_version(void);

#ifdef __cplusplus
}
#endif
#endif
