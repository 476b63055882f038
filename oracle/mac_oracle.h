/*
 * mac_oracle.h -- FP64 CPU restatement of MAESTRO's MAC projection.
 * TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the CPU
 * baseline of bench.py as the checker; never linked by the product library.
 */
#ifndef MAC_ORACLE_H
#define MAC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* physical BC codes (FBoxLib bc_module values; OUTLET=12 and SYMMETRY=13
 * also appear in Exec/SCIENCE/sub_chandra/paper_II/PyRoot/subchandra.py) */
enum {
    ORC_PERIODIC = -1,
    ORC_INTERIOR = 0,
    ORC_INLET = 11,
    ORC_OUTLET = 12,
    ORC_SYMMETRY = 13,
    ORC_SLIP_WALL = 14,
    ORC_NO_SLIP_WALL = 15
};
enum { ORC_ELL_PER = -1, ORC_ELL_DIR = 1, ORC_ELL_NEU = 2 };
enum { ORC_OK = 0, ORC_ERR_INVALID = 1, ORC_ERR_NOT_CONVERGED = 4 };

#define ORC_MAX_HIST 256

/* cycle shapes, numbered as mg_cycle_type (Source/_parameters:376-377) */
enum { ORC_CYCLE_F = 1, ORC_CYCLE_W = 2, ORC_CYCLE_V = 3 };
/* bottom solvers, numbered as mg_bottom_solver (Docs/mg/mg.tex:239-252) */
enum { ORC_BOTTOM_SMOOTH = 0, ORC_BOTTOM_BICGSTAB = 1, ORC_BOTTOM_CG = 2 };

typedef struct {
    int dim;          /* 2 or 3 */
    int n[3];         /* cells per direction; n[2] = 1 in 2D */
    double dx[3];
    int bc[3][2];     /* physical BC code per direction and side (lo, hi) */
} orc_grid;

typedef struct {
    double eps_rel, eps_abs;
    int max_iter, nu1, nu2, nub;
    double bottom_eps;
    int bottom_maxit;
    int max_levels;   /* 0 = coarsen as far as possible */
    int use_alt_energy_fix;
    int cycle_type;   /* ORC_CYCLE_F/W/V (mg_cycle_type); 0 = V */
    int bottom_solver; /* ORC_BOTTOM_SMOOTH/BICGSTAB/CG (mg_bottom_solver) */
    double omega;     /* over-relaxation of the red-black sweeps on the levels
                         above ORC_SOR_MAX_CELLS cells; 1 = Gauss-Seidel (the
                         reference's smoother).  No reference counterpart:
                         the smoother of the library's FP32 cycles.          */
} orc_opts;

/* levels of at most this many cells (3D; 2D: ORC_SOR_MAX_CELLS_2D) are swept
 * without over-relaxation: the library runs them in its single-workgroup
 * coarse cycle, which keeps plain Gauss-Seidel                               */
#define ORC_SOR_MAX_CELLS 512
#define ORC_SOR_MAX_CELLS_2D 1024

typedef struct {
    int iters, converged, nlev;
    double res0, res;
    double hist[ORC_MAX_HIST];
} orc_stats;

typedef struct {
    int n[3];
    double h[3];
    double *b[3];     /* raw face coefficients, natural (n+1)-in-d layout */
    double *phi, *rhs, *res;
    double *alpha;    /* cell coefficient of (alpha - D beta G); NULL: 0 */
} orc_level;

typedef struct {
    orc_grid g;
    orc_opts opts;
    int nlev;
    orc_level *lev;
    int singular;     /* no Dirichlet side and alpha == 0 */
} orc_mg;

void orc_default_opts(orc_opts *o);
int orc_num_threads(void);
int orc_ell_bc(int phys);
int orc_is_singular(const orc_grid *g);
double orc_sum(const double *x, long n);

void orc_mk_mac_coeffs(const orc_grid *g, const double *rho, double *bx,
                       double *by, double *bz);
void orc_scale_faces_1d(const orc_grid *g, double *fx, double *fy, double *fz,
                        const double *beta0, const double *beta0_edge, int do_mult);
void orc_scale_faces_cart(const orc_grid *g, double *fx, double *fy, double *fz,
                          const double *cx, const double *cy, const double *cz,
                          int do_mult);
void orc_divumac(const orc_grid *g, const double *u, const double *v,
                 const double *w, const double *macrhs, double *rh);

void orc_apply(const orc_grid *g, const orc_level *lev, const double *phi,
               double *out);
double orc_residual(const orc_grid *g, const orc_level *lev, const double *phi,
                    const double *rhs, double *res);
void orc_gsrb(const orc_grid *g, orc_level *lev);
void orc_gsrb_omega(const orc_grid *g, orc_level *lev, double omega);
void orc_gsrb_folded_f32(int dim, const int *n, const int *per, const float *ih2,
                         const float *bx, const float *by, const float *bz,
                         const float *bxh, const float *byh, const float *bzh,
                         float *phi, const float *rhs);
void orc_gsrb_folded_f32_omega(int dim, const int *n, const int *per, const float *ih2,
                               const float *bx, const float *by, const float *bz,
                               const float *bxh, const float *byh, const float *bzh,
                               float *phi, const float *rhs, float omega);

void orc_level_restrict(const orc_grid *g, const orc_level *f, orc_level *c);
void orc_level_prolong(const orc_grid *g, const orc_level *c, orc_level *f);
void orc_level_coarsen(const orc_grid *g, const orc_level *f, orc_level *c);
void orc_mg_vcycle(orc_mg *mg);

int orc_mg_nlevels(const orc_grid *g, int max_levels);
orc_mg *orc_mg_create(const orc_grid *g, const double *bx, const double *by,
                      const double *bz, const orc_opts *o);
orc_mg *orc_mg_create_alpha(const orc_grid *g, const double *alpha, const double *bx,
                            const double *by, const double *bz, const orc_opts *o);
void orc_mg_destroy(orc_mg *mg);
void orc_dirichlet_rhs(const orc_grid *g, const double *bx, const double *by, const double *bz,
                       const double *phi_g, double *rh);
int orc_solve(const orc_grid *g, const double *alpha, const double *bx, const double *by,
              const double *bz, const double *rh, double *phi_g, const orc_opts *o, orc_stats *st);
int orc_mg_solve(orc_mg *mg, double *phi, double *rh, orc_stats *st);

void orc_mkumac(const orc_grid *g, double *u, double *v, double *w,
                const double *phi, const double *bx, const double *by,
                const double *bz);

int orc_macproject(const orc_grid *g, double *u, double *v, double *w,
                   double *phi, const double *rho, const double *macrhs,
                   const double *beta0, const double *beta0_edge,
                   const double *cart_x, const double *cart_y,
                   const double *cart_z, const orc_opts *o, orc_stats *st,
                   double *bx_out, double *by_out, double *bz_out,
                   double *rh_out);

/* producers of the projection's inputs (mac_inputs_oracle.c) */
void orc_cell_to_edge(const double *c, int nr, double *e);
double orc_sphr_cell_value(const double *s0, int nr, double dr, const double *center,
                           const double *prob_lo, const double *dx, int type, int i, int j, int k);
void orc_make_macrhs(int dim, const int *n, const double *S, const double *dg, const double *Sbar,
                     const double *beta0, const double *gamma1bar, const double *p0,
                     const double *rho0, const double *dp, double *dchi, double dt,
                     double dpdt_factor, int base_cutoff_coord, double base_cutoff_density,
                     int is_predictor, double sph_dr, int sph_nr, const double *center,
                     const double *prob_lo, const double *dx, int interp, double *rhs);
void orc_make_s0mac(const int *n, const double *s0, int nr, double dr, const double *center,
                    const double *prob_lo, const double *dx, int interp, double *fx, double *fy,
                    double *fz);

#ifdef __cplusplus
}
#endif
#endif
