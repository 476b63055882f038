"""ctypes wrapper of the FP64 CPU oracle (``liboracle.so``).

TEST INFRASTRUCTURE ONLY.  Imported by ``tests/``, ``__graft_entry__.smoke``
and the CPU-baseline leg of ``bench.py`` as the checker; the product package
``maestro_amd`` never imports it.

Arrays follow the oracle layout (C order, x fastest):
cells ``(nz, ny, nx)``; rho with one ghost layer ``(nz+2gz, ny+2, nx+2)``;
faces ``x:(nz,ny,nx+1)``, ``y:(nz,ny+1,nx)``, ``z:(nz+1,ny,nx)``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

PERIODIC, INTERIOR, INLET, OUTLET, SYMMETRY, SLIP_WALL, NO_SLIP_WALL = -1, 0, 11, 12, 13, 14, 15
ELL_PER, ELL_DIR, ELL_NEU = -1, 1, 2
MAX_HIST = 256


CYCLE_F, CYCLE_W, CYCLE_V = 1, 2, 3   # mg_cycle_type numbering
BOTTOM_SMOOTH, BOTTOM_BICGSTAB, BOTTOM_CG = 0, 1, 2   # mg_bottom_solver numbering


class Grid(C.Structure):
    _fields_ = [("dim", C.c_int), ("n", C.c_int * 3), ("dx", C.c_double * 3),
                ("bc", (C.c_int * 2) * 3)]


class Opts(C.Structure):
    _fields_ = [("eps_rel", C.c_double), ("eps_abs", C.c_double),
                ("max_iter", C.c_int), ("nu1", C.c_int), ("nu2", C.c_int),
                ("nub", C.c_int), ("bottom_eps", C.c_double),
                ("bottom_maxit", C.c_int), ("max_levels", C.c_int),
                ("use_alt_energy_fix", C.c_int), ("cycle_type", C.c_int),
                ("bottom_solver", C.c_int), ("omega", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("iters", C.c_int), ("converged", C.c_int), ("nlev", C.c_int),
                ("res0", C.c_double), ("res", C.c_double),
                ("hist", C.c_double * MAX_HIST)]


class Level(C.Structure):
    _fields_ = [("n", C.c_int * 3), ("h", C.c_double * 3),
                ("b", C.POINTER(C.c_double) * 3),
                ("phi", C.POINTER(C.c_double)), ("rhs", C.POINTER(C.c_double)),
                ("res", C.POINTER(C.c_double)), ("alpha", C.POINTER(C.c_double))]


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        dp = C.POINTER(C.c_double)
        L.orc_default_opts.argtypes = [C.POINTER(Opts)]
        L.orc_num_threads.restype = C.c_int
        L.orc_sum.restype = C.c_double
        L.orc_sum.argtypes = [dp, C.c_long]
        L.orc_mk_mac_coeffs.argtypes = [C.POINTER(Grid), dp, dp, dp, dp]
        L.orc_scale_faces_1d.argtypes = [C.POINTER(Grid), dp, dp, dp, dp, dp, C.c_int]
        L.orc_scale_faces_cart.argtypes = [C.POINTER(Grid), dp, dp, dp, dp, dp, dp, C.c_int]
        L.orc_divumac.argtypes = [C.POINTER(Grid), dp, dp, dp, dp, dp]
        L.orc_apply.argtypes = [C.POINTER(Grid), C.POINTER(Level), dp, dp]
        L.orc_residual.restype = C.c_double
        L.orc_residual.argtypes = [C.POINTER(Grid), C.POINTER(Level), dp, dp, dp]
        L.orc_gsrb.argtypes = [C.POINTER(Grid), C.POINTER(Level)]
        L.orc_gsrb_omega.argtypes = [C.POINTER(Grid), C.POINTER(Level), C.c_double]
        fp = C.POINTER(C.c_float)
        ip = C.POINTER(C.c_int)
        L.orc_gsrb_folded_f32.argtypes = [C.c_int, ip, ip, fp, fp, fp, fp, fp, fp, fp, fp, fp]
        L.orc_gsrb_folded_f32_omega.argtypes = [C.c_int, ip, ip, fp, fp, fp, fp, fp, fp, fp, fp, fp, C.c_float]
        L.orc_mg_nlevels.restype = C.c_int
        L.orc_mg_nlevels.argtypes = [C.POINTER(Grid), C.c_int]
        L.orc_mg_create.restype = C.c_void_p
        L.orc_mg_create.argtypes = [C.POINTER(Grid), dp, dp, dp, C.POINTER(Opts)]
        L.orc_mg_destroy.argtypes = [C.c_void_p]
        L.orc_mg_solve.restype = C.c_int
        L.orc_mg_solve.argtypes = [C.c_void_p, dp, dp, C.POINTER(Stats)]
        L.orc_mkumac.argtypes = [C.POINTER(Grid), dp, dp, dp, dp, dp, dp, dp]
        L.orc_dirichlet_rhs.argtypes = [C.POINTER(Grid), dp, dp, dp, dp, dp]
        L.orc_solve.restype = C.c_int
        L.orc_solve.argtypes = [C.POINTER(Grid), dp, dp, dp, dp, dp, dp, C.POINTER(Opts), C.POINTER(Stats)]
        for f in ("orc_level_restrict", "orc_level_prolong", "orc_level_coarsen"):
            getattr(L, f).argtypes = [C.POINTER(Grid), C.POINTER(Level), C.POINTER(Level)]
        L.orc_mg_vcycle.argtypes = [C.c_void_p]
        L.orc_macproject.restype = C.c_int
        L.orc_macproject.argtypes = [C.POINTER(Grid), dp, dp, dp, dp, dp, dp, dp, dp,
                                     dp, dp, dp, C.POINTER(Opts), C.POINTER(Stats),
                                     dp, dp, dp, dp]
        L.orc_cell_to_edge.argtypes = [dp, C.c_int, dp]
        L.orc_sphr_cell_value.restype = C.c_double
        L.orc_sphr_cell_value.argtypes = [dp, C.c_int, C.c_double, dp, dp, dp, C.c_int, C.c_int, C.c_int,
                                          C.c_int]
        L.orc_make_macrhs.argtypes = [C.c_int, ip, dp, dp, dp, dp, dp, dp, dp, dp, dp, C.c_double,
                                      C.c_double, C.c_int, C.c_double, C.c_int, C.c_double, C.c_int,
                                      dp, dp, dp, C.c_int, dp]
        L.orc_make_s0mac.argtypes = [ip, dp, C.c_int, C.c_double, dp, dp, dp, C.c_int, dp, dp, dp]
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_grid(n, dx, bc, dim=None) -> Grid:
    dim = dim or len(n)
    g = Grid()
    g.dim = dim
    nn = list(n) + [1] * (3 - len(n))
    dd = list(dx) + [1.0] * (3 - len(dx))
    for d in range(3):
        g.n[d] = int(nn[d])
        g.dx[d] = float(dd[d])
        for s in range(2):
            g.bc[d][s] = int(bc[d][s]) if d < dim else PERIODIC
    return g


def default_opts(**kw) -> Opts:
    o = Opts()
    lib().orc_default_opts(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def cell_shape(g):
    return (g.n[2], g.n[1], g.n[0])


def face_shape(g, d):
    n = [g.n[0], g.n[1], g.n[2]]
    n[d] += 1
    return (n[2], n[1], n[0])


def rho_shape(g):
    gz = 1 if g.dim == 3 else 0
    return (g.n[2] + 2 * gz, g.n[1] + 2, g.n[0] + 2)


def comp_sum(x):
    """The oracle's compensated (Neumaier) sum of a float64 vector, in its
    order: the sum behind every mean the oracle removes."""
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    return float(lib().orc_sum(_p(x), x.size))


def mk_mac_coeffs(g, rho):
    b = [np.zeros(face_shape(g, d)) for d in range(3)]
    lib().orc_mk_mac_coeffs(C.byref(g), _p(np.ascontiguousarray(rho)), _p(b[0]), _p(b[1]),
                            _p(b[2]) if g.dim == 3 else None)
    return b[:g.dim]


def divumac(g, u, v, w=None, macrhs=None):
    rh = np.zeros(cell_shape(g))
    lib().orc_divumac(C.byref(g), _p(u), _p(v), _p(w), _p(macrhs), _p(rh))
    return rh


def scale_faces_1d(g, faces, beta0, beta0_edge, do_mult):
    f = [np.array(x, dtype=np.float64, copy=True) for x in faces] + [None] * (3 - len(faces))
    lib().orc_scale_faces_1d(C.byref(g), _p(f[0]), _p(f[1]), _p(f[2]),
                             _p(np.ascontiguousarray(beta0, dtype=np.float64)),
                             _p(np.ascontiguousarray(beta0_edge, dtype=np.float64)),
                             int(do_mult))
    return f[:g.dim]


class LevelView:
    """Keeps numpy buffers alive behind an oracle ``orc_level`` (extents
    ``n`` and spacing ``h`` default to the grid's)."""

    def __init__(self, g, b, phi, rhs, h=None, n=None, res=None, alpha=None):
        self.g = g
        self.alpha = None if alpha is None else np.ascontiguousarray(alpha, dtype=np.float64)
        self.b = [np.ascontiguousarray(x, dtype=np.float64) for x in b]
        self.phi = np.ascontiguousarray(phi, dtype=np.float64)
        self.rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        self.res = np.zeros_like(self.phi) if res is None else np.ascontiguousarray(res, dtype=np.float64)
        L = Level()
        for d in range(3):
            L.n[d] = (n[d] if n is not None else g.n[d])
            L.h[d] = (h[d] if h is not None else g.dx[d]) if d < g.dim else 1.0
            L.b[d] = _p(self.b[d]) if d < g.dim else C.POINTER(C.c_double)()
        L.phi = _p(self.phi)
        L.rhs = _p(self.rhs)
        L.res = _p(self.res)
        L.alpha = _p(self.alpha) if self.alpha is not None else C.POINTER(C.c_double)()
        self.L = L


def level_restrict(g, fine: LevelView, coarse: LevelView):
    """coarse.rhs = conservative average of fine.res"""
    lib().orc_level_restrict(C.byref(g), C.byref(fine.L), C.byref(coarse.L))


def level_prolong(g, coarse: LevelView, fine: LevelView):
    """fine.phi += P coarse.phi"""
    lib().orc_level_prolong(C.byref(g), C.byref(coarse.L), C.byref(fine.L))


def level_coarsen(g, fine: LevelView, coarse: LevelView):
    """coarse.b = average of the fine faces covering each coarse face"""
    lib().orc_level_coarsen(C.byref(g), C.byref(fine.L), C.byref(coarse.L))


def apply(g, b, phi, alpha=None):
    lv = LevelView(g, b, phi, np.zeros(cell_shape(g)), alpha=alpha)
    out = np.zeros(cell_shape(g))
    lib().orc_apply(C.byref(g), C.byref(lv.L), _p(lv.phi), _p(out))
    return out


def residual(g, b, phi, rhs, alpha=None):
    lv = LevelView(g, b, phi, rhs, alpha=alpha)
    res = np.zeros(cell_shape(g))
    m = lib().orc_residual(C.byref(g), C.byref(lv.L), _p(lv.phi), _p(lv.rhs), _p(res))
    return res, m


def gsrb(g, b, phi, rhs, sweeps=1, alpha=None, omega=None):
    """Red-black sweeps (over-relaxed by ``omega`` if given)."""
    lv = LevelView(g, b, np.array(phi, dtype=np.float64, copy=True), rhs, alpha=alpha)
    for _ in range(sweeps):
        if omega is None:
            lib().orc_gsrb(C.byref(g), C.byref(lv.L))
        else:
            lib().orc_gsrb_omega(C.byref(g), C.byref(lv.L), float(omega))
    return lv.phi


def gsrb_folded_f32(dim, n, per, ih2, w, phi, rhs, sweeps=1, omega=None):
    """Red-black sweeps in FP32 on folded weights ``w`` (dict of bx, by, bz,
    bxh, byh, bzh float32 arrays as the GPU library stores a level); returns
    the new phi (float32, shape of ``phi``)."""
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    fptr = lambda a: None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))
    nn = (C.c_int * 3)(*(list(n) + [1] * (3 - len(n))))
    pp = (C.c_int * 3)(*(list(per) + [0] * (3 - len(per))))
    hh = (C.c_float * 3)(*(list(ih2) + [0.0] * (3 - len(ih2))))
    arrs = {k: f32(w[k]) for k in ("bx", "by", "bxh", "byh")}
    for k in ("bz", "bzh"):
        arrs[k] = f32(w[k]) if dim == 3 else None
    out = np.array(phi, dtype=np.float32, copy=True, order="C")
    r = f32(rhs)
    for _ in range(sweeps):
        args = (dim, nn, pp, hh, fptr(arrs["bx"]), fptr(arrs["by"]), fptr(arrs["bz"]),
                fptr(arrs["bxh"]), fptr(arrs["byh"]), fptr(arrs["bzh"]), fptr(out), fptr(r))
        if omega is None:
            lib().orc_gsrb_folded_f32(*args)
        else:
            lib().orc_gsrb_folded_f32_omega(*args, float(omega))
    return out


def mg_solve(g, b, phi, rh, opts=None):
    """FP64 MG solve of A phi = rh; returns (phi, rh_used, stats)."""
    opts = opts or default_opts()
    phi = np.array(phi, dtype=np.float64, copy=True)
    rh = np.array(rh, dtype=np.float64, copy=True)
    bb = [np.ascontiguousarray(x, dtype=np.float64) for x in b] + [None] * (3 - len(b))
    mg = lib().orc_mg_create(C.byref(g), _p(bb[0]), _p(bb[1]), _p(bb[2]), C.byref(opts))
    st = Stats()
    try:
        rc = lib().orc_mg_solve(mg, _p(phi), _p(rh), C.byref(st))
    finally:
        lib().orc_mg_destroy(mg)
    return phi, rh, st, rc


def ghost_shape(g):
    """phi with one ghost layer in each direction of the problem"""
    return (g.n[2] + 2, g.n[1] + 2, g.n[0] + 2) if g.dim == 3 else (g.n[1] + 2, g.n[0] + 2)


def ell_grid(n, dx, ell):
    """Grid whose physical codes give the elliptic kinds ``ell`` (per
    direction and side: ELL_PER, ELL_DIR or ELL_NEU) under the press_comp
    mapping of the oracle (Dirichlet <- OUTLET, Neumann <- SLIP_WALL)."""
    code = {ELL_PER: PERIODIC, ELL_DIR: OUTLET, ELL_NEU: SLIP_WALL}
    return make_grid(n, dx, [[code[k] for k in side] for side in ell])


def dirichlet_rhs(g, b, phi_g, rh):
    """rh plus the Dirichlet boundary terms (2 beta / h^2) g of the boundary
    values g held in phi_g's ghost cells"""
    out = np.array(rh, dtype=np.float64, copy=True)
    bb = [np.ascontiguousarray(x, dtype=np.float64) for x in b] + [None] * (3 - len(b))
    lib().orc_dirichlet_rhs(C.byref(g), _p(bb[0]), _p(bb[1]), _p(bb[2]),
                            _p(np.ascontiguousarray(phi_g, dtype=np.float64)), _p(out))
    return out


def solve(g, alpha, b, rh, phi_g, opts=None):
    """mac_multigrid with alpha: solve (alpha - D beta G) phi = rh (plus the
    Dirichlet boundary terms of phi_g's ghost cells).  Returns (phi_g with
    the solution in its valid cells, stats, rc)."""
    opts = opts or default_opts()
    phi_g = np.array(phi_g, dtype=np.float64, copy=True)
    bb = [np.ascontiguousarray(x, dtype=np.float64) for x in b] + [None] * (3 - len(b))
    a = None if alpha is None else np.ascontiguousarray(alpha, dtype=np.float64)
    st = Stats()
    rc = lib().orc_solve(C.byref(g), _p(a), _p(bb[0]), _p(bb[1]), _p(bb[2]),
                         _p(np.ascontiguousarray(rh, dtype=np.float64)), _p(phi_g), C.byref(opts),
                         C.byref(st))
    return phi_g, st, rc


def mkumac(g, faces, phi, b):
    f = [np.array(x, dtype=np.float64, copy=True) for x in faces] + [None] * (3 - len(faces))
    bb = list(b) + [None] * (3 - len(b))
    lib().orc_mkumac(C.byref(g), _p(f[0]), _p(f[1]), _p(f[2]), _p(phi), _p(bb[0]), _p(bb[1]),
                     _p(bb[2]))
    return f[:g.dim]


def macproject(g, faces, phi, rho, macrhs=None, beta0=None, beta0_edge=None,
               cart=None, opts=None):
    """Whole-domain oracle macproject.  Returns dict with the projected faces,
    phi, face coefficients, rh and stats."""
    opts = opts or default_opts()
    f = [np.array(x, dtype=np.float64, copy=True) for x in faces] + [None] * (3 - len(faces))
    phi = np.array(phi, dtype=np.float64, copy=True)
    b = [np.zeros(face_shape(g, d)) for d in range(g.dim)] + [None] * (3 - g.dim)
    rh = np.zeros(cell_shape(g))
    cc = list(cart) + [None] * (3 - len(cart)) if cart is not None else [None] * 3
    st = Stats()
    rc = lib().orc_macproject(
        C.byref(g), _p(f[0]), _p(f[1]), _p(f[2]), _p(phi),
        _p(np.ascontiguousarray(rho, dtype=np.float64)),
        _p(None if macrhs is None else np.ascontiguousarray(macrhs, dtype=np.float64)),
        _p(None if beta0 is None else np.ascontiguousarray(beta0, dtype=np.float64)),
        _p(None if beta0_edge is None else np.ascontiguousarray(beta0_edge, dtype=np.float64)),
        _p(cc[0]), _p(cc[1]), _p(cc[2]), C.byref(opts), C.byref(st),
        _p(b[0]), _p(b[1]), _p(b[2]), _p(rh))
    return {"faces": f[:g.dim], "phi": phi, "beta": b[:g.dim], "rh": rh, "stats": st, "rc": rc}


def num_threads() -> int:
    return lib().orc_num_threads()


# ---- producers of the projection's inputs (mac_inputs_oracle.c)
def _d3(v):
    return (C.c_double * 3)(*[float(x) for x in v])


def cell_to_edge(s0_cell):
    """cell_to_edge (Source/cell_to_edge.f90:14-68), one level, planar."""
    c = np.ascontiguousarray(s0_cell, dtype=np.float64)
    e = np.zeros(c.size + 1)
    lib().orc_cell_to_edge(_p(c), c.size, _p(e))
    return e


def sphr_cell_value(s0, dr, center, prob_lo, dx, interp, i, j, k):
    s = np.ascontiguousarray(s0, dtype=np.float64)
    return lib().orc_sphr_cell_value(_p(s), s.size, dr, _d3(center), _d3(prob_lo), _d3(dx), interp, i, j, k)


def make_macrhs(n, S, dg, Sbar, beta0, gamma1bar=None, p0=None, rho0=None, dp=None, dchi=None, dt=1.0,
                dpdt_factor=0.0, base_cutoff_coord=0, base_cutoff_density=0.0, is_predictor=True,
                sph=None, dx=None):
    """make_macrhs (Source/make_macrhs.f90) on whole-domain arrays; sph =
    (dr, center, prob_lo, interp) selects the spherical form (3D).  Returns
    rhs (and updates dchi in place)."""
    dim = len(n)
    ni = (C.c_int * 3)(*(list(n) + [1] * (3 - dim)))
    shp = tuple(reversed(n))
    rhs = np.zeros(shp)
    a = lambda x: None if x is None else np.ascontiguousarray(x, dtype=np.float64)
    S, dg, Sbar, beta0 = a(S), a(dg), a(Sbar), a(beta0)
    gamma1bar, p0, rho0, dp = a(gamma1bar), a(p0), a(rho0), a(dp)
    if dchi is not None:
        assert dchi.dtype == np.float64 and dchi.flags.c_contiguous
    dr, center, prob_lo, interp = sph if sph is not None else (0.0, (0, 0, 0), (0, 0, 0), 1)
    dxa = dx if dx is not None else (1.0, 1.0, 1.0)
    lib().orc_make_macrhs(dim, ni, _p(S), _p(dg), _p(Sbar), _p(beta0), _p(gamma1bar), _p(p0), _p(rho0),
                          _p(dp), _p(dchi), dt, dpdt_factor, base_cutoff_coord, base_cutoff_density,
                          int(is_predictor), dr, Sbar.size, _d3(center), _d3(prob_lo), _d3(dxa), interp,
                          _p(rhs))
    return rhs


def make_s0mac(n, s0, dr, center, prob_lo, dx, interp):
    """make_s0mac (type 1) on the grown face ranges: returns (fx, fy, fz) of
    extents (nz+2, ny+2, nx+3), (nz+2, ny+3, nx+2), (nz+3, ny+2, nx+2);
    face index -1 is at position 0."""
    s = np.ascontiguousarray(s0, dtype=np.float64)
    nx, ny, nz = n
    f = [np.zeros((nz + 2 + (d == 2), ny + 2 + (d == 1), nx + 2 + (d == 0))) for d in range(3)]
    lib().orc_make_s0mac((C.c_int * 3)(*n), _p(s), s.size, dr, _d3(center), _d3(prob_lo), _d3(dx), interp,
                         _p(f[0]), _p(f[1]), _p(f[2]))
    return f
