"""Python binding of ``libmaestro_mac.so`` (the gfx950 MAC projection).

Mirrors the reference's MAC-projection interface,
``macproject(mla, umac, phi, rho, dx, the_bc_tower[, macrhs][, beta0_1d,
beta0_1d_edge][, beta0_cart_edge])`` (Source/macproject.f90:51-76), over the
per-box arrays of :mod:`maestro_amd.boxes`:

* ``Layout`` plays the role of ``mla`` + ``dx`` + ``the_bc_tower``: the boxes,
  the problem domain, the cell size and the physical BC codes.
* :func:`macproject` takes per-box ``umac`` (list of ``dim`` arrays per box),
  ``phi`` and ``rho`` (lists of arrays), modifies ``umac`` and ``phi`` in
  place and returns the solver statistics.  Invalid option combinations raise
  :class:`MacError` as the reference calls ``bl_error`` (:99-109).

Arrays may be numpy (host) or torch tensors on the GPU; a call uses one kind
throughout.  There is no CPU fallback: a missing library raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmaestro_mac.so")

PERIODIC, INTERIOR, INLET, OUTLET, SYMMETRY, SLIP_WALL, NO_SLIP_WALL = -1, 0, 11, 12, 13, 14, 15
PREC_FP64, PREC_MIXED, PREC_MIXED_BF16 = 0, 1, 2
CYCLE_F, CYCLE_W, CYCLE_V = 1, 2, 3                # mg_cycle_type (Source/_parameters:376-377)
BOTTOM_SMOOTH, BOTTOM_BICGSTAB, BOTTOM_CG = 0, 1, 2  # mg_bottom_solver (Docs/mg/mg.tex:239-252)
ELL_PER, ELL_DIR, ELL_NEU = -1, 1, 2               # elliptic boundary kinds (mac_solve)
COMP_PRESS, COMP_SCALAR, COMP_EXTRAP = 0, 1, 2     # component classes of mac_ell_bc
ERRORS = {1: "invalid argument or option combination", 2: "bad box layout",
          3: "non-finite input", 4: "not converged", 5: "HIP runtime error", 6: "out of memory",
          7: "the call needs more than the fixed scratch of mac_project_noalloc"}
SYMBOLS = ("mac_default_opts", "mac_plan_create", "mac_plan_destroy", "mac_plan_set_opts",
           "mac_project", "mac_project_noalloc", "mac_project_level", "mac_adjoint_level", "mac_version",
           "mac_cell_to_edge", "mac_make_macrhs", "mac_make_s0mac", "mac_solve", "mac_ell_bc",
           "mac_debug_level_dims", "mac_debug_get", "mac_debug_set", "mac_debug_op",
           "mac_plan_profile", "mac_plan_profile_read", "mac_ml_plan_create", "mac_ml_plan_destroy",
           "mac_ml_project", "mac_ml_plan_set_opts")
PROF = {"gsrb64": 0, "gsrb32": 1, "resid64": 2, "setup": 3, "correct": 4, "coarse": 5,
        "gsrb64_zero": 6, "gsrb32_zero": 7, "restrict64": 8, "restrict32": 9, "prolong64": 10,
        "prolong32": 11, "ir_update": 12}
ARR = {"phi": 0, "rhs": 1, "bx": 2, "by": 3, "bz": 4, "bxh": 5, "byh": 6, "bzh": 7, "res": 8}
ARR_F32, ARR_BF16 = 0x10, 0x20   # debug_get(level, what, copy="f32" | "bf16")
OP = {"gsrb": 0, "residual": 1, "restrict": 2, "prolong": 3, "coarsen": 4, "vcycle": 5,
      "resid_restrict": 6, "prolong_sweep": 7, "gsrb32": 8, "gsrb32_rho": 9}


class MacError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        super().__init__(f"{what}: {ERRORS.get(code, 'error')} (code {code})")


class CLayout(C.Structure):
    _fields_ = [("dim", C.c_int), ("nboxes", C.c_int), ("box_lo", C.POINTER(C.c_int)),
                ("box_hi", C.POINTER(C.c_int)), ("domain_lo", C.c_int * 3),
                ("domain_hi", C.c_int * 3), ("dx", C.c_double * 3),
                ("phys_bc", (C.c_int * 2) * 3), ("ng_umac", C.c_int), ("ng_phi", C.c_int),
                ("ng_rho", C.c_int), ("ng_macrhs", C.c_int), ("ng_beta0_cart", C.c_int)]


class COpts(C.Structure):
    _fields_ = [("eps_rel", C.c_double), ("eps_abs", C.c_double), ("max_iter", C.c_int),
                ("nu1", C.c_int), ("nu2", C.c_int), ("nub", C.c_int),
                ("bottom_eps", C.c_double), ("bottom_max_iter", C.c_int),
                ("max_levels", C.c_int), ("use_alt_energy_fix", C.c_int),
                ("precision", C.c_int), ("inner_cycles", C.c_int), ("device_ptrs", C.c_int),
                ("verbose", C.c_int), ("cycle_type", C.c_int), ("bottom_solver", C.c_int),
                ("omega", C.c_double)]


class CSphr(C.Structure):
    """mac_sphr: spherical base-state geometry."""
    _fields_ = [("dr", C.c_double), ("nr", C.c_int), ("center", C.c_double * 3),
                ("prob_lo", C.c_double * 3), ("interp", C.c_int)]


class CMacrhsIn(C.Structure):
    """mac_macrhs_in: the inputs of make_macrhs."""
    _fields_ = [("S_cc", C.c_void_p), ("ng_S", C.c_int),
                ("delta_gamma1_term", C.c_void_p), ("ng_dg", C.c_int),
                ("delta_p_term", C.c_void_p), ("ng_dp", C.c_int),
                ("delta_chi", C.c_void_p), ("ng_dchi", C.c_int),
                ("Sbar", C.c_void_p), ("beta0", C.c_void_p), ("gamma1bar", C.c_void_p),
                ("p0", C.c_void_p), ("rho0", C.c_void_p),
                ("dt", C.c_double), ("dpdt_factor", C.c_double),
                ("base_cutoff_coord", C.c_int), ("base_cutoff_density", C.c_double),
                ("is_predictor", C.c_int)]


class CSolveIn(C.Structure):
    """mac_solve_in: the arguments of mac_multigrid (per-box tables)."""
    _fields_ = [("rh", C.c_void_p), ("ng_rh", C.c_int), ("phi", C.c_void_p), ("ng_phi", C.c_int),
                ("alpha", C.c_void_p), ("ng_alpha", C.c_int), ("beta", C.c_void_p), ("ng_beta", C.c_int),
                ("ell_bc", (C.c_int * 2) * 3), ("eps_rel", C.c_double), ("eps_abs", C.c_double)]


def sphr(dr, nr, center, prob_lo=(0.0, 0.0, 0.0), interp=3) -> CSphr:
    g = CSphr()
    g.dr, g.nr, g.interp = float(dr), int(nr), int(interp)
    for d in range(3):
        g.center[d] = float(center[d])
        g.prob_lo[d] = float(prob_lo[d])
    return g


class CStats(C.Structure):
    _fields_ = [("iters", C.c_int), ("inner_cycles", C.c_int), ("converged", C.c_int),
                ("nlevels", C.c_int), ("rhs_norm", C.c_double), ("res0", C.c_double),
                ("res", C.c_double), ("t_h2d_ms", C.c_double), ("t_setup_ms", C.c_double),
                ("t_solve_ms", C.c_double), ("t_correct_ms", C.c_double),
                ("t_d2h_ms", C.c_double), ("t_total_ms", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_lib = None


def lib():
    """Load libmaestro_mac.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with maestro_amd.build.build()")
        L = C.CDLL(LIB_PATH)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.c_int
        L.mac_default_opts.argtypes = [C.POINTER(COpts)]
        L.mac_plan_create.argtypes = [C.POINTER(CLayout), C.POINTER(COpts), C.POINTER(vp)]
        L.mac_plan_destroy.argtypes = [vp]
        L.mac_plan_set_opts.argtypes = [vp, C.POINTER(COpts)]
        L.mac_project.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, C.POINTER(CStats)]
        L.mac_project_noalloc.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, C.POINTER(CStats)]
        L.mac_project_level.argtypes = [vp] * 9 + [C.POINTER(CStats)]
        L.mac_adjoint_level.argtypes = [vp] * 16 + [C.POINTER(CStats)]
        L.mac_version.restype = C.c_char_p
        L.mac_cell_to_edge.argtypes = [vp, vp, ip, vp]
        L.mac_make_macrhs.argtypes = [vp, vp, C.POINTER(CMacrhsIn), C.POINTER(CSphr)]
        L.mac_make_s0mac.argtypes = [vp, vp, C.POINTER(CSphr), vp]
        L.mac_solve.argtypes = [vp, C.POINTER(CSolveIn), C.POINTER(CStats)]
        L.mac_ell_bc.argtypes = [ip, ip]
        L.mac_debug_level_dims.argtypes = [vp, ip, C.POINTER(C.c_int)]
        L.mac_debug_get.argtypes = [vp, ip, ip, vp]
        L.mac_debug_set.argtypes = [vp, ip, ip, vp]
        L.mac_debug_op.argtypes = [vp, ip, ip, ip]
        L.mac_ml_plan_create.argtypes = [ip, C.POINTER(CLayout), C.POINTER(COpts), C.POINTER(vp)]
        L.mac_ml_plan_destroy.argtypes = [vp]
        L.mac_ml_plan_set_opts.argtypes = [vp, C.POINTER(COpts)]
        L.mac_ml_project.argtypes = [vp, vp, vp, vp, vp, vp, vp, C.POINTER(CStats)]
        L.mac_plan_profile.argtypes = [vp, ip]
        L.mac_plan_profile_read.argtypes = [vp, ip, C.POINTER(C.c_long), C.POINTER(C.c_double)]
        for f in ("mac_debug_level_dims", "mac_debug_get", "mac_debug_set", "mac_debug_op",
                  "mac_plan_profile", "mac_plan_profile_read", "mac_ml_plan_create", "mac_ml_plan_destroy",
           "mac_ml_project", "mac_ml_plan_set_opts"):
            getattr(L, f).restype = ip
        for f in ("mac_default_opts", "mac_plan_create", "mac_plan_destroy", "mac_plan_set_opts",
                  "mac_project", "mac_project_level", "mac_adjoint_level", "mac_cell_to_edge",
                  "mac_make_macrhs", "mac_make_s0mac", "mac_solve", "mac_ell_bc",
                  "mac_ml_plan_create", "mac_ml_plan_destroy", "mac_ml_project", "mac_ml_plan_set_opts"):
            getattr(L, f).restype = ip
        _lib = L
    return _lib


def default_opts(**kw) -> COpts:
    o = COpts()
    lib().mac_default_opts(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


@dataclass
class Layout:
    """Boxes, domain, cell size and physical BCs of one level (``mla`` +
    ``dx`` + ``the_bc_tower`` of the reference call)."""
    dim: int
    boxes: list            # [(lo, hi)] inclusive cell indices
    domain_lo: tuple
    domain_hi: tuple
    dx: tuple
    phys_bc: list          # [[lo, hi]] per direction
    ng_umac: int = 1
    ng_phi: int = 1
    ng_rho: int = 1
    ng_macrhs: int = 0
    ng_beta0_cart: int = 1
    _keep: list = field(default_factory=list, repr=False)

    def to_c(self) -> CLayout:
        nb = len(self.boxes)
        lo = (C.c_int * (3 * nb))()
        hi = (C.c_int * (3 * nb))()
        for b, (l, h) in enumerate(self.boxes):
            for d in range(3):
                lo[3 * b + d] = int(l[d]) if d < self.dim else 0
                hi[3 * b + d] = int(h[d]) if d < self.dim else 0
        self._keep[:] = [lo, hi]
        c = CLayout()
        c.dim, c.nboxes = self.dim, nb
        c.box_lo = C.cast(lo, C.POINTER(C.c_int))
        c.box_hi = C.cast(hi, C.POINTER(C.c_int))
        for d in range(3):
            c.domain_lo[d] = int(self.domain_lo[d]) if d < self.dim else 0
            c.domain_hi[d] = int(self.domain_hi[d]) if d < self.dim else 0
            c.dx[d] = float(self.dx[d]) if d < self.dim else 1.0
            for s in range(2):
                c.phys_bc[d][s] = int(self.phys_bc[d][s]) if d < self.dim else PERIODIC
        c.ng_umac, c.ng_phi, c.ng_rho = self.ng_umac, self.ng_phi, self.ng_rho
        c.ng_macrhs, c.ng_beta0_cart = self.ng_macrhs, self.ng_beta0_cart
        return c

    @staticmethod
    def uniform(n, box, dx, phys_bc, **ng):
        from .workloads import box_list
        dim = len(n)
        return Layout(dim=dim, boxes=box_list(n, box), domain_lo=(0,) * dim,
                      domain_hi=tuple(x - 1 for x in n), dx=tuple(dx), phys_bc=phys_bc, **ng)


def _is_torch(x):
    return type(x).__module__.startswith("torch")


def _ptr(x):
    if x is None:
        return None
    if _is_torch(x):
        if not x.is_contiguous() or x.dtype.itemsize != 8:
            raise ValueError("device arrays must be contiguous float64 tensors")
        return C.c_void_p(x.data_ptr())
    if not (isinstance(x, np.ndarray) and x.dtype == np.float64 and x.flags.c_contiguous):
        raise ValueError("host arrays must be C-contiguous float64 numpy arrays")
    return C.c_void_p(x.ctypes.data)


def _ptr_array(xs):
    if xs is None:
        return None
    if isinstance(xs, PtrTable):
        return xs.arr
    arr = (C.c_void_p * len(xs))(*[_ptr(x).value for x in xs])
    return arr


class PtrTable:
    """The data pointers of a list of arrays (all torch device tensors or all
    numpy host arrays), checked once and kept with the arrays they point
    into.  A caller that passes the same per-box arrays on every call (a
    MultiFab's boxes, step after step) builds the table once: building it
    from some thousands of Python array objects costs about as much as a
    millisecond per call.  Accepted wherever ``project`` takes a list of
    arrays; for ``umac`` it holds the flattened per-box list (box-major,
    ``dim`` faces per box)."""

    def __init__(self, xs):
        self.items = list(xs)
        self.device = bool(self.items) and _is_torch(self.items[0])
        if any(_is_torch(x) != self.device for x in self.items):
            raise ValueError("a pointer table holds device arrays or host arrays, not both")
        self.arr = _ptr_array(self.items)

    def __len__(self):
        return len(self.items)


class MacPlan:
    """A plan: validated layout + device MG hierarchy (``mac_plan_create``)."""

    def __init__(self, layout: Layout, opts: COpts | None = None, **kw):
        self.layout = layout
        self.opts = opts or default_opts(**kw)
        self._c_layout = layout.to_c()
        h = C.c_void_p()
        rc = lib().mac_plan_create(C.byref(self._c_layout), C.byref(self.opts), C.byref(h))
        if rc:
            raise MacError(rc, "mac_plan_create")
        self._h = h

    def set_opts(self, **kw):
        for k, v in kw.items():
            setattr(self.opts, k, v)
        rc = lib().mac_plan_set_opts(self._h, C.byref(self.opts))
        if rc:
            raise MacError(rc, "mac_plan_set_opts")

    def close(self):
        if getattr(self, "_h", None):
            lib().mac_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _sync_torch(self, xs):
        if any(_is_torch(x) for x in xs if x is not None):
            import torch
            torch.cuda.current_stream().synchronize()

    def project(self, umac, phi, rho, macrhs=None, beta0_1d=None, beta0_1d_edge=None,
                beta0_cart_edge=None, check=True, noalloc=False):
        """macproject on per-box arrays; ``umac`` is a list (per box) of ``dim``
        face arrays.  Returns the stats dict.  noalloc: through
        mac_project_noalloc (device arrays; the plan's solver storage for its
        precision already allocated by an earlier call)."""
        dim = self.layout.dim
        nb = len(self.layout.boxes)

        def table(xs, per_box, name):
            if xs is None:
                return None
            if isinstance(xs, PtrTable):
                t = xs
            elif per_box == 1:
                t = list(xs)
            else:
                if any(len(box) != per_box for box in xs):
                    raise ValueError(f"{name}: {per_box} arrays per box")
                t = [a for box in xs for a in box]
            if len(t) != nb * per_box:
                raise ValueError(f"{name}: {nb * per_box} arrays for {nb} boxes, got {len(t)}")
            return t

        flat_u = table(umac, dim, "umac")
        phi_t, rho_t = table(phi, 1, "phi"), table(rho, 1, "rho")
        rhs_t, flat_c = table(macrhs, 1, "macrhs"), table(beta0_cart_edge, dim, "beta0_cart_edge")
        # one kind of array for the whole call: the library reads every pointer
        # as a device pointer or every one as a host pointer
        mixed = ValueError("mac_project: all arrays must be device tensors or all host arrays")

        def kind(t):
            if isinstance(t, PtrTable):
                return t.device
            if hasattr(t, "shape"):
                return _is_torch(t)
            ks = {_is_torch(x) for x in t}
            if len(ks) != 1:
                raise mixed
            return ks.pop()

        dev = kind(flat_u)
        others = [t for t in (phi_t, rho_t, rhs_t, flat_c) if t is not None and len(t)]
        others += [x for x in (beta0_1d, beta0_1d_edge) if x is not None]
        if any(kind(t) != dev for t in others):
            raise mixed
        if bool(self.opts.device_ptrs) != dev:
            self.set_opts(device_ptrs=int(dev))
        if dev:
            self._sync_torch([flat_u.items[0] if isinstance(flat_u, PtrTable) else flat_u[0]])
        keep = [_ptr_array(flat_u), _ptr_array(phi_t), _ptr_array(rho_t),
                _ptr_array(rhs_t), _ptr_array(flat_c)]
        b0 = _ptr(beta0_1d) if beta0_1d is not None else None
        b0e = _ptr(beta0_1d_edge) if beta0_1d_edge is not None else None
        st = CStats()
        entry = lib().mac_project_noalloc if noalloc else lib().mac_project
        rc = entry(self._h, keep[0], keep[1], keep[2], keep[3], b0, b0e, keep[4], C.byref(st))
        if rc and (check or rc != 4):
            raise MacError(rc, "mac_project")
        d = st.as_dict()
        d["rc"] = rc
        return d

    # ---- producers of the projection's inputs
    def _mode(self, xs):
        dev = any(_is_torch(x) for x in xs if x is not None)
        if bool(self.opts.device_ptrs) != dev:
            self.set_opts(device_ptrs=int(dev))
        self._sync_torch(list(xs))

    def cell_to_edge(self, s0_cell, s0_edge):
        """cell_to_edge into s0_edge (nr + 1 values)."""
        self._mode([s0_cell, s0_edge])
        rc = lib().mac_cell_to_edge(self._h, _ptr(s0_cell), int(s0_cell.shape[0]), _ptr(s0_edge))
        if rc:
            raise MacError(rc, "mac_cell_to_edge")

    def make_macrhs(self, macrhs, S_cc, delta_gamma1_term, Sbar, beta0, ng_S=0, ng_dg=0,
                    delta_p_term=None, ng_dp=0, delta_chi=None, ng_dchi=0, gamma1bar=None, p0=None,
                    rho0=None, dt=1.0, dpdt_factor=0.0, base_cutoff_coord=0, base_cutoff_density=0.0,
                    is_predictor=True, sph=None):
        """make_macrhs into the per-box arrays macrhs (lists of per-box
        arrays; 1-D profiles as arrays)."""
        flat = list(macrhs) + list(S_cc) + list(delta_gamma1_term) + list(delta_p_term or []) + \
            list(delta_chi or []) + [Sbar, beta0, gamma1bar, p0, rho0]
        self._mode(flat)
        keep = [_ptr_array(list(macrhs)), _ptr_array(list(S_cc)), _ptr_array(list(delta_gamma1_term)),
                _ptr_array(delta_p_term), _ptr_array(delta_chi)]
        a = CMacrhsIn()
        a.S_cc, a.ng_S = C.cast(keep[1], C.c_void_p), ng_S
        a.delta_gamma1_term, a.ng_dg = C.cast(keep[2], C.c_void_p), ng_dg
        a.delta_p_term, a.ng_dp = (C.cast(keep[3], C.c_void_p) if keep[3] is not None else None), ng_dp
        a.delta_chi, a.ng_dchi = (C.cast(keep[4], C.c_void_p) if keep[4] is not None else None), ng_dchi
        for f, x in (("Sbar", Sbar), ("beta0", beta0), ("gamma1bar", gamma1bar), ("p0", p0), ("rho0", rho0)):
            setattr(a, f, _ptr(x).value if x is not None else None)
        a.dt, a.dpdt_factor = float(dt), float(dpdt_factor)
        a.base_cutoff_coord, a.base_cutoff_density = int(base_cutoff_coord), float(base_cutoff_density)
        a.is_predictor = int(bool(is_predictor))
        rc = lib().mac_make_macrhs(self._h, keep[0], C.byref(a), C.byref(sph) if sph is not None else None)
        if rc:
            raise MacError(rc, "mac_make_macrhs")

    def make_s0mac(self, s0, sph, s0mac):
        """make_s0mac into s0mac: per box, three face arrays (one ghost layer)."""
        flat = [f for box in s0mac for f in box]
        self._mode(flat + [s0])
        tab = _ptr_array(flat)
        rc = lib().mac_make_s0mac(self._h, _ptr(s0), C.byref(sph), tab)
        if rc:
            raise MacError(rc, "mac_make_s0mac")

    # ---- solver-level entry (mac_multigrid with alpha)
    def solve(self, rh, phi, beta, alpha=None, ell_bc=None, comp=COMP_PRESS, eps_rel=1e-10, eps_abs=-1.0,
              ng_rh=0, ng_phi=None, ng_alpha=0, ng_beta=0, check=True):
        """mac_multigrid: solve (alpha - D beta G) phi = rh on per-box arrays
        (``beta``: per box, ``dim`` face arrays; ``alpha`` None for 0).  The
        elliptic BCs are ``ell_bc`` ([[lo, hi]] per direction, ELL_*) or those
        of component class ``comp`` for the layout's physical BCs.  ``phi``
        (warm start; Dirichlet values in its ghost cells) receives the
        solution.  Returns the stats dict."""
        lay = self.layout
        dim = lay.dim
        if ng_phi is None:
            ng_phi = lay.ng_phi
        if ell_bc is None:
            ell_bc = [[ell_bc_of(lay.phys_bc[d][s], comp) for s in range(2)] for d in range(dim)]
        flat_b = [f for box in beta for f in box]
        assert all(len(box) == dim for box in beta)
        self._mode(list(rh) + list(phi) + flat_b + list(alpha or []))
        keep = [_ptr_array(list(rh)), _ptr_array(list(phi)), _ptr_array(list(alpha) if alpha else None),
                _ptr_array(flat_b)]
        a = CSolveIn()
        a.rh, a.ng_rh = C.cast(keep[0], C.c_void_p), ng_rh
        a.phi, a.ng_phi = C.cast(keep[1], C.c_void_p), ng_phi
        a.alpha, a.ng_alpha = (C.cast(keep[2], C.c_void_p) if keep[2] is not None else None), ng_alpha
        a.beta, a.ng_beta = C.cast(keep[3], C.c_void_p), ng_beta
        for d in range(3):
            for s in range(2):
                a.ell_bc[d][s] = int(ell_bc[d][s]) if d < dim else ELL_PER
        a.eps_rel, a.eps_abs = float(eps_rel), float(eps_abs)
        st = CStats()
        rc = lib().mac_solve(self._h, C.byref(a), C.byref(st))
        if rc and (check or rc != 4):
            raise MacError(rc, "mac_solve")
        d = st.as_dict()
        d["rc"] = rc
        return d

    # ---- kernel timing on the plan's stream
    def profile(self, enable=True):
        rc = lib().mac_plan_profile(self._h, int(enable))
        if rc:
            raise MacError(rc, "mac_plan_profile")

    def profile_read(self):
        out = {}
        for name, cls in PROF.items():
            cnt, ms = C.c_long(), C.c_double()
            rc = lib().mac_plan_profile_read(self._h, cls, C.byref(cnt), C.byref(ms))
            if rc:
                raise MacError(rc, "mac_plan_profile_read")
            out[name] = (cnt.value, ms.value)
        return out

    # ---- diagnostics: internal FP64 level arrays and single operators
    def level_dims(self, level):
        d = (C.c_int * 3)()
        rc = lib().mac_debug_level_dims(self._h, level, d)
        if rc:
            raise MacError(rc, "mac_debug_level_dims")
        return tuple(d)

    def _arr_size(self, level, what):
        nx, ny, nz = self.level_dims(level)
        return {"bxh": ny * nz, "byh": nx * nz, "bzh": nx * ny}.get(what, nx * ny * nz)

    def debug_get(self, level, what, copy=None):
        """An internal level array (FP64); copy="f32" / "bf16": the FP32
        hierarchy's / the compact level-0 copy, widened to FP64."""
        out = np.zeros(self._arr_size(level, what))
        code = ARR[what] | {None: 0, "f32": ARR_F32, "bf16": ARR_BF16}[copy]
        rc = lib().mac_debug_get(self._h, level, code, C.c_void_p(out.ctypes.data))
        if rc:
            raise MacError(rc, "mac_debug_get")
        return out

    def debug_set(self, level, what, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        assert a.size == self._arr_size(level, what)
        rc = lib().mac_debug_set(self._h, level, ARR[what], C.c_void_p(a.ctypes.data))
        if rc:
            raise MacError(rc, "mac_debug_set")

    def debug_op(self, level, op, count=1):
        rc = lib().mac_debug_op(self._h, level, OP[op], count)
        if rc:
            raise MacError(rc, "mac_debug_op")

    def project_level(self, ux, uy, uz, phi, bx, by, bz, macrhs=None, check=True):
        """Whole-domain device entry (natural face layout, torch tensors)."""
        if not self.opts.device_ptrs:
            self.set_opts(device_ptrs=1)
        self._sync_torch([ux])
        st = CStats()
        rc = lib().mac_project_level(self._h, _ptr(ux), _ptr(uy), _ptr(uz), _ptr(phi), _ptr(bx),
                                     _ptr(by), _ptr(bz), _ptr(macrhs), C.byref(st))
        if rc and (check or rc != 4):
            raise MacError(rc, "mac_project_level")
        d = st.as_dict()
        d["rc"] = rc
        return d

    def adjoint_level(self, phi, bx, by, bz, gphi, gux, guy, guz, lam, oux, ouy, ouz, obx, oby,
                      obz, check=True):
        if not self.opts.device_ptrs:
            self.set_opts(device_ptrs=1)
        self._sync_torch([phi])
        st = CStats()
        args = [_ptr(a) for a in (phi, bx, by, bz, gphi, gux, guy, guz, lam, oux, ouy, ouz, obx,
                                  oby, obz)]
        rc = lib().mac_adjoint_level(self._h, *args, C.byref(st))
        if rc and (check or rc != 4):
            raise MacError(rc, "mac_adjoint_level")
        d = st.as_dict()
        d["rc"] = rc
        return d


class MacMLPlan:
    """Two-level plan (``mac_ml_plan_create``): ``layouts`` = [coarse, fine],
    the fine level's boxes disjoint and each of whole coarse cells.  The
    arrays of ``project`` are all device tensors or all host numpy arrays
    (per level, per box)."""

    def __init__(self, layouts, opts: COpts | None = None, **kw):
        if len(layouts) != 2:
            raise ValueError("two levels")
        self.layouts = list(layouts)
        # the relative tolerance of a two-level projection: min(eps_mac *
        # 10^(nlevs-1), eps_mac_max) = 1e-9 (Source/macproject.f90:217)
        self.opts = opts or default_opts(**{"eps_rel": 1e-9, **kw})
        self.opts.device_ptrs = 1
        self._c = (CLayout * 2)(*[l.to_c() for l in self.layouts])
        h = C.c_void_p()
        rc = lib().mac_ml_plan_create(2, self._c, C.byref(self.opts), C.byref(h))
        if rc:
            raise MacError(rc, "mac_ml_plan_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().mac_ml_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def project(self, umac, phi, rho, macrhs=None, beta0_1d=None, beta0_1d_edge=None, check=True):
        """macproject on two levels.  Per level l: umac[l] a list per box of
        ``dim`` face tensors (or a PtrTable of the flattened list), phi[l],
        rho[l] and macrhs[l] per-box lists, beta0_1d[l] / beta0_1d_edge[l]
        level l's planar profile.  Returns the stats dict."""
        dim = self.layouts[0].dim
        keep = []
        kinds = set()

        def table(xs, l, per_box, name):
            nb = len(self.layouts[l].boxes)
            if isinstance(xs, PtrTable):
                t = xs
            elif per_box == 1:
                t = list(xs)
            else:
                if any(len(box) != per_box for box in xs):
                    raise ValueError(f"{name}[{l}]: {per_box} arrays per box")
                t = [a for box in xs for a in box]
            if len(t) != nb * per_box:
                raise ValueError(f"{name}[{l}]: {nb * per_box} arrays for {nb} boxes, got {len(t)}")
            items = t.items if isinstance(t, PtrTable) else t
            kinds.update(_is_torch(x) for x in items)
            a = _ptr_array(t)
            keep.append(a)
            return C.cast(a, C.c_void_p).value

        def per_level(xs, per_box, name):
            if xs is None:
                return None
            arr = (C.c_void_p * 2)(*[table(xs[l], l, per_box, name) for l in range(2)])
            keep.append(arr)
            return arr

        def vec(xs):
            if xs is None:
                return None
            kinds.update(_is_torch(x) for x in xs)
            arr = (C.c_void_p * 2)(*[_ptr(x).value for x in xs])
            keep.append(arr)
            return arr

        u = per_level(umac, dim, "umac")
        p = per_level(phi, 1, "phi")
        r = per_level(rho, 1, "rho")
        m = per_level(macrhs, 1, "macrhs")
        b0, b0e = vec(beta0_1d), vec(beta0_1d_edge)
        if len(kinds) != 1:
            raise ValueError("mac_ml_project: all arrays must be device tensors or all host arrays")
        dev = kinds.pop()
        if bool(self.opts.device_ptrs) != dev:
            self.opts.device_ptrs = int(dev)
            rc = lib().mac_ml_plan_set_opts(self._h, C.byref(self.opts))
            if rc:
                raise MacError(rc, "mac_ml_plan_set_opts")
        if dev:
            import torch
            torch.cuda.current_stream().synchronize()
        st = CStats()
        rc = lib().mac_ml_project(self._h, u, p, r, m, b0, b0e, C.byref(st))
        if rc and (check or rc != 4):
            raise MacError(rc, "mac_ml_project")
        d = st.as_dict()
        d["rc"] = rc
        return d


def ell_bc_of(phys, comp=COMP_PRESS) -> int:
    """Elliptic kind of a physical boundary code for a component class
    (ell_bc_level_build, Source/define_bc_tower.f90:296-402)."""
    return lib().mac_ell_bc(int(phys), int(comp))


def macproject(layout: Layout, umac, phi, rho, macrhs=None, beta0_1d=None, beta0_1d_edge=None,
               beta0_cart_edge=None, opts: COpts | None = None, plan: MacPlan | None = None, **kw):
    """``macproject`` (Source/macproject.f90:51-328) on per-box arrays.
    ``umac`` and ``phi`` are updated in place; returns the solver stats."""
    if beta0_1d is not None and beta0_cart_edge is not None:
        raise MacError(1, "macproject: CANT HAVE 1D and 3D DIV_COEFF IN MACPROJECT")
    p = plan or MacPlan(layout, opts, **kw)
    try:
        return p.project(umac, phi, rho, macrhs, beta0_1d, beta0_1d_edge, beta0_cart_edge)
    finally:
        if plan is None:
            p.close()
