"""Seeded synthetic MAC-projection inputs (whole-domain arrays).

Every workload is returned in the whole-domain layout (C order, x fastest):
cells ``(nz, ny, nx)``; ``rho`` with one ghost layer ``(nz+2gz, ny+2, nx+2)``
(``gz = 0`` in 2D); x/y/z faces ``(nz,ny,nx+1)``, ``(nz,ny+1,nx)``,
``(nz+1,ny,nx)``.  :func:`split_boxes` cuts them into the per-box,
ghosted, Fortran-ordered arrays that ``macproject`` takes.

Input rules that keep the singular (periodic/wall) problems solvable, as in
the reference's own test (Exec/UNIT_TESTS/test_projection/test_projection.f90:
787-823): one value per physical face, equal periodic duplicates, and zero
wall-normal velocity.

Workloads (sizes are defaults; every one can be shrunk for tests):
  0  2D 128x128, x periodic, y slip walls, beta0 = exp(-y/H), H = L/4,
     rho = beta0 (1 + 0.01 G), G a Gaussian bubble of radius 0.1 at a seeded
     centre, U ~ U[-1, 1].
  1  3D 256^3, x/y periodic, z slip walls, isothermal beta0 = exp(-z/H) with
     exp(L/H) = 1e4, rho = beta0 (1 + 0.01 N(0,1)), U ~ U[-1, 1].
  2  workload 1 inputs (solved in mixed precision).
  3  3D 512^3 spherical star: beta0(r) = theta(r)^1.5 of the n = 1.5
     Lane-Emden polytrope with surface at r = 0.45 L, floored at 1e-6 of the
     centre; rho = beta0 (1 + 0.005 N(0,1)); U ~ U[-1,1] rho_face^-1/2;
     outlet on every face; beta0 put on faces as the average of the two
     cell values (make_s0mac type 1, Source/fill_3d_data.f90:1041-1075).
  4  workload 1 inputs plus loss weights w ~ N(0,1).
"""
from __future__ import annotations

import math

import numpy as np

PERIODIC, INTERIOR, INLET, OUTLET, SYMMETRY, SLIP_WALL, NO_SLIP_WALL = -1, 0, 11, 12, 13, 14, 15

DEFAULT_N = {0: (128, 128), 1: (256, 256, 256), 2: (256, 256, 256), 3: (512, 512, 512),
             4: (256, 256, 256)}
DEFAULT_BOX = {0: 32, 1: 64, 2: 64, 3: 64, 4: 64}


def cell_to_edge(b0: np.ndarray) -> np.ndarray:
    """beta0 on radial edges from cell values, Source/cell_to_edge.f90:36-61:
    end edges copy the cell value, the next ones in take the mean, elsewhere
    the limited fourth-order average."""
    nr = b0.shape[0]
    e = np.empty(nr + 1)
    for r in range(nr + 1):
        if r == 0:
            e[r] = b0[0]
        elif r == 1:
            e[r] = 0.5 * (b0[0] + b0[1])
        elif r == nr:
            e[r] = b0[nr - 1]
        elif r == nr - 1:
            e[r] = 0.5 * (b0[r] + b0[r - 1])
        else:
            tmp = 7.0 / 12.0 * (b0[r] + b0[r - 1]) - 1.0 / 12.0 * (b0[r + 1] + b0[r - 2])
            lo, hi = min(b0[r], b0[r - 1]), max(b0[r], b0[r - 1])
            e[r] = min(max(tmp, lo), hi)
    return e


def _fill_rho_ghosts(rho_in: np.ndarray, bc, dim: int) -> np.ndarray:
    """Ghost layer for rho: periodic image or first-order extrapolation
    (density uses FOEXTRAP/HOEXTRAP at walls and outlets,
    Source/define_bc_tower.f90:213-294)."""
    if dim == 2:
        pad = [(0, 0), (1, 1), (1, 1)]
        modes = {1: bc[1], 2: bc[0]}
    else:
        pad = [(1, 1), (1, 1), (1, 1)]
        modes = {0: bc[2], 1: bc[1], 2: bc[0]}
    r = np.pad(rho_in, pad, mode="edge")
    for ax, b in modes.items():
        if b[0] == PERIODIC:
            sl_g = [slice(None)] * 3
            sl_s = [slice(None)] * 3
            sl_g[ax] = 0
            sl_s[ax] = -2
            r[tuple(sl_g)] = r[tuple(sl_s)]
            sl_g[ax] = -1
            sl_s[ax] = 1
            r[tuple(sl_g)] = r[tuple(sl_s)]
    return r


def _random_faces(rng, n, dim, bc, scale=None):
    """U ~ U[-1,1] per physical face: periodic duplicates equal, wall-normal
    faces zero (Neumann pressure faces carry no normal flow)."""
    nx, ny, nz = n[0], n[1], (n[2] if dim == 3 else 1)
    faces = []
    shapes = [(nz, ny, nx + 1), (nz, ny + 1, nx), (nz + 1, ny, nx)]
    for d in range(dim):
        f = rng.uniform(-1.0, 1.0, size=shapes[d])
        ax = 2 - d
        sl_lo = [slice(None)] * 3
        sl_hi = [slice(None)] * 3
        sl_lo[ax] = 0
        sl_hi[ax] = -1
        if bc[d][0] == PERIODIC:
            f[tuple(sl_hi)] = f[tuple(sl_lo)]
        else:
            if bc[d][0] != OUTLET:
                f[tuple(sl_lo)] = 0.0
            if bc[d][1] != OUTLET:
                f[tuple(sl_hi)] = 0.0
        faces.append(f)
    if scale is not None:
        faces = [f * s for f, s in zip(faces, scale)]
    return faces


def lane_emden_theta(xi: np.ndarray, n: float = 1.5):
    """theta(xi) of the Lane-Emden equation (RK4 on a fine grid, then linear
    interpolation); returns (theta, xi_1) with theta(xi_1) = 0."""
    h = 1e-3
    x = [1e-6]
    y, z = 1.0 - x[0] ** 2 / 6.0, -x[0] / 3.0
    th = [y]
    while y > 0.0:
        def f(xx, yy, zz):
            return zz, -max(yy, 0.0) ** n - 2.0 * zz / xx
        k1 = f(x[-1], y, z)
        k2 = f(x[-1] + h / 2, y + h / 2 * k1[0], z + h / 2 * k1[1])
        k3 = f(x[-1] + h / 2, y + h / 2 * k2[0], z + h / 2 * k2[1])
        k4 = f(x[-1] + h, y + h * k3[0], z + h * k3[1])
        y = y + h / 6 * (k1[0] + 2 * k2[0] + 2 * k3[0] + k4[0])
        z = z + h / 6 * (k1[1] + 2 * k2[1] + 2 * k3[1] + k4[1])
        x.append(x[-1] + h)
        th.append(y)
    x = np.array(x)
    th = np.array(th)
    # surface by linear interpolation of the last step
    xi1 = x[-2] + (x[-1] - x[-2]) * th[-2] / (th[-2] - th[-1])
    return np.interp(xi, x, np.maximum(th, 0.0), right=0.0), xi1


def make_workload(cfg: int, n=None, seed=None, box=None):
    """Seeded inputs of workload ``cfg``.  Returns a dict with grid info
    (``dim, n, dx, bc``), ``faces`` (list), ``rho`` (ghosted), ``beta0``,
    ``beta0_edge`` or ``cart`` (face beta0), ``macrhs`` (None), ``phi``
    (zeros), ``box`` (box edge length) and, for workload 4, ``w``."""
    n = tuple(n or DEFAULT_N[cfg])
    seed = 1234 + cfg if seed is None else seed
    box = box or min(DEFAULT_BOX[cfg], *n)
    rng = np.random.default_rng(seed)
    dim = len(n)
    L = 1.0
    dx = [L / n[d] for d in range(dim)]
    out = {"cfg": cfg, "dim": dim, "n": n, "dx": dx, "box": box, "macrhs": None,
           "beta0": None, "beta0_edge": None, "cart": None}
    if cfg == 0:
        assert dim == 2
        bc = [[PERIODIC, PERIODIC], [SLIP_WALL, SLIP_WALL]]
        H = 0.25 * L
        yc = (np.arange(n[1]) + 0.5) * dx[1]
        beta0 = np.exp(-yc / H)
        cx, cy = rng.uniform(0.2, 0.8, size=2)
        xc = (np.arange(n[0]) + 0.5) * dx[0]
        X, Y = np.meshgrid(xc, yc, indexing="xy")
        G = np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / 0.1 ** 2)
        rho = (beta0[:, None] * (1.0 + 0.01 * G))[None]
        out["beta0"], out["beta0_edge"] = beta0, cell_to_edge(beta0)
        faces = _random_faces(rng, n, dim, bc)
    elif cfg in (1, 2, 4):
        assert dim == 3
        bc = [[PERIODIC, PERIODIC], [PERIODIC, PERIODIC], [SLIP_WALL, SLIP_WALL]]
        H = L / math.log(1e4)
        zc = (np.arange(n[2]) + 0.5) * dx[2]
        beta0 = np.exp(-zc / H)
        noise = rng.standard_normal(size=(n[2], n[1], n[0]))
        rho = beta0[:, None, None] * (1.0 + 0.01 * noise)
        out["beta0"], out["beta0_edge"] = beta0, cell_to_edge(beta0)
        faces = _random_faces(rng, n, dim, bc)
        if cfg == 4:
            out["w"] = rng.standard_normal(size=(n[2], n[1], n[0]))
    elif cfg == 3:
        assert dim == 3
        bc = [[OUTLET, OUTLET]] * 3
        c = [(np.arange(n[d]) + 0.5) * dx[d] - 0.5 * L for d in range(3)]
        Z, Y, X = np.meshgrid(c[2], c[1], c[0], indexing="ij")
        r = np.sqrt(X * X + Y * Y + Z * Z)
        R = 0.45 * L
        _, xi1 = lane_emden_theta(np.array([0.0]))
        th, _ = lane_emden_theta(r.ravel() / R * xi1)
        b0c = np.maximum(th.reshape(r.shape) ** 1.5, 1e-6)
        noise = rng.standard_normal(size=b0c.shape)
        rho = b0c * (1.0 + 0.005 * noise)
        # beta0 and rho on faces: average of the two cells (copy at the edge)
        cart = []
        rho_f = []
        for d in range(3):
            ax = 2 - d
            for src, dst in ((b0c, cart), (rho, rho_f)):
                p = np.concatenate([np.take(src, [0], axis=ax), src,
                                    np.take(src, [-1], axis=ax)], axis=ax)
                lo = [slice(None)] * 3
                hi = [slice(None)] * 3
                lo[ax] = slice(0, -1)
                hi[ax] = slice(1, None)
                dst.append(0.5 * (p[tuple(lo)] + p[tuple(hi)]))
        out["cart"] = cart
        faces = _random_faces(rng, n, dim, bc, scale=[f ** -0.5 for f in rho_f])
        del rho_f
    else:
        raise ValueError(f"unknown workload {cfg}")
    out["bc"] = bc
    out["rho"] = _fill_rho_ghosts(rho if rho.ndim == 3 else rho[None], bc, dim)
    out["faces"] = faces
    out["phi"] = np.zeros((n[2] if dim == 3 else 1, n[1], n[0]))
    return out


def box_list(n, box):
    """Boxes of edge ``box`` tiling the domain, x fastest (list of (lo, hi))."""
    dim = len(n)
    counts = [n[d] // box for d in range(dim)]
    assert all(counts[d] * box == n[d] for d in range(dim)), "box must divide the domain"
    boxes = []
    for idx in np.ndindex(*reversed(counts)):
        idx = tuple(reversed(idx))
        lo = [idx[d] * box for d in range(dim)]
        hi = [lo[d] + box - 1 for d in range(dim)]
        boxes.append((lo, hi))
    return boxes
