"""Write the golden fixtures of the MAC projection (run here, committed).

Each fixture is an independent construction of the discrete problem that
Source/macproject.f90 hands to the solver, built from its definitions with
scipy.sparse instead of stencil loops:

  beta_f  = 2 / (rho_L + rho_R)               (mk_mac_coeffs, :698-738)
            times beta0 on the face, twice under use_alt_energy_fix
            (mult_edge_by_1d_coeff :577-609; :185-203)
  U'      = beta0_face * U                    (:154-158)
  rh      = macrhs - D U'   or  -D U'         (divumac :457-480, :389-398)
  A       = -D B G   (D: faces -> cells divergence, G: cells -> faces
            gradient with periodic wrap / zero-flux walls / phi = 0 on an
            outlet face, B = diag(beta_f)); solve A phi = rh directly (with
            the mean fixed to zero on singular problems)
  U_proj  = (U' - B G phi) / beta0_face        (mkumac :908-975, :232-236)

Usage: python tests/golden/make_golden.py   (writes tests/golden/*.npz)
"""
from __future__ import annotations

import os
import sys

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from maestro_amd import workloads as W  # noqa: E402

P, WALL, OUT = W.PERIODIC, W.SLIP_WALL, W.OUTLET


def cidx(n, i, j, k):
    return i + n[0] * (j + n[1] * k)


def face_count(n, d):
    m = list(n)
    m[d] += 1
    return m[0] * m[1] * m[2], m


def build_ops(n, dx, bc, dim):
    """Sparse divergence D (cells x faces, per direction blocks) and gradient
    G (faces x cells) over natural face layouts; also the boundary mask."""
    nc = n[0] * n[1] * n[2]
    Ds, Gs = [], []
    for d in range(dim):
        nf, m = face_count(n, d)
        Dr, Dc, Dv = [], [], []
        Gr, Gc, Gv = [], [], []
        for k in range(m[2]):
            for j in range(m[1]):
                for i in range(m[0]):
                    f = i + m[0] * (j + m[1] * k)
                    c = [i, j, k]
                    lo = list(c)
                    lo[d] -= 1
                    # divergence: face f is the high face of lo and low face of c
                    if c[d] < n[d]:
                        Dr.append(cidx(n, *c)); Dc.append(f); Dv.append(-1.0 / dx[d])
                    if lo[d] >= 0:
                        Dr.append(cidx(n, *lo)); Dc.append(f); Dv.append(1.0 / dx[d])
                    # gradient (phi_hi - phi_lo)/dx with boundary rules
                    inside_hi = c[d] < n[d]
                    inside_lo = lo[d] >= 0
                    if inside_hi and inside_lo:
                        Gr += [f, f]; Gc += [cidx(n, *c), cidx(n, *lo)]
                        Gv += [1.0 / dx[d], -1.0 / dx[d]]
                    else:
                        side = 0 if not inside_lo else 1
                        if bc[d][side] == P:
                            hi = list(c); lw = list(lo)
                            hi[d] %= n[d]; lw[d] %= n[d]
                            Gr += [f, f]; Gc += [cidx(n, *hi), cidx(n, *lw)]
                            Gv += [1.0 / dx[d], -1.0 / dx[d]]
                        elif bc[d][side] == OUT:
                            # phi = 0 on the face: one-sided (phi_c - 0)/(dx/2)
                            if inside_hi:
                                Gr.append(f); Gc.append(cidx(n, *c)); Gv.append(2.0 / dx[d])
                            else:
                                Gr.append(f); Gc.append(cidx(n, *lo)); Gv.append(-2.0 / dx[d])
                        # walls: zero gradient (no entries)
        Ds.append(sp.csr_matrix((Dv, (Dr, Dc)), shape=(nc, nf)))
        Gs.append(sp.csr_matrix((Gv, (Gr, Gc)), shape=(nf, nc)))
    return Ds, Gs


def faces_natural(a, d, n):
    """whole-domain face array (nz,ny,nx+..) -> flat natural layout"""
    return np.ascontiguousarray(a).ravel()


def golden_case(name, cfg, n, alt, macrhs_seed=None):
    wl = W.make_workload(cfg, n=n, box=n[0] // 2)
    dim = wl["dim"]
    nn = list(wl["n"]) + [1] * (3 - dim)
    dx = list(wl["dx"]) + [1.0] * (3 - dim)
    bc = wl["bc"] + [[P, P]] * (3 - dim)
    rho = wl["rho"]
    gz = 1 if dim == 3 else 0
    rr = rho if dim == 3 else rho
    # coefficients (raw) from rho with ghosts
    beta = []
    for d in range(dim):
        ax = 2 - d
        r = rr
        hi = [slice(None)] * 3
        lo = [slice(None)] * 3
        # interior region in the other directions, both neighbours along d
        for e in range(3):
            ae = 2 - e
            if e == d:
                hi[ae] = slice(1, None) if (e < dim) else slice(None)
                lo[ae] = slice(0, -1) if (e < dim) else slice(None)
            else:
                g = 1 if (e < dim and (e < 2 or gz)) else 0
                hi[ae] = slice(g, r.shape[ae] - g) if g else slice(None)
                lo[ae] = hi[ae]
        b = 2.0 / (r[tuple(hi)] + r[tuple(lo)])
        beta.append(b)
    faces = [f.copy() for f in wl["faces"]]
    if wl["beta0"] is not None:
        rd = dim - 1
        for d in range(dim):
            if d == rd:
                c = wl["beta0_edge"]
            else:
                c = wl["beta0"]
            shape = [1, 1, 1]
            shape[2 - rd] = -1
            c = c.reshape(shape)
            faces[d] = faces[d] * c
            beta[d] = beta[d] * c
            if alt:
                beta[d] = beta[d] * c
    elif wl["cart"] is not None:
        for d in range(dim):
            faces[d] = faces[d] * wl["cart"][d]
            beta[d] = beta[d] * wl["cart"][d]
            if alt:
                beta[d] = beta[d] * wl["cart"][d]
    macrhs = None
    if macrhs_seed is not None:
        macrhs = np.random.default_rng(macrhs_seed).standard_normal((nn[2], nn[1], nn[0]))
        if all(b[0] != OUT for b in wl["bc"]):
            macrhs -= macrhs.mean()
    Ds, Gs = build_ops(nn, dx, bc, dim)
    DU = sum(Ds[d] @ faces[d].ravel() for d in range(dim))
    rh = (macrhs.ravel() - DU) if macrhs is not None else -DU
    A = -sum(Ds[d] @ sp.diags(beta[d].ravel()) @ Gs[d] for d in range(dim))
    singular = all(b[0] != OUT for b in wl["bc"])
    nc = nn[0] * nn[1] * nn[2]
    if singular:
        rhs = rh - rh.mean()
        ones = np.ones((nc, 1))
        K = sp.bmat([[A, sp.csr_matrix(ones)], [sp.csr_matrix(ones.T), None]]).tocsc()
        sol = spla.spsolve(K, np.concatenate([rhs, [0.0]]))
        phi = sol[:nc]
    else:
        phi = spla.spsolve(A.tocsc(), rh)
    proj = []
    for d in range(dim):
        u = faces[d].ravel() - beta[d].ravel() * (Gs[d] @ phi)
        proj.append(u.reshape(faces[d].shape))
    if wl["beta0"] is not None:
        rd = dim - 1
        for d in range(dim):
            c = wl["beta0_edge"] if d == rd else wl["beta0"]
            shape = [1, 1, 1]
            shape[2 - rd] = -1
            proj[d] = proj[d] / c.reshape(shape)
    elif wl["cart"] is not None:
        proj = [proj[d] / wl["cart"][d] for d in range(dim)]
    out = {"cfg": cfg, "n": np.array(wl["n"]), "alt": alt,
           "rh": rh.reshape(nn[2], nn[1], nn[0]),
           "phi": phi.reshape(nn[2], nn[1], nn[0])}
    for d in range(dim):
        out[f"beta{d}"] = beta[d]
        out[f"proj{d}"] = proj[d]
    if macrhs is not None:
        out["macrhs"] = macrhs
        out["macrhs_seed"] = macrhs_seed
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
    return out


CASES = [
    ("c0_2d_16_alt1", 0, (16, 16), 1, None),
    ("c0_2d_16_alt0_macrhs", 0, (16, 16), 0, 77),
    ("c1_3d_8_alt1", 1, (8, 8, 8), 1, None),
    ("c1_3d_8_alt0_macrhs", 1, (8, 8, 8), 0, 78),
    ("c3_3d_16_alt1", 3, (16, 16, 16), 1, None),
    ("c3_3d_8_alt0", 3, (8, 8, 8), 0, None),
]

if __name__ == "__main__":
    for name, cfg, n, alt, seed in CASES:
        o = golden_case(name, cfg, n, alt, seed)
        print(name, "phi norm", np.linalg.norm(o["phi"]))
