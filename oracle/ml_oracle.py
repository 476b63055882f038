"""Two-level (AMR) MAC projection: FP64 CPU restatement.

TEST INFRASTRUCTURE ONLY.  Imported by ``tests/`` as the checker of the
library's two-level projection; the product package never imports it.

What it restates (MAESTRO, two levels, refinement ratio 2):
  Source/macproject.f90:154-158   U <- beta0 U on every level
  Source/macproject.f90:357-361   divumac: ml_edge_restriction of U first
                                  (coarse faces under the fine level take the
                                  mean of the 2^(dim-1) fine faces; where the
                                  region meets a periodic seam, the face's
                                  periodic image too)
  Source/macproject.f90:389-404   rh = macrhs - D per level; ml_cc_restriction
                                  (covered coarse cells take the mean of their
                                  2^dim fine cells)
  Source/macproject.f90:178-193   beta = beta0 / rho (beta0^2 / rho, alt)
  Source/macproject.f90:217       rel_eps = min(eps_mac 10^(nlevs-1), eps_mac_max)
  Source/macproject.f90:227-229   the composite solve (FBoxLib ml_cc_solve)
  Source/macproject.f90:756-833   mkumac per level, then ml_edge_restriction
  Source/macproject.f90:229-240   U <- U / beta0
  Docs/mg/mg.tex:9-19,124-156     the composite operator: fine-grid stencils
                                  see coarse data through interpolated ghost
                                  values at coarse-fine (C-F) faces, and the
                                  coarse cell next to a C-F face takes the
                                  fine fluxes through that face (refluxing),
                                  "no correction step needed".

The interpolation FBoxLib uses at C-F faces is not in the reference tree
(FBoxLib is absent), so this restatement fixes its own, a standard
second-order one, and every consumer follows it:
  * the coarse value is moved to the fine ghost cell's transverse position
    with the coarse cell's linear slopes (central differences of the coarse
    neighbours along each transverse direction; one-sided where one
    neighbour lies beyond a non-periodic wall);
  * along the face normal, the quadratic through that coarse value (its
    centre 1 fine cell beyond the face), the first and the second fine cell
    inside gives the ghost value
        g = 8/15 C + 2/3 phi_0 - 1/5 phi_1        (Lagrange weights);
  * the fine flux through the face is beta_f (phi_0 - g) / h_f (outward
    sign by side); the coarse cell uses the mean of the 2^(dim-1) fine
    fluxes on the face.
Pinned by: exactness on fields of the class the interpolation reproduces
(the composite operator of x^2 along the C-F normal is -2, of the bilinear
x y is 0), conservation (the volume sum of A phi vanishes on a closed
domain), the sparse direct solution, the projection's own invariants and
the two-level test_projection known answers (Util/regression_testing/
Maestro-tests.ini:245-257, gr0_2d_2lev_coarse).  Parity with FBoxLib's own
interpolation, its ghost fill create_umac_grown and its treatment of the
singular mean stays unpinned (FBoxLib absent).

Arrays are in the oracle layout (C order (nz, ny, nx), nz = 1 in 2D); the
fine level's arrays cover the refined region only: its cells, its faces
(extent + 1 in the face direction) and rho with one ghost layer.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import oracle as O

W_COARSE, W_FINE0, W_FINE1 = 8.0 / 15.0, 2.0 / 3.0, -1.0 / 5.0


class TwoLevel:
    """Coarse domain of ``nc`` cells (spacing ``hc``, physical BCs ``bc``) and
    a fine level made of the boxes ``boxes`` (fine cell indices, inclusive
    (lo, hi) pairs, lo even and hi odd: each covers whole coarse cells; any
    union, disjoint boxes).  ``rlo``/``rhi`` alone give one box.  Fine
    arrays cover the boxes' bounding rectangle (``rlo``..``rhi``); its cells
    outside the boxes (holes) carry no unknown."""

    def __init__(self, dim, nc, hc, bc, rlo=None, rhi=None, boxes=None):
        self.dim = dim
        self.nc = [int(x) for x in nc] + [1] * (3 - len(nc))
        self.hc = [float(x) for x in hc] + [1.0] * (3 - len(hc))
        self.hf = [h / 2 for h in self.hc]
        self.bc = [list(bc[d]) if d < dim else [O.PERIODIC, O.PERIODIC] for d in range(3)]
        if boxes is None:
            boxes = [(list(rlo), list(rhi))]
        self.boxes = [([int(x) for x in lo] + [0] * (3 - len(lo)), [int(x) for x in hi] + [0] * (3 - len(hi)))
                      for lo, hi in boxes]
        self.rlo = [min(b[0][d] for b in self.boxes) for d in range(3)]
        self.rhi = [max(b[1][d] for b in self.boxes) for d in range(3)]
        self.nr = [self.rhi[d] - self.rlo[d] + 1 for d in range(3)]
        self.kind = [[O.lib().orc_ell_bc(self.bc[d][s]) for s in range(2)] for d in range(3)]
        for lo, hi in self.boxes:
            for d in range(dim):
                nf = 2 * self.nc[d]
                assert lo[d] % 2 == 0 and hi[d] % 2 == 1, "boxes must cover whole coarse cells"
                assert 0 <= lo[d] <= hi[d] < nf
        # fine cells of the boxes over the bounding rectangle, and the
        # covered coarse cells over the coarse domain
        self.mask = np.zeros((self.nr[2], self.nr[1], self.nr[0]), dtype=bool)
        self.cmask = np.zeros((self.nc[2], self.nc[1], self.nc[0]), dtype=bool)
        for lo, hi in self.boxes:
            sl = tuple(slice(lo[d] - self.rlo[d], hi[d] - self.rlo[d] + 1) for d in (2, 1, 0))
            assert not self.mask[sl].any(), "boxes overlap"
            self.mask[sl] = True
            self.cmask[tuple(slice(lo[d] // 2, hi[d] // 2 + 1) for d in (2, 1, 0))] = True
        self.ncc = self.nc[0] * self.nc[1] * self.nc[2]
        self.nfc = self.nr[0] * self.nr[1] * self.nr[2]

    @property
    def side(self):
        """Class of each side of the bounding rectangle: "per" (it spans a
        periodic direction), "phys" (on a non-periodic domain side) or
        "cf" (inside the domain, or against a periodic seam it does not
        span)."""
        out = [["per", "per"] for _ in range(3)]
        for d in range(self.dim):
            per = self.kind[d][0] == O.ELL_PER
            touch = (self.rlo[d] == 0, self.rhi[d] == 2 * self.nc[d] - 1)
            for s in range(2):
                if per and touch[0] and touch[1]:
                    out[d][s] = "per"
                elif touch[s] and not per:
                    out[d][s] = "phys"
                else:
                    out[d][s] = "cf"
        return out

    # ------------------------------------------------------------ indexing
    def covered(self, c):
        return bool(self.cmask[c[2], c[1], c[0]])

    def cid(self, c):
        return c[0] + self.nc[0] * (c[1] + self.nc[1] * c[2])

    def fid(self, f):
        r = [f[d] - self.rlo[d] for d in range(3)]
        return self.ncc + r[0] + self.nr[0] * (r[1] + self.nr[1] * r[2])

    def in_region(self, f):
        if not all(self.rlo[d] <= f[d] <= self.rhi[d] for d in range(self.dim)):
            return False
        return bool(self.mask[f[2] - self.rlo[2], f[1] - self.rlo[1], f[0] - self.rlo[0]])

    def fine_nb(self, f, d, step):
        """Fine neighbour of f along d in the fine domain (periodic wrap), or
        None beyond a non-periodic side."""
        q = list(f)
        q[d] += step
        nf = 2 * self.nc[d]
        if 0 <= q[d] < nf:
            return q
        if self.kind[d][0] == O.ELL_PER:
            q[d] %= nf
            return q
        return None

    def coarse_nb(self, c, d, step):
        """Coarse neighbour of c along d (periodic wrap) or None beyond a wall."""
        q = list(c)
        q[d] += step
        if 0 <= q[d] < self.nc[d]:
            return q
        if self.kind[d][0] == O.ELL_PER:
            q[d] %= self.nc[d]
            return q
        return None

    def covered_mask(self):
        return self.cmask.copy()

    # ------------------------------------------------------------- fluxes
    # Each flux is a dict {unknown: coefficient} of beta d(phi)/dx_d at a
    # face, positive along +d.
    @staticmethod
    def _acc(out, terms, a):
        for k, v in terms.items():
            out[k] = out.get(k, 0.0) + a * v

    def cf_ghost(self, f0, d, s):
        """Ghost value beyond the C-F face on side s of fine cell f0 (dict)."""
        inward = 1 if s == 0 else -1
        f1 = list(f0)
        f1[d] += inward
        assert self.in_region(f1)
        C = [f0[e] // 2 for e in range(3)]
        C = self.coarse_nb(C, d, -inward)
        assert C is not None and not self.covered(C)
        cint = {self.cid(C): 1.0}
        for t in range(self.dim):
            if t == d:
                continue
            sig = -0.25 if f0[t] % 2 == 0 else 0.25
            # transverse neighbours that are composite unknowns (not beyond a
            # wall, not covered by the fine level)
            cp, cm = self.coarse_nb(C, t, 1), self.coarse_nb(C, t, -1)
            if cp is not None and self.covered(cp):
                cp = None
            if cm is not None and self.covered(cm):
                cm = None
            if cp is not None and cm is not None:
                self._acc(cint, {self.cid(cp): 0.5, self.cid(cm): -0.5}, sig)
            elif cp is not None:
                self._acc(cint, {self.cid(cp): 1.0, self.cid(C): -1.0}, sig)
            elif cm is not None:
                self._acc(cint, {self.cid(C): 1.0, self.cid(cm): -1.0}, sig)
        g = {}
        self._acc(g, cint, W_COARSE)
        self._acc(g, {self.fid(f0): 1.0}, W_FINE0)
        self._acc(g, {self.fid(f1): 1.0}, W_FINE1)
        return g

    def fine_flux(self, bf, f, d, s):
        """Flux through the face on side s (0 low, 1 high) of fine cell f."""
        r = [f[e] - self.rlo[e] for e in range(3)]
        fr = list(r)
        fr[d] += s
        beta = bf[d][fr[2], fr[1], fr[0]]
        h = self.hf[d]
        nb = self.fine_nb(f, d, 1 if s else -1)
        if nb is None:
            # a physical side of the domain
            if self.kind[d][s] == O.ELL_DIR:
                return {self.fid(f): (-2.0 if s else 2.0) * beta / h}
            return {}
        if self.in_region(nb):
            lo, hi = (f, nb) if s else (nb, f)
            return {self.fid(hi): beta / h, self.fid(lo): -beta / h}
        g = self.cf_ghost(f, d, s)
        out = {}
        if s == 0:
            self._acc(out, {self.fid(f): 1.0}, beta / h)
            self._acc(out, g, -beta / h)
        else:
            self._acc(out, g, beta / h)
            self._acc(out, {self.fid(f): 1.0}, -beta / h)
        return out

    def coarse_flux(self, bc_, bf, c, d, s):
        """Flux through the face on side s of uncovered coarse cell c."""
        fc = list(c)
        fc[d] += s
        beta = bc_[d][fc[2], fc[1], fc[0]]
        h = self.hc[d]
        nb = self.coarse_nb(c, d, 1 if s else -1)
        if nb is None:
            if self.kind[d][s] == O.ELL_DIR:
                return {self.cid(c): (-2.0 if s else 2.0) * beta / h}
            return {}
        if not self.covered(nb):
            lo, hi = (c, nb) if s else (nb, c)
            return {self.cid(hi): beta / h, self.cid(lo): -beta / h}
        # C-F face: mean of the fine fluxes through it, taken from the fine
        # cells just inside (their face on the far side from c)
        fs = 0 if s else 1
        f0d = 2 * nb[d] if s else 2 * nb[d] + 1
        tr = [e for e in range(self.dim) if e != d]
        out = {}
        cnt = 2 ** len(tr)
        for m in range(cnt):
            f = [0, 0, 0]
            f[d] = f0d
            for q, t in enumerate(tr):
                f[t] = 2 * c[t] + ((m >> q) & 1)
            self._acc(out, self.fine_flux(bf, f, d, fs), 1.0 / cnt)
        return out

    # ----------------------------------------------------------- operator
    def fine_in_mask(self):
        """Fine cells of the boxes, in bounding-rectangle order."""
        for f in self.fine_cells():
            if self.in_region(f):
                yield f

    def cells(self):
        nz = self.nc[2]
        for K in range(nz):
            for J in range(self.nc[1]):
                for I in range(self.nc[0]):
                    yield (I, J, K)

    def fine_cells(self):
        for k in range(self.rlo[2], self.rhi[2] + 1):
            for j in range(self.rlo[1], self.rhi[1] + 1):
                for i in range(self.rlo[0], self.rhi[0] + 1):
                    yield (i, j, k)

    def matrix(self, bc_, bf):
        """Composite operator A (alpha = 0): rows and columns are all coarse
        cells (covered ones carry an identity row and are referenced by no
        other row), then the region's fine cells."""
        rows, cols, vals = [], [], []

        def put(r, terms, a):
            for k, v in terms.items():
                rows.append(r)
                cols.append(k)
                vals.append(a * v)

        for c in self.cells():
            r = self.cid(c)
            if self.covered(c):
                put(r, {r: 1.0}, 1.0)
                continue
            for d in range(self.dim):
                put(r, self.coarse_flux(bc_, bf, c, d, 1), -1.0 / self.hc[d])
                put(r, self.coarse_flux(bc_, bf, c, d, 0), 1.0 / self.hc[d])
        for f in self.fine_cells():
            r = self.fid(f)
            if not self.in_region(f):
                put(r, {r: 1.0}, 1.0)     # a hole of the bounding rectangle
                continue
            for d in range(self.dim):
                put(r, self.fine_flux(bf, f, d, 1), -1.0 / self.hf[d])
                put(r, self.fine_flux(bf, f, d, 0), 1.0 / self.hf[d])
        n = self.ncc + self.nfc
        return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))

    def pack(self, phic, phif):
        return np.concatenate([np.asarray(phic, dtype=np.float64).ravel(),
                               np.asarray(phif, dtype=np.float64).ravel()])

    def unpack(self, x):
        c = x[: self.ncc].reshape(self.nc[2], self.nc[1], self.nc[0])
        f = x[self.ncc:].reshape(self.nr[2], self.nr[1], self.nr[0])
        return c, f

    def singular(self):
        return all(self.kind[d][s] != O.ELL_DIR for d in range(self.dim) for s in range(2))

    def volume_weights(self):
        """Composite cell volumes in fine-cell units (covered coarse cells 0)."""
        wc = np.where(self.covered_mask(), 0.0, float(2 ** self.dim)).ravel()
        return np.concatenate([wc, np.where(self.mask, 1.0, 0.0).ravel()])

    # ------------------------------------------------------ level transfers
    def _block_mean(self, a, fz, fy, fx):
        """Means over blocks of fz x fy x fx of a (nz, ny, nx) array."""
        nz, ny, nx = a.shape
        return a.reshape(nz // fz, fz, ny // fy, fy, nx // fx, fx).mean(axis=(1, 3, 5))

    def _cbox(self, extra=(0, 0, 0)):
        """Slices of the bounding rectangle's coarse range (plus extra cells
        per direction) in (z, y, x) order."""
        return tuple(slice(self.rlo[d] // 2, self.rhi[d] // 2 + 1 + extra[d]) if d < self.dim else slice(None)
                     for d in (2, 1, 0))

    def restrict_cells(self, xc, xf):
        """ml_cc_restriction: covered coarse cells <- mean of their fine cells."""
        out = np.array(xc, dtype=np.float64, copy=True)
        f = (2 if self.dim == 3 else 1, 2, 2)
        m = self._block_mean(np.asarray(xf, dtype=np.float64), *f)
        sl = self._cbox()
        cov = self.cmask[sl]
        sub = out[sl]
        sub[cov] = m[cov]
        out[sl] = sub
        return out

    def face_under_fine(self, d):
        """Coarse faces of direction d (over the bounding rectangle's coarse
        face range) with a covered cell on at least one side."""
        sl = self._cbox(tuple(1 if e == d else 0 for e in range(3)))
        shp = [x.stop - x.start if x.start is not None else n for x, n in zip(sl, self.cmask.shape)]
        m = np.zeros(shp, dtype=bool)
        ax = 2 - d
        lo = self.rlo[d] // 2
        for side in (-1, 0):
            idx = np.arange(shp[ax]) + lo + side
            per = self.kind[d][0] == O.ELL_PER
            ok = (idx >= 0) & (idx < self.nc[d])
            if per:
                idx = np.mod(idx, self.nc[d])
                ok[:] = True
            idx = np.clip(idx, 0, self.nc[d] - 1)
            csl = list(self._cbox())
            csl[ax] = idx
            c = self.cmask[tuple(csl) if ax != 0 else (idx,) + tuple(csl[1:])]
            c = np.where(np.expand_dims(ok, tuple(e for e in range(3) if e != ax)), c, False)
            m |= c
        return sl, m

    def restrict_faces(self, fc, ff):
        """ml_edge_restriction: coarse faces with a covered cell on either side
        (faces of the fine level's boundary included) <- mean of the
        2^(dim-1) fine faces; on a periodic seam the face's periodic image
        follows where only one of the two images lies under the fine level."""
        out = [np.array(x, dtype=np.float64, copy=True) for x in fc]
        for d in range(self.dim):
            a = ff[d][:, :, ::2] if d == 0 else (ff[d][:, ::2, :] if d == 1 else ff[d][::2])
            f = (1 if d == 2 else (2 if self.dim == 3 else 1), 1 if d == 1 else 2, 1 if d == 0 else 2)
            m = self._block_mean(np.asarray(a, dtype=np.float64), *f)
            sl, under = self.face_under_fine(d)
            sub = out[d][sl]
            sub[under] = m[under]
            out[d][sl] = sub
            if self.kind[d][0] == O.ELL_PER:
                ax = 2 - d
                n = self.nc[d]
                full = np.zeros(out[d].shape, dtype=bool)
                fsub = full[sl]
                fsub[under] = True
                full[sl] = fsub
                lo_i = [slice(None)] * 3
                hi_i = [slice(None)] * 3
                lo_i[ax], hi_i[ax] = 0, n
                w0, wn = full[tuple(lo_i)], full[tuple(hi_i)]
                v0, vn = out[d][tuple(lo_i)].copy(), out[d][tuple(hi_i)].copy()
                out[d][tuple(hi_i)] = np.where(w0 & ~wn, v0, vn)
                out[d][tuple(lo_i)] = np.where(wn & ~w0, vn, v0)
        return out

    def fine_face_fluxes(self, bf, x):
        """beta G phi on every face of the fine level (composite fluxes); 0 on
        faces of the bounding rectangle between two holes."""
        out = []
        for d in range(self.dim):
            shp = [self.nr[2], self.nr[1], self.nr[0]]
            shp[2 - d] += 1
            F = np.zeros(shp)
            for f in self.fine_in_mask():
                r = [f[e] - self.rlo[e] for e in range(3)]
                F[r[2], r[1], r[0]] = sum(v * x[k] for k, v in self.fine_flux(bf, f, d, 0).items())
                nb = self.fine_nb(f, d, 1)
                if r[d] == self.nr[d] - 1 or nb is None or not self.in_region(nb):
                    rr = list(r)
                    rr[d] += 1
                    F[rr[2], rr[1], rr[0]] = sum(v * x[k] for k, v in self.fine_flux(bf, f, d, 1).items())
            out.append(F)
        return out

    def grids(self):
        gc = O.make_grid(self.nc[:self.dim], self.hc[:self.dim], self.bc[:self.dim])
        gf = O.make_grid(self.nr[:self.dim], self.hf[:self.dim], self.bc[:self.dim])
        return gc, gf


def rel_eps(eps_mac=1e-10, nlevs=2, eps_max=1e-8, factor=10.0):
    """Source/macproject.f90:217 with mg_eps.f90:22-28."""
    return min(eps_mac * factor ** (nlevs - 1), eps_max)


def macproject2(tl: TwoLevel, faces_c, faces_f, rho_c, rho_f, macrhs_c=None, macrhs_f=None,
                beta0=None, beta0_edge=None, beta0_f=None, beta0_edge_f=None, alt=False):
    """Two-level macproject with a sparse direct composite solve.  beta0*
    are the planar profiles of each level over its whole domain (radial =
    last direction); the fine level's are sliced to the region here."""
    gc, gf = tl.grids()
    dim = tl.dim
    rd = dim - 1
    r0 = tl.rlo[rd]
    fc = [np.array(x, dtype=np.float64, copy=True) for x in faces_c]
    ff = [np.array(x, dtype=np.float64, copy=True) for x in faces_f]
    use_b0 = beta0 is not None
    if use_b0:
        b0f = np.ascontiguousarray(beta0_f[r0:r0 + tl.nr[rd]])
        b0ef = np.ascontiguousarray(beta0_edge_f[r0:r0 + tl.nr[rd] + 1])
        fc = O.scale_faces_1d(gc, fc, beta0, beta0_edge, 1)
        ff = O.scale_faces_1d(gf, ff, b0f, b0ef, 1)
    fc = tl.restrict_faces(fc, ff)
    rh_c = O.divumac(gc, *fc[:2], fc[2] if dim == 3 else None, macrhs=macrhs_c)
    rh_f = O.divumac(gf, *ff[:2], ff[2] if dim == 3 else None, macrhs=macrhs_f)
    rh_c = tl.restrict_cells(rh_c, rh_f)
    bc_ = O.mk_mac_coeffs(gc, rho_c)
    bf = O.mk_mac_coeffs(gf, rho_f)
    if use_b0:
        for _ in range(2 if alt else 1):
            bc_ = O.scale_faces_1d(gc, bc_, beta0, beta0_edge, 1)
            bf = O.scale_faces_1d(gf, bf, b0f, b0ef, 1)
    b = tl.pack(rh_c, rh_f)
    off = np.concatenate([tl.covered_mask().ravel(), ~tl.mask.ravel()])
    b[off] = 0.0
    wv = tl.volume_weights()
    mean = 0.0
    if tl.singular():
        mean = math.fsum(wv * b) / wv.sum()
        b = b - mean
        b[off] = 0.0
    A = tl.matrix(bc_, bf)
    if tl.singular():
        # pin the first fine cell of the boxes, then shift to composite mean zero
        A = A.tolil()
        k = tl.ncc + int(np.argmax(tl.mask.ravel()))
        A.rows[k] = [k]
        A.data[k] = [1.0]
        A = A.tocsr()
        bb = b.copy()
        bb[k] = 0.0
        x = spla.spsolve(A.tocsc(), bb)
        x = x - math.fsum(wv * x) / wv.sum()
        A = tl.matrix(bc_, bf)
    else:
        x = spla.spsolve(A.tocsc(), b)
    res = b - A @ x
    res[off] = 0.0
    phic, phif = tl.unpack(x.copy())
    phic = tl.restrict_cells(phic, phif)
    x[: tl.ncc] = phic.ravel()
    # mkumac: coarse with the single-level rule (faces under the region are
    # overwritten by the edge restriction), fine with the composite fluxes
    fc = O.mkumac(gc, fc, phic, bc_)
    F = tl.fine_face_fluxes(bf, x)
    ff = [ff[d] - F[d] for d in range(dim)]
    fc = tl.restrict_faces(fc, ff)
    if use_b0:
        fc = O.scale_faces_1d(gc, fc, beta0, beta0_edge, 0)
        ff = O.scale_faces_1d(gf, ff, b0f, b0ef, 0)
    return {"faces_c": fc, "faces_f": ff, "phi_c": phic, "phi_f": phif, "beta_c": bc_,
            "beta_f": bf, "rh_c": rh_c, "rh_f": rh_f, "mean": mean,
            "res": float(np.abs(res).max()), "A": A}


def composite_divergence(tl: TwoLevel, faces_c, faces_f, beta0=None, beta0_edge=None,
                         beta0_f=None, beta0_edge_f=None):
    """D(beta0 U) on uncovered coarse cells (covered ones 0) and fine cells."""
    gc, gf = tl.grids()
    dim = tl.dim
    fc, ff = list(faces_c), list(faces_f)
    if beta0 is not None:
        rd = dim - 1
        r0 = tl.rlo[rd]
        fc = O.scale_faces_1d(gc, fc, beta0, beta0_edge, 1)
        ff = O.scale_faces_1d(gf, ff, np.ascontiguousarray(beta0_f[r0:r0 + tl.nr[rd]]),
                              np.ascontiguousarray(beta0_edge_f[r0:r0 + tl.nr[rd] + 1]), 1)
    dc = -O.divumac(gc, *fc[:2], fc[2] if dim == 3 else None)
    df = -O.divumac(gf, *ff[:2], ff[2] if dim == 3 else None)
    dc[tl.covered_mask()] = 0.0
    df[~tl.mask] = 0.0
    return dc, df
