"""CPU tests that pin the two-level (AMR) oracle: the composite operator's
exactness on the fields its coarse-fine interpolation reproduces,
conservation, the projection's invariants and the two-level
test_projection known answers (Maestro-tests.ini:245-257)."""
import numpy as np
import pytest

from maestro_amd import workloads as W

P, WALL, OUT = W.PERIODIC, W.SLIP_WALL, W.OUTLET


def ml(oracle_lib):
    from oracle import ml_oracle as M
    return M


def centres(tl, level):
    """Cell-centre coordinates (x, y, z) of a level's cells, oracle layout."""
    if level == 0:
        n, h, off = tl.nc, tl.hc, [0, 0, 0]
    else:
        n, h, off = tl.nr, tl.hf, tl.rlo
    ax = [(np.arange(n[d]) + off[d] + 0.5) * h[d] for d in range(3)]
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    return X, Y, Z


def ones_beta(tl):
    bc_ = [np.ones(tuple(tl.nc[2 - e] + (2 - e == d) for e in range(3))) for d in range(tl.dim)]
    bf = [np.ones(tuple(tl.nr[2 - e] + (2 - e == d) for e in range(3))) for d in range(tl.dim)]
    return bc_, bf


def apply(tl, A, fc, ff):
    y = A @ tl.pack(fc, ff)
    return tl.unpack(y)


def interior_mask(tl, level, margin):
    """Cells whose coarse-cell distance to every periodic seam is >= margin."""
    X, Y, Z = centres(tl, level)
    m = np.ones(X.shape, dtype=bool)
    for d, C in enumerate((X, Y, Z)[:tl.dim]):
        L = tl.nc[d] * tl.hc[d]
        m &= (C > margin * tl.hc[d]) & (C < L - margin * tl.hc[d])
    return m


@pytest.mark.parametrize("dim", [2, 3])
def test_composite_operator_exact_on_reproduced_fields(oracle_lib, dim):
    """The ghost value is exact for fields quadratic along the C-F normal and
    linear along the face, so A x^2 = -2 and A (x y) = 0 (beta = 1) on every
    composite cell, coarse cells next to the fine region (their flux is the
    mean of the fine fluxes) and fine cells next to it included."""
    M = ml(oracle_lib)
    n = 16
    per = [[P, P]] * dim
    # a slab across x: C-F faces normal to x only (it spans y, z)
    lo = [10] + [0] * (dim - 1)
    hi = [19] + [2 * n - 1] * (dim - 1)
    tl = M.TwoLevel(dim, [n] * dim, [1.0 / n] * dim, per, lo, hi)
    assert tl.side[0] == ["cf", "cf"] and tl.side[1] == ["per", "per"]
    A = tl.matrix(*ones_beta(tl))
    Xc, _, _ = centres(tl, 0)
    Xf, _, _ = centres(tl, 1)
    cov = tl.covered_mask()
    yc, yf = apply(tl, A, Xc ** 2, Xf ** 2)
    mc, mf = interior_mask(tl, 0, 2) & ~cov, interior_mask(tl, 1, 2)
    assert np.abs(yc[mc] + 2).max() < 1e-9 and np.abs(yf[mf] + 2).max() < 1e-9
    # the weights are pinned: moving a twentieth between the coarse weight
    # and the first fine one breaks the result
    saved = (M.W_COARSE, M.W_FINE0)
    try:
        M.W_COARSE, M.W_FINE0 = saved[0] + 0.05, saved[1] - 0.05
        yc2, yf2 = apply(tl, tl.matrix(*ones_beta(tl)), Xc ** 2, Xf ** 2)
        assert np.abs(yf2[mf] + 2).max() > 1e-2
    finally:
        M.W_COARSE, M.W_FINE0 = saved
    # a box (2D) or a column spanning z (3D): C-F faces normal to x and y;
    # the bilinear x y is reproduced (tangential slope exact)
    lo = [10, 8] + [0] * (dim - 2)
    hi = [19, 21] + [2 * n - 1] * (dim - 2)
    tl = M.TwoLevel(dim, [n] * dim, [1.0 / n] * dim, per, lo, hi)
    A = tl.matrix(*ones_beta(tl))
    Xc, Yc, _ = centres(tl, 0)
    Xf, Yf, _ = centres(tl, 1)
    cov = tl.covered_mask()
    mc, mf = interior_mask(tl, 0, 2) & ~cov, interior_mask(tl, 1, 2)
    for fc, ff in ((Xc * Yc, Xf * Yf), (Xc, Xf), (Yc, Yf), (np.ones_like(Xc), np.ones_like(Xf))):
        yc, yf = apply(tl, A, fc, ff)
        assert np.abs(yc[mc]).max() < 1e-8 and np.abs(yf[mf]).max() < 1e-8


REGIONS = [
    # (dim, n, bc, lo, hi): interior box, box on a wall, box across the
    # periodic seam, and the reference's two-level grid (left half, all y)
    (2, 16, [[P, P], [P, P]], [8, 10], [19, 21]),
    (2, 16, [[WALL, WALL], [WALL, OUT]], [0, 10], [11, 31]),
    (2, 16, [[P, P], [WALL, WALL]], [0, 4], [15, 13]),
    (2, 16, [[P, P], [P, P]], [0, 0], [15, 31]),
    (3, 8, [[P, P], [P, P], [WALL, WALL]], [4, 2, 0], [11, 9, 7]),
    (3, 8, [[WALL, WALL], [P, P], [WALL, OUT]], [0, 4, 6], [7, 11, 15]),
]


@pytest.mark.parametrize("dim,n,bc,lo,hi", REGIONS)
def test_composite_operator_conserves(oracle_lib, dim, n, bc, lo, hi):
    """Volume sum of A phi = the flux through the domain boundary: zero
    without Dirichlet sides for any phi and beta (the coarse side of a C-F
    face takes exactly the fine fluxes), and A 1 = 0 there (the ghost
    weights sum to one)."""
    M = ml(oracle_lib)
    rng = np.random.default_rng(7)
    tl = M.TwoLevel(dim, [n] * dim, [1.0 / n] * dim, bc, lo, hi)
    bc_, bf = ones_beta(tl)
    bc_ = [0.5 + rng.random(b.shape) for b in bc_]
    bf = [0.5 + rng.random(b.shape) for b in bf]
    # a periodic direction's last face is its first one (as beta from a
    # periodic rho is)
    for d in range(dim):
        if bc[d][0] == P:
            lo_, hi_ = [slice(None)] * 3, [slice(None)] * 3
            lo_[2 - d], hi_[2 - d] = 0, -1
            bc_[d][tuple(hi_)] = bc_[d][tuple(lo_)]
            if tl.side[d][0] == "per":
                bf[d][tuple(hi_)] = bf[d][tuple(lo_)]
    A = tl.matrix(bc_, bf)
    x = rng.standard_normal(tl.ncc + tl.nfc)
    y = A @ x
    w = tl.volume_weights()
    scale = np.abs(w * y).sum()
    if tl.singular():
        assert abs(np.dot(w, y)) < 1e-12 * scale
        one = np.concatenate([np.where(tl.covered_mask().ravel(), 0.0, 1.0), np.ones(tl.nfc)])
        assert np.abs(A @ one)[w > 0].max() < 1e-10
    else:
        # only the Dirichlet faces carry flux out: beta phi / (h/2) per face
        assert abs(np.dot(w, y)) > 1e-6 * scale


# fine levels made of several boxes: an L, two disjoint boxes (one on a
# wall), a 3D step, and two boxes that meet across the periodic seam
UNIONS = [
    (2, 16, [[P, P], [P, P]], [([8, 8], [15, 23]), ([16, 8], [23, 15])]),
    (2, 16, [[WALL, WALL], [P, P]], [([0, 2], [7, 9]), ([18, 20], [27, 27])]),
    (3, 8, [[P, P], [P, P], [WALL, WALL]], [([4, 4, 0], [11, 11, 7]), ([12, 4, 4], [15, 11, 11])]),
    (2, 16, [[P, P], [WALL, OUT]], [([0, 4], [5, 11]), ([26, 4], [31, 11])]),
]


def random_two_level(M, O, dim, n, bc, lo, hi, seed, beta0=True, boxes=None):
    rng = np.random.default_rng(seed)
    tl = M.TwoLevel(dim, [n] * dim, [1.0 / n] * dim, bc, lo, hi, boxes=boxes)
    gc, gf = tl.grids()
    fc = [rng.uniform(-1, 1, O.face_shape(gc, d)) for d in range(dim)]
    ff = [rng.uniform(-1, 1, O.face_shape(gf, d)) for d in range(dim)]
    rc = 1 + 0.1 * rng.random(O.rho_shape(gc))
    rf = 1 + 0.1 * rng.random(O.rho_shape(gf))
    # periodic images agree, as in a caller's filled arrays: the last face
    # of a periodic direction is its first, rho's ghosts are wrapped cells
    for d in range(dim):
        if bc[d][0] != P:
            continue
        lvls = [(fc, rc)] + ([(ff, rf)] if tl.side[d][0] == "per" else [])
        for faces, rho in lvls:
            lo_, hi_ = [slice(None)] * 3, [slice(None)] * 3
            lo_[2 - d], hi_[2 - d] = 0, -1
            faces[d][tuple(hi_)] = faces[d][tuple(lo_)]
            a = [slice(None)] * 3
            b = [slice(None)] * 3
            a[2 - d], b[2 - d] = 0, -2
            rho[tuple(a)] = rho[tuple(b)]
            a[2 - d], b[2 - d] = -1, 1
            rho[tuple(a)] = rho[tuple(b)]
    kw = {}
    if beta0:
        rd = dim - 1
        zc = (np.arange(n) + 0.5) / n
        ze = np.arange(n + 1) / n
        zf = (np.arange(2 * n) + 0.5) / (2 * n)
        zfe = np.arange(2 * n + 1) / (2 * n)
        kw = dict(beta0=np.exp(-2 * zc), beta0_edge=np.exp(-2 * ze), beta0_f=np.exp(-2 * zf),
                  beta0_edge_f=np.exp(-2 * zfe))
        assert rd == dim - 1
    return tl, fc, ff, rc, rf, kw


@pytest.mark.parametrize("dim,n,bc,lo,hi", REGIONS)
def test_two_level_projection_invariants(oracle_lib, dim, n, bc, lo, hi):
    """After the projection the composite divergence of beta0 U is zero on
    every composite cell (coarse cells next to the fine region see the fine
    faces through the edge restriction); projecting again changes nothing;
    wall-normal faces are untouched."""
    M = ml(oracle_lib)
    O = oracle_lib
    # planar beta0 along the last direction where that one is not periodic
    tl, fc, ff, rc, rf, kw = random_two_level(M, O, dim, n, bc, lo, hi, 11, beta0=bc[dim - 1][0] != P)
    for d in range(dim):
        for s, sl in ((0, 0), (1, -1)):
            if bc[d][s] == WALL:
                idx = [slice(None)] * 3
                idx[2 - d] = sl
                fc[d][tuple(idx)] = 0.0
                if tl.side[d][s] == "phys":
                    ff[d][tuple(idx)] = 0.0
    r = M.macproject2(tl, fc, ff, rc, rf, alt=True, **kw)
    assert r["res"] < 1e-9 * max(np.abs(r["rh_c"]).max(), np.abs(r["rh_f"]).max())
    dc, df = M.composite_divergence(tl, r["faces_c"], r["faces_f"], **kw)
    assert max(np.abs(dc).max(), np.abs(df).max()) < 1e-9
    r2 = M.macproject2(tl, r["faces_c"], r["faces_f"], rc, rf, alt=True, **kw)
    for d in range(dim):
        assert np.abs(r2["faces_c"][d] - r["faces_c"][d]).max() < 1e-9
        assert np.abs(r2["faces_f"][d] - r["faces_f"][d]).max() < 1e-9
        for s, sl in ((0, 0), (1, -1)):
            if bc[d][s] == WALL:
                idx = [slice(None)] * 3
                idx[2 - d] = sl
                assert np.abs(r["faces_c"][d][tuple(idx)]).max() == 0.0


def test_two_level_test_projection_known_answers(oracle_lib):
    """Exec/UNIT_TESTS/test_projection on the reference's two-level grid
    (gr0_2d_2lev_coarse: the left half refined, all of y; inputs_2d_per_mac
    and inputs_2d_wall_mac): the divergence-free field plus a gradient
    projects back onto the field up to the truncation error of the C-F
    stencil, which falls with the mesh."""
    M = ml(oracle_lib)
    O = oracle_lib

    def fields(n, lo_f, nr):
        h = 1.0 / n
        xf, xc = np.arange(n + 1) * h, (np.arange(n) + 0.5) * h
        X, Y = np.meshgrid(xf, xc)
        u = -np.sin(np.pi * X) ** 2 * np.sin(2 * np.pi * Y)
        X, Y = np.meshgrid(xc, xf)
        v = np.sin(np.pi * Y) ** 2 * np.sin(2 * np.pi * X)
        return u, v, h

    errs = {}
    for kind, bc in (("per", [[P, P], [P, P]]), ("wall", [[WALL, WALL], [WALL, WALL]])):
        errs[kind] = []
        for n in (16, 32):
            tl = M.TwoLevel(2, [n, n], [1.0 / n] * 2, bc, [0, 0], [n - 1, 2 * n - 1])
            out = []
            for lev, m in ((0, n), (1, 2 * n)):
                u, v, h = fields(m, 0, 0)
                if lev == 1:
                    u, v = u[:, : n + 1], v[:, :n]
                xfc = [np.arange(s + 1) * h for s in (u.shape[1] - 1, v.shape[1])]
                if kind == "per":
                    def ph(x, y):
                        return 0.1 * np.cos(2 * np.pi * x) * np.cos(2 * np.pi * y)
                    X, Y = np.meshgrid(np.arange(u.shape[1]) * h, (np.arange(u.shape[0]) + 0.5) * h)
                    gx = -2 * np.pi * 0.1 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y)
                    X, Y = np.meshgrid((np.arange(v.shape[1]) + 0.5) * h, np.arange(v.shape[0]) * h)
                    gy = -2 * np.pi * 0.1 * np.cos(2 * np.pi * X) * np.sin(2 * np.pi * Y)
                else:
                    X = np.arange(u.shape[1]) * h
                    gx = np.broadcast_to(4 * X * (1 - X), u.shape).copy()
                    Y = np.arange(v.shape[0]) * h
                    gy = np.broadcast_to((4 * Y * (1 - Y))[:, None], v.shape).copy()
                out.append((u[None], v[None], gx[None], gy[None]))
                del xfc
            (uc, vc, gxc, gyc), (uf, vf, gxf, gyf) = out
            gc, gf = tl.grids()
            r = M.macproject2(tl, [uc + gxc, vc + gyc], [uf + gxf, vf + gyf],
                              np.ones(O.rho_shape(gc)), np.ones(O.rho_shape(gf)))
            e = max(np.abs(r["faces_c"][0] - uc).max(), np.abs(r["faces_c"][1] - vc).max(),
                    np.abs(r["faces_f"][0] - uf).max(), np.abs(r["faces_f"][1] - vf).max())
            errs[kind].append(e)
    for kind, (e16, e32) in errs.items():
        assert e32 < 0.06, (kind, errs)
        assert e32 < e16 / 1.8, (kind, errs)


@pytest.mark.parametrize("dim,n,bc,boxes", UNIONS)
def test_composite_operator_on_box_unions(oracle_lib, dim, n, bc, boxes):
    """Fine levels of several boxes: the composite operator conserves (the
    volume sum of A phi vanishes without Dirichlet sides), keeps constants
    in its null space, reproduces x y in 2D and linear fields in 3D (A = 0,
    beta = 1) away from the domain's sides, re-entrant corners included, and
    the projection
    leaves zero composite divergence and is idempotent."""
    M = ml(oracle_lib)
    O = oracle_lib
    rng = np.random.default_rng(17)
    tl = M.TwoLevel(dim, [n] * dim, [1.0 / n] * dim, bc, boxes=boxes)
    assert tl.mask.sum() < tl.nfc or len(boxes) == 1
    bc_, bf = ones_beta(tl)
    w = tl.volume_weights()
    A = tl.matrix(bc_, bf)
    if tl.singular():
        one = np.concatenate([np.where(tl.covered_mask().ravel(), 0.0, 1.0), tl.mask.ravel().astype(float)])
        assert np.abs(A @ one)[w > 0].max() < 1e-10
        x = rng.standard_normal(tl.ncc + tl.nfc)
        y = A @ x
        assert abs(np.dot(w, y)) < 1e-12 * np.abs(w * y).sum()
    Xc, Yc, _ = centres(tl, 0)
    Xf, Yf, _ = centres(tl, 1)
    mc = interior_mask(tl, 0, 2) & ~tl.covered_mask()
    mf = interior_mask(tl, 1, 2) & tl.mask
    # a field the interpolation reproduces on every C-F face: the bilinear
    # x y in 2D (linear along each face), a linear field in 3D (x y is
    # bilinear along a z-face, beyond the linear transverse slopes)
    if dim == 2:
        fc_, ff_ = Xc * Yc, Xf * Yf
    else:
        _, _, Zc = centres(tl, 0)
        _, _, Zf = centres(tl, 1)
        fc_, ff_ = Xc + 2 * Yc + 3 * Zc, Xf + 2 * Yf + 3 * Zf
    yc, yf = apply(tl, A, fc_, ff_)
    assert np.abs(yc[mc]).max() < 1e-8 and (not mf.any() or np.abs(yf[mf]).max() < 1e-8)
    # projection invariants
    lo = [min(b[0][d] for b in boxes) for d in range(dim)]
    hi = [max(b[1][d] for b in boxes) for d in range(dim)]
    use_b0 = bc[dim - 1][0] != P
    tl, fc, ff, rc, rf, kw = random_two_level(M, O, dim, n, bc, lo, hi, 19, beta0=use_b0, boxes=boxes)
    for d in range(dim):
        for s_, sl in ((0, 0), (1, -1)):
            if bc[d][s_] == WALL:
                idx = [slice(None)] * 3
                idx[2 - d] = sl
                fc[d][tuple(idx)] = 0.0
                if tl.side[d][s_] == "phys":
                    ff[d][tuple(idx)] = 0.0
    r = M.macproject2(tl, fc, ff, rc, rf, alt=True, **kw)
    dc, df = M.composite_divergence(tl, r["faces_c"], r["faces_f"], **kw)
    assert max(np.abs(dc).max(), np.abs(df).max()) < 1e-9
    r2 = M.macproject2(tl, r["faces_c"], r["faces_f"], rc, rf, alt=True, **kw)
    for d in range(dim):
        assert np.abs(r2["faces_c"][d] - r["faces_c"][d]).max() < 1e-9
