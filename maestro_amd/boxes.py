"""Per-box (multifab-like) arrays for the ``macproject`` boundary.

A box array is a C-contiguous numpy array indexed ``[k, j, i]`` (2D:
``[j, i]``) whose memory equals FBoxLib's Fortran fab ``a(i, j, k)`` over
``lo-ng : hi+ng`` (``hi+1+ng`` along the face direction of an edge array),
Docs/architecture/architecture.tex:617-644.  These helpers move data between
the whole-domain layout of :mod:`maestro_amd.workloads` and per-box arrays.
"""
from __future__ import annotations

import numpy as np

from .workloads import PERIODIC


def _axes(dim):
    # numpy axis of direction d in a [k, j, i] / [j, i] array
    return [dim - 1 - d for d in range(dim)]


def box_shape(lo, hi, ng, dim, face_dir=None):
    ext = [hi[d] - lo[d] + 1 + 2 * ng + (1 if face_dir == d else 0) for d in range(dim)]
    return tuple(reversed(ext))


def _wrap_index(idx, n, periodic, faces):
    """Map global indices into the domain; -1 where outside a non-periodic side.
    For face arrays index n is valid (the domain's high face)."""
    top = n if faces else n - 1
    out = np.array(idx)
    bad = (out < 0) | (out > top)
    if periodic:
        out = np.where(bad, np.mod(out, n), out)
        return out, np.zeros_like(bad)
    return np.where(bad, 0, out), bad


def split_cells(glob, lo, hi, ng, dim, bc, ghost_fill=0.0, glob_ghost=0):
    """Cut a whole-domain cell array into a box array with ``ng`` ghosts.
    ``glob`` may carry ``glob_ghost`` ghost layers of its own (e.g. rho),
    used for positions outside non-periodic domain sides."""
    src = glob[0] if (dim == 2 and glob.ndim == 3) else glob
    n = [src.shape[dim - 1 - d] - 2 * glob_ghost for d in range(dim)]
    idx = []
    masks = []
    for d in range(dim):
        g = np.arange(lo[d] - ng, hi[d] + ng + 1)
        if glob_ghost:
            per = bc[d][0] == PERIODIC
            if per:
                w = np.mod(g, n[d]) + glob_ghost
                bad = np.zeros(g.shape, bool)
            else:
                w = g + glob_ghost
                bad = (w < 0) | (w >= n[d] + 2 * glob_ghost)
                w = np.clip(w, 0, n[d] + 2 * glob_ghost - 1)
        else:
            w, bad = _wrap_index(g, n[d], bc[d][0] == PERIODIC, False)
        idx.append(w)
        masks.append(bad)
    ix = np.ix_(*reversed(idx))
    out = np.array(src[ix], dtype=np.float64)
    m = np.zeros(out.shape, bool)
    for d in range(dim):
        shp = [1] * dim
        shp[dim - 1 - d] = -1
        m |= masks[d].reshape(shp)
    out[m] = ghost_fill
    return np.ascontiguousarray(out)


def split_faces(glob, d, lo, hi, ng, dim, bc, fill=0.0):
    """Cut a whole-domain face array of direction ``d`` into a box array."""
    shp = glob.shape if dim == 3 else glob.shape[1:]
    n = [shp[dim - 1 - e] - (1 if e == d else 0) for e in range(dim)]
    idx, masks = [], []
    for e in range(dim):
        g = np.arange(lo[e] - ng, hi[e] + ng + 1 + (1 if e == d else 0))
        w, bad = _wrap_index(g, n[e], bc[e][0] == PERIODIC, e == d)
        idx.append(w)
        masks.append(bad)
    src = glob if dim == 3 else glob[0]
    out = np.array(src[np.ix_(*reversed(idx))], dtype=np.float64)
    m = np.zeros(out.shape, bool)
    for e in range(dim):
        s = [1] * dim
        s[dim - 1 - e] = -1
        m |= masks[e].reshape(s)
    out[m] = fill
    return np.ascontiguousarray(out)


def split_workload(wl, ng_u=1, ng_phi=1, ng_rho=1, ng_rhs=0, ng_cart=1, macrhs=None):
    """Per-box arrays of a :func:`maestro_amd.workloads.make_workload` dict:
    returns (boxes, umac, phi, rho, macrhs, beta0_cart_edge)."""
    from .workloads import box_list
    dim, n, bc = wl["dim"], wl["n"], wl["bc"]
    boxes = box_list(n, wl["box"])
    umac = [[split_faces(wl["faces"][d], d, lo, hi, ng_u, dim, bc) for d in range(dim)]
            for lo, hi in boxes]
    phi = [split_cells(wl["phi"], lo, hi, ng_phi, dim, bc) for lo, hi in boxes]
    rho = [split_cells(wl["rho"], lo, hi, ng_rho, dim, bc, glob_ghost=1) for lo, hi in boxes]
    rhs = None
    if macrhs is not None:
        rhs = [split_cells(macrhs, lo, hi, ng_rhs, dim, bc) for lo, hi in boxes]
    cart = None
    if wl.get("cart") is not None:
        cart = [[split_faces(wl["cart"][d], d, lo, hi, ng_cart, dim, bc, fill=1.0)
                 for d in range(dim)] for lo, hi in boxes]
    return boxes, umac, phi, rho, rhs, cart


def gather_cells(boxes_arr, boxes, ng, n, dim):
    """Assemble the valid regions of box cell arrays into a whole-domain array."""
    out = np.zeros((n[2] if dim == 3 else 1, n[1], n[0]))
    for a, (lo, hi) in zip(boxes_arr, boxes):
        v = a[(slice(ng, a.shape[0] - ng),) * dim] if ng else a
        if dim == 3:
            out[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = v
        else:
            out[0, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = v
    return out


def gather_faces(boxes_arr, boxes, d, ng, n, dim):
    """Assemble the valid faces of box face arrays (direction ``d``)."""
    shp = [n[0], n[1], n[2] if dim == 3 else 1]
    shp[d] += 1
    out = np.zeros((shp[2], shp[1], shp[0]))
    for a, (lo, hi) in zip(boxes_arr, boxes):
        sl_src = []
        for e in reversed(range(dim)):
            ln = hi[e] - lo[e] + 1 + (1 if e == d else 0)
            sl_src.append(slice(ng, ng + ln))
        v = a[tuple(sl_src)]
        if dim == 3:
            out[lo[2]:lo[2] + v.shape[0], lo[1]:lo[1] + v.shape[1], lo[0]:lo[0] + v.shape[2]] = v
        else:
            out[0, lo[1]:lo[1] + v.shape[0], lo[0]:lo[0] + v.shape[1]] = v
    return out
