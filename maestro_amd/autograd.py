"""Differentiable MAC projection: a ``torch.autograd.Function`` over the
whole-domain C entry points of libmaestro_mac.so.

Forward (``mac_project_level``): for face velocities U (already scaled by
beta0) and solver face coefficients beta (beta0/rho or beta0^2/rho),
  rh = macrhs - D U;  solve A phi = rh with A = -D beta G;  U_p = U - beta G phi
and return phi (mean-removed on singular problems) and U_p.

Backward (``mac_adjoint_level``): a second solve with the same self-adjoint
operator, A lambda = dL/dphi - G^T(beta dL/dU_p), then per face
  dL/dU    = dL/dU_p - D^T lambda
  dL/dbeta = G(phi) ((D^T lambda) - dL/dU_p)
and dL/dmacrhs = lambda.  The forward phi is reused, not recomputed.

Layout: whole-domain float64 CUDA tensors, C order with x fastest: cells
(nz,ny,nx); x/y/z faces (nz,ny,nx+1), (nz,ny+1,nx), (nz+1,ny,nx).
"""
from __future__ import annotations

import torch

from . import mac


def level_plan(n, dx, phys_bc, **opts) -> mac.MacPlan:
    """A plan over one box covering the domain, for the whole-domain entries."""
    dim = len(n)
    lay = mac.Layout(dim=dim, boxes=[((0,) * dim, tuple(x - 1 for x in n))],
                     domain_lo=(0,) * dim, domain_hi=tuple(x - 1 for x in n), dx=tuple(dx),
                     phys_bc=phys_bc)
    opts.setdefault("device_ptrs", 1)
    return mac.MacPlan(lay, mac.default_opts(**opts))


def _singular(plan: mac.MacPlan) -> bool:
    return all(b[0] != mac.OUTLET and b[1] != mac.OUTLET for b in plan.layout.phys_bc)


class MacProjectFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, plan, ux, uy, uz, bx, by, bz, macrhs):
        dim = plan.layout.dim
        n = [plan.layout.domain_hi[d] - plan.layout.domain_lo[d] + 1 for d in range(dim)]
        cells = (n[2] if dim == 3 else 1, n[1], n[0])
        ux_o, uy_o = ux.detach().clone(), uy.detach().clone()
        uz_o = uz.detach().clone() if uz is not None else None
        phi = torch.zeros(cells, dtype=torch.float64, device=ux.device)
        mr = macrhs.detach().contiguous() if macrhs is not None else None
        plan.project_level(ux_o, uy_o, uz_o, phi, bx.detach().contiguous(),
                           by.detach().contiguous(),
                           bz.detach().contiguous() if bz is not None else None, mr)
        if _singular(plan):
            phi = phi - phi.mean()
        ctx.plan = plan
        ctx.has_macrhs = macrhs is not None
        ctx.dim = dim
        ctx.save_for_backward(phi, bx.detach(), by.detach(),
                              bz.detach() if bz is not None else torch.empty(0))
        return phi, ux_o, uy_o, uz_o

    @staticmethod
    def backward(ctx, gphi, gux, guy, guz):
        phi, bx, by, bz = ctx.saved_tensors
        dim = ctx.dim
        if dim == 2:
            bz = None
        z = lambda ref, g: g.contiguous() if g is not None else torch.zeros_like(ref)
        gphi = z(phi, gphi)
        gux, guy = z(bx, gux), z(by, guy)
        guz = z(bz, guz) if dim == 3 else None
        lam = torch.zeros_like(phi)
        oux, ouy = torch.empty_like(bx), torch.empty_like(by)
        obx, oby = torch.empty_like(bx), torch.empty_like(by)
        ouz = torch.empty_like(bz) if dim == 3 else None
        obz = torch.empty_like(bz) if dim == 3 else None
        ctx.plan.adjoint_level(phi.contiguous(), bx.contiguous(), by.contiguous(),
                               bz.contiguous() if bz is not None else None, gphi, gux, guy, guz,
                               lam, oux, ouy, ouz, obx, oby, obz)
        gm = lam if ctx.has_macrhs else None
        return None, oux, ouy, ouz, obx, oby, obz, gm


def mac_project(plan, ux, uy, uz, bx, by, bz, macrhs=None):
    """Differentiable whole-domain MAC projection; returns (phi, ux, uy, uz)."""
    return MacProjectFn.apply(plan, ux, uy, uz, bx, by, bz, macrhs)
