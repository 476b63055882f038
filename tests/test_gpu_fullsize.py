"""The workloads at their full sizes: the 2D 128^2 case against the oracle,
and the 256^3 plane-parallel case (FP64 and mixed), the 512^3 star (mixed
and FP64) and the 256^3 differentiable case through properties that hold at
any size:

- convergence: the final residual is at most 1e-10 of the initial one;
- the projection identity D(beta0 U_p) = -(rh - A phi): the discrete
  divergence of the projected velocity, recomputed here on the GPU from the
  returned faces, is bounded by the solver's final residual (plus rounding),
  where the input velocity's divergence is ~1e10 times larger;
- the FP64 and mixed paths agree (phi and U_p within 1e-8 relative);
- periodic duplicate faces are equal, wall-normal faces stay zero;
- the adjoint gradient matches central finite differences (16 seeded
  directions, step 1e-6, <= 1e-6 relative; the difference L(+h) - L(-h) is
  summed term by term so that it does not cancel).
"""
import numpy as np
import pytest

from maestro_amd import boxes as BX
from maestro_amd import mac
from maestro_amd import workloads as W

from .helpers import rel_err, run_gpu, run_oracle

pytestmark = pytest.mark.gpu


def project_device(torch, wl, prec, alt=1, **opts):
    """macproject on device-resident per-box arrays; returns whole-domain
    device tensors of the projected faces and phi, and the stats."""
    boxes, umac, phi, rho, rhs, cart = BX.split_workload(wl)
    cv = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    du = [[cv(a) for a in box] for box in umac]
    dp = [cv(a) for a in phi]
    dr = [cv(a) for a in rho]
    dc = None if cart is None else [[cv(a) for a in box] for box in cart]
    b0 = None if wl["beta0"] is None else cv(wl["beta0"])
    b0e = None if wl["beta0_edge"] is None else cv(wl["beta0_edge"])
    lay = mac.Layout.uniform(wl["n"], wl["box"], wl["dx"], wl["bc"])
    plan = mac.MacPlan(lay, mac.default_opts(device_ptrs=1, precision=prec, use_alt_energy_fix=alt, **opts))
    st = plan.project(du, dp, dr, None, b0, b0e, dc)
    plan.close()
    n = wl["n"]
    faces = []
    for d in range(3):
        shp = [n[2], n[1], n[0]]
        shp[2 - d] += 1
        f = torch.empty(shp, dtype=torch.float64, device="cuda")
        for (lo, hi), box in zip(boxes, du):
            a = box[d]
            e = [hi[q] - lo[q] + 1 + (1 if q == d else 0) for q in range(3)]
            f[lo[2]:lo[2] + e[2], lo[1]:lo[1] + e[1], lo[0]:lo[0] + e[0]] = \
                a[1:1 + e[2], 1:1 + e[1], 1:1 + e[0]]
        faces.append(f)
    ph = torch.empty((n[2], n[1], n[0]), dtype=torch.float64, device="cuda")
    for (lo, hi), a in zip(boxes, dp):
        ph[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = a[1:-1, 1:-1, 1:-1]
    del du, dp, dr, dc
    return faces, ph, st


def beta0_faces(torch, wl):
    """beta0 on the faces as the projection scales them (tangential faces the
    cell value, radial faces the edge value; or the 3-D face arrays)."""
    if wl["cart"] is not None:
        return [torch.from_numpy(np.ascontiguousarray(c)).cuda() for c in wl["cart"]]
    b0 = torch.from_numpy(wl["beta0"]).cuda()
    b0e = torch.from_numpy(wl["beta0_edge"]).cuda()
    return [b0[:, None, None], b0[:, None, None], b0e[:, None, None]]


def divergence(faces, b0f, dx):
    ux, uy, uz = (faces[d] * b0f[d] for d in range(3))
    return ((ux[:, :, 1:] - ux[:, :, :-1]) / dx[0] + (uy[:, 1:, :] - uy[:, :-1, :]) / dx[1]
            + (uz[1:] - uz[:-1]) / dx[2])


def check_projection(torch, wl, faces, st):
    assert st["converged"] and st["res"] <= 1e-10 * st["res0"] * (1 + 1e-12)
    div = divergence(faces, beta0_faces(torch, wl), wl["dx"])
    dmax = float(div.abs().max())
    print(f"  res0 {st['res0']:.3e} res {st['res']:.3e} max|D(beta0 U_p)| {dmax:.3e}")
    assert dmax <= 2.0 * st["res"] + 1e-12 * st["res0"]
    return dmax


def test_config0_2d_128_matches_oracle(gpu):
    """The 2D case at its own size: 128^2 in 16 boxes of 32^2."""
    wl = W.make_workload(0, n=(128, 128), box=32)
    truth = run_oracle(wl, eps=1e-13)
    for prec in (mac.PREC_FP64, mac.PREC_MIXED):
        out = run_gpu(wl, precision=prec)
        st = out["stats"]
        assert st["converged"] and st["res"] <= 1e-10 * st["res0"] * (1 + 1e-12)
        e_phi = rel_err(out["phi"], truth["phi"], remove_mean=True)
        e_u = max(rel_err(out["faces"][d], truth["faces"][d]) for d in range(2))
        print(f"  prec {prec}: iterations {st['iters']} phi {e_phi:.2e} U {e_u:.2e}")
        assert e_phi <= 1e-8 and e_u <= 1e-8


def oracle_truth(wl):
    """The oracle's FP64 solve of the same inputs to 1e-13 (mean-removed phi
    and the projected faces, whole-domain numpy arrays)."""
    ref = run_oracle(wl, eps=1e-13, alt=1)
    assert ref["stats"].converged
    return ref["faces"], ref["phi"] - ref["phi"].mean()


def against_truth(faces, ph, truth):
    """Relative 2-norm errors of phi (mean-removed) and of each face
    component against the oracle's 1e-13 solve."""
    ft, pt = truth
    p = ph.cpu().numpy()
    p = p - p.mean()
    e_phi = float(np.linalg.norm(p - pt) / np.linalg.norm(pt))
    e_u = max(float(np.linalg.norm(faces[d].cpu().numpy() - ft[d]) / np.linalg.norm(ft[d])) for d in range(3))
    return e_phi, e_u


def test_plane_parallel_256_fp64_and_mixed(gpu):
    """256^3 in 64 boxes of 64^3, isothermal beta0 (contrast 1e4), x/y
    periodic, z walls: every precision converges, satisfies the projection
    identity, agrees with the others and is within 1e-8 of the oracle's
    1e-13 solve of the same inputs (phi mean-removed, and U_p); so does the
    benchmark's headline, the mixed path with its FP32 sweeps over-relaxed by
    1.25 (measured: phi 6.5e-9, U 8.8e-11 from the 1e-13 solve)."""
    torch = gpu
    wl = W.make_workload(2, n=(256,) * 3, box=64)
    truth = oracle_truth(wl)
    res = {}
    for prec in (mac.PREC_FP64, mac.PREC_MIXED, mac.PREC_MIXED_BF16, "mixed_over_relaxed"):
        if prec == "mixed_over_relaxed":
            faces, ph, st = project_device(torch, wl, mac.PREC_MIXED, omega=1.25)
        else:
            faces, ph, st = project_device(torch, wl, prec)
        print(f"  prec {prec}: iterations {st['iters']}")
        check_projection(torch, wl, faces, st)
        # periodic duplicates equal, wall-normal faces zero
        assert torch.equal(faces[0][:, :, 0], faces[0][:, :, -1])
        assert torch.equal(faces[1][:, 0, :], faces[1][:, -1, :])
        assert float(faces[2][0].abs().max()) == 0.0 and float(faces[2][-1].abs().max()) == 0.0
        e_phi, e_u = against_truth(faces, ph, truth)
        print(f"  prec {prec} vs the oracle's 1e-13 solve: phi {e_phi:.2e}  U {e_u:.2e}")
        assert e_phi <= 1e-8 and e_u <= 1e-8
        res[prec] = (faces, ph - ph.mean())
    f64, p64 = res[mac.PREC_FP64]
    for prec in (mac.PREC_MIXED, mac.PREC_MIXED_BF16, "mixed_over_relaxed"):
        fmx, pmx = res[prec]
        e_phi = float((pmx - p64).norm() / p64.norm())
        e_u = max(float((fmx[d] - f64[d]).norm() / f64[d].norm()) for d in range(3))
        print(f"  prec {prec} vs FP64: phi {e_phi:.2e}  U {e_u:.2e}")
        assert e_phi <= 1e-8 and e_u <= 1e-8


@pytest.mark.parametrize("cycle", [3, 1])
def test_star_256_matches_oracle(gpu, cycle):
    """The star inputs at 256^3 (64 boxes of 64^3, outflow on every side,
    coefficient contrast 1e6), V- and F-cycles, against the oracle's solves of
    the same inputs.  With this contrast a 1e-10 relative residual leaves
    ~5e-8 in phi, for the reference's algorithm as for this one, so:
    - FP64 runs the oracle's algorithm: within 1e-9 of the oracle's own
      1e-10 solve with the same cycle shape (phi and U_p);
    - mixed precision is within max(1e-8, 1.5 x) the oracle's own 1e-10 error
      against its 1e-13 solve."""
    torch = gpu
    wl = W.make_workload(3, n=(256,) * 3, box=64)
    truth = oracle_truth(wl)
    ref = run_oracle(wl, eps=1e-10, alt=1, cycle_type=cycle, max_iter=200)
    assert ref["stats"].converged
    ref10 = (ref["faces"], ref["phi"] - ref["phi"].mean())
    rf, rp = ref10
    ft, pt = truth
    e_ref = (float(np.linalg.norm(rp - pt) / np.linalg.norm(pt)),
             max(float(np.linalg.norm(rf[d] - ft[d]) / np.linalg.norm(ft[d])) for d in range(3)))
    print(f"  cycle {cycle}: oracle at 1e-10 ({ref['stats'].iters} cycles) vs its 1e-13 solve: "
          f"phi {e_ref[0]:.2e}  U {e_ref[1]:.2e}")
    for prec in (mac.PREC_FP64, mac.PREC_MIXED):
        faces, ph, st = project_device(torch, wl, prec, cycle_type=cycle, max_iter=200)
        check_projection(torch, wl, faces, st)
        e_same = against_truth(faces, ph, ref10)
        e_true = against_truth(faces, ph, truth)
        print(f"  cycle {cycle} prec {prec}: {st['iters']} iterations ({st['inner_cycles']} FP32 cycles); "
              f"vs the oracle at 1e-10: phi {e_same[0]:.2e} U {e_same[1]:.2e}; "
              f"vs 1e-13: phi {e_true[0]:.2e} U {e_true[1]:.2e}")
        if prec == mac.PREC_FP64:
            assert abs(st["iters"] - ref["stats"].iters) <= 1
            assert e_same[0] <= 1e-9 and e_same[1] <= 1e-9
        for k in range(2):
            assert e_true[k] <= max(1e-8, 1.5 * e_ref[k])


def test_star_512_mixed_and_fp64(gpu):
    """512^3 in 512 boxes of 64^3, n = 1.5 polytrope with a density floor
    (coefficient contrast 1e6), outflow on every side (Dirichlet phi = 0)."""
    torch = gpu
    wl = W.make_workload(3, n=(512,) * 3, box=64)
    res = {}
    for prec in (mac.PREC_MIXED, mac.PREC_FP64):
        faces, ph, st = project_device(torch, wl, prec)
        print(f"  prec {prec}: iterations {st['iters']}")
        check_projection(torch, wl, faces, st)
        res[prec] = (faces, ph)
        del faces, ph
    (fmx, pmx), (f64, p64) = res[mac.PREC_MIXED], res[mac.PREC_FP64]
    # the truth: the FP64 solve to 1e-12.  With the 1e6 contrast a 1e-10
    # residual leaves ~1e-7 in phi, so each 1e-10 answer is held to a small
    # multiple of the FP64 path's own error rather than to the other answer
    ft, pt, st = project_device(torch, wl, mac.PREC_FP64, eps_rel=1e-12)
    assert st["converged"]
    err = lambda f, p: (float((p - pt).norm() / pt.norm()),
                        max(float((f[d] - ft[d]).norm() / ft[d].norm()) for d in range(3)))
    e64, emx = err(f64, p64), err(fmx, pmx)
    print(f"  vs the 1e-12 solve: FP64 phi {e64[0]:.2e} U {e64[1]:.2e}; mixed phi {emx[0]:.2e} U {emx[1]:.2e}")
    for k in range(2):
        # measured: FP64 phi 1.2e-7, U 1.8e-9; mixed phi 2.8e-8, U 4.1e-10
        assert e64[k] <= (3e-7, 5e-9)[k] and emx[k] <= max(1e-8, 3.0 * e64[k])


def test_adjoint_256_matches_finite_differences(gpu, oracle_lib):
    """The differentiable projection at 256^3 (the plane-parallel inputs):
    L = sum(w phi) + 1/2 ||U_p||^2, gradients with respect to U and the
    face coefficients, mixed-precision forward and adjoint."""
    torch = gpu
    O = oracle_lib
    from maestro_amd import autograd as AG
    from .test_gpu_adjoint import build, consistent_dir
    wl, plan, U, B, b0f, w = build(torch, O, 4, (256, 256, 256), alt=1, prec=1, eps=1e-13)
    dim = 3

    def forward(Uin, Bin):
        Us = [Uin[d] * b0f[d] for d in range(dim)]
        phi, ux, uy, uz = AG.mac_project(plan, Us[0], Us[1], Us[2], Bin[0], Bin[1], Bin[2])
        return phi, [ux / b0f[0], uy / b0f[1], uz / b0f[2]]

    Ul = [u.clone().requires_grad_(True) for u in U]
    Bl = [b.clone().requires_grad_(True) for b in B]
    phi, Up = forward(Ul, Bl)
    L = (w * phi).sum() + 0.5 * sum((u * u).sum() for u in Up)
    L.backward()
    gU = [u.grad.clone() for u in Ul]
    gB = [b.grad.clone() for b in Bl]
    rng = np.random.default_rng(2024)
    h = 1e-6
    worst = 0.0
    with torch.no_grad():
        for it in range(16):
            dU = [torch.from_numpy(rng.standard_normal(tuple(u.shape))).cuda() for u in U]
            dB = [torch.from_numpy(consistent_dir(rng, tuple(b.shape), d, wl["bc"])).cuda() * B[d]
                  for d, b in enumerate(B)]
            ad = sum(float((gU[d] * dU[d]).sum()) for d in range(dim)) + \
                sum(float((gB[d] * dB[d]).sum()) for d in range(dim))
            pp, up = forward([U[d] + h * dU[d] for d in range(dim)], [B[d] + h * dB[d] for d in range(dim)])
            pm, um = forward([U[d] - h * dU[d] for d in range(dim)], [B[d] - h * dB[d] for d in range(dim)])
            # L(+h) - L(-h) summed term by term (no cancellation of two large sums)
            dl = float((w * (pp - pm)).sum()) + \
                sum(0.5 * float(((up[d] - um[d]) * (up[d] + um[d])).sum()) for d in range(dim))
            fd = dl / (2 * h)
            worst = max(worst, abs(fd - ad) / abs(ad))
    print(f"  worst FD relative error over 16 directions {worst:.2e}")
    assert worst <= 1e-6
