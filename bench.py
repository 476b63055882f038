#!/usr/bin/env python3
"""MAC-projection benchmark on one MI355X.

Workload: 3D 256^3 plane-parallel MAC projection, 64 boxes of 64^3, x/y
periodic, z slip walls, isothermal beta0 with 1e4 contrast, rho = beta0 (1 +
0.01 N(0,1)), U ~ U[-1,1], solved to 1e-10 relative residual.  One step is
one complete ``macproject`` call (beta0 scaling, divergence, coefficients,
solve, face correction, ghost fill) on device-resident per-box arrays, each
step on its own fresh copy of the inputs (staged in HBM before timing).

``value`` is the mixed-precision path of the configuration (FP32 V-cycle
with FP32 iterate, coefficients and residual inside FP64 iterative
refinement; the finest level's FP32 coefficients are formed in the sweep
from the FP32 density, 2 beta0^2 / (rho_L + rho_R), the coarser levels'
stored; the FP32 red-black sweeps over-relaxed by 1.25, mac_opts.omega) in
solves/s, against the plain FP64 V-cycle (Gauss-Seidel red-black, the
reference algorithm).  Measured the same way and reported beside them: the
mixed path with the library's default Gauss-Seidel sweeps, the FP64 V-cycle
with the same over-relaxed sweeps (what precision alone gains), and the
variant that stores the finest level's FP32-rounded weights and right side
as bfloat16.  Every path stops at the same 1e-10 relative residual;
``accuracy`` reports each against the FP64 path, the oracle at the same
tolerance and the oracle's 1e-13 solve (the over-relaxed iterates are
further from the exact discrete solution: 6.5e-9 in phi against 5.1e-10).  ``roofline`` covers the dominant kernel of the
headline path: the finest-level fused red-black sweep (one full sweep per
launch, the launches that read the old iterate), timed with HIP events on
the library's own stream during the timed steps.  Its ``traffic`` figure is
read from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary
(profiles/pmc_gsrb_sweep_256.txt) when present.  ``cpu_baseline`` is the
FP64 oracle (C/OpenMP restatement) on the same inputs, on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# over-relaxation of the headline mixed path's FP32 sweeps (mac_opts.omega)
OMEGA_HEADLINE = 1.25


def sweep_kernel_name(s: int, b16: bool = False, nx: int = 256, rho: bool = False, sor: bool = False,
                      pz: bool = False) -> str:
    """Name of the finest-level sweep kernel the library launches (b16: the
    compact bfloat16-storage form of the FP32 full-row kernel; rho: the FP32
    form that forms its weights from the density, which the plane-parallel
    mixed path runs).  Full rows,
    W = nx / 128 waves per row; rows x threads per block: FP64 6 x 768, FP32
    6 x 768 with the weights fetched one plane ahead (16 x 1024 at nx = 128),
    FP32 with bfloat16 storage 6 x 768; the template's last two arguments: no
    alpha term, the weights' prefetch distance; the last: over-relaxed (sor).
    The rho form's last argument: z periodic (pz; the workloads here have z
    walls)."""
    t = "double" if s == 8 else "float"
    c = "true" if b16 else "false"
    o = "true" if sor else "false"
    w = max(1, nx // 128)
    if rho:   # FP32, the weights formed from rho: 6 rows x 768 (12 x 768 at nx = 128)
        z = "true" if pz else "false"
        return f"k_gsrb_rows_rho<0, 6, 768, {w}, {o}, {z}>" if w == 2 else f"k_gsrb_rows_rho<0, 12, 768, 1, {o}, {z}>"
    if s == 8 or b16:
        return f"k_gsrb_rows<{t}, 0, 6, 768, {c}, {w}, false, 2, {o}>"
    return (f"k_gsrb_rows<{t}, 0, 16, 1024, {c}, 1, false, 2, {o}>" if w == 1
            else f"k_gsrb_rows<{t}, 0, 6, 768, {c}, {w}, false, 1, {o}>")


def measured_traffic(name: str):
    """HBM bytes per launch of ``name`` from the committed PMC summary:
    lines 'bench-traffic <kernel> <bytes>' (FETCH_SIZE x2 + WRITE_SIZE x1,
    calibrated on a plain copy)."""
    path = os.path.join(ROOT, "profiles", "pmc_gsrb_sweep_256.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("bench-traffic "):
                rest = line[len("bench-traffic "):].rsplit(None, 1)
                if len(rest) == 2 and rest[0].strip() == name:
                    return float(rest[1])
    return None


def gsrb_launch_bytes(ncell: int, s: int, ws: int = None, rho: bool = False) -> float:
    """Algorithmic bytes of one launch of the fused 3D red-black sweep (one
    full sweep): three face weights and rhs (element size ws), the old
    iterate read and the new one written (element size s), each once per
    cell -> 2 s + 4 ws per cell; in the rho form (the weights formed from
    the FP32 density) rho and rhs -> 2 s + 2 ws."""
    ws = s if ws is None else ws
    return (2.0 * s + (2.0 if rho else 4.0) * ws) * ncell


def stage(per_box, copies, torch):
    """Device copies of the per-box inputs: ``copies`` independent sets of
    umac/phi (modified in place by a call) and one shared rho."""
    boxes, umac, phi, rho, _, _ = per_box
    sets = []
    for _ in range(copies):
        du = [[torch.from_numpy(a).cuda() for a in box] for box in umac]
        dp = [torch.from_numpy(a).cuda() for a in phi]
        sets.append((du, dp))
    dr = [torch.from_numpy(a).cuda() for a in rho]
    return sets, dr


def run_precision(plan, sets, dr, b0, b0e, warmup, steps, prec, torch, inner=0, profile=True, omega=0.0):
    """Warm-up and timed projections, one input set per call.  The per-box
    pointer tables are built once per set, outside the timed region, as a
    Fortran or C caller holds its MultiFab's data pointers from step to step."""
    from maestro_amd import mac
    plan.set_opts(precision=prec, device_ptrs=1, inner_cycles=inner, omega=omega)
    tabs = [(mac.PtrTable([a for box in du for a in box]), mac.PtrTable(dp)) for du, dp in sets]
    drt = mac.PtrTable(dr)
    stats = []
    for w in range(warmup):
        du, dp = tabs[w]
        stats.append(plan.project(du, dp, drt, None, b0, b0e, None))
    # per-kernel HIP events bracket the launches of the last timed step only:
    # each bracket stalls the stream for a few microseconds, so instrumenting
    # every step would slow the measured throughput itself
    if profile:   # create the event pool outside the timed region
        plan.profile(True)
    plan.profile(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        if profile and k == steps - 1:
            plan.profile(True)
        du, dp = tabs[warmup + k]
        stats.append(plan.project(du, dp, drt, None, b0, b0e, None))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = plan.profile_read()
    plan.profile(False)
    return (t1 - t0), stats[warmup:], prof


def time_producers(plan, n, box, torch, reps=10):
    """Producers of the projection's inputs on the bench layout (device
    pointers): make_macrhs (planar, dp/dt term on, corrector: reads S,
    delta_gamma1_term, delta_p_term, delta_chi, writes macrhs and delta_chi,
    48 B/cell) and make_s0mac (spherical geometry on the same grid, quadratic
    profile, three face arrays with one ghost layer written, 24 B/face).
    Wall time per call (each call returns after its kernel), median of reps."""
    from maestro_amd import mac
    from maestro_amd import workloads as W
    from maestro_amd import boxes as BX
    boxes = W.box_list(n, box)
    ncell = n[0] * n[1] * n[2]
    cells = lambda ng: [torch.rand(BX.box_shape(lo, hi, ng, 3), dtype=torch.float64, device="cuda")
                        for lo, hi in boxes]
    rhs, S, dg, dp, dchi = cells(0), cells(0), cells(0), cells(0), cells(0)
    one = lambda v: torch.full((n[2],), v, dtype=torch.float64, device="cuda")
    out = {}
    ts = []
    for _ in range(reps + 1):
        torch.cuda.synchronize()
        t = time.perf_counter()
        plan.make_macrhs(rhs, S, dg, one(0.1), one(1.0), delta_p_term=dp, delta_chi=dchi, gamma1bar=one(1.4),
                         p0=one(2.0), dt=0.01, dpdt_factor=0.5, base_cutoff_coord=n[2] // 2, is_predictor=False)
        ts.append(time.perf_counter() - t)
    tm = float(np.median(ts[1:]))
    out["make_macrhs"] = {"us": round(1e6 * tm, 1), "bytes": 48 * ncell,
                          "GB_per_s": round(48 * ncell / tm / 1e9, 1)}
    dx = 1.0 / n[0]
    dr = dx / 5.0
    nr = int(np.sqrt(3) * 0.5 / dr) + 1
    sph = mac.sphr(dr, nr, (0.5, 0.5, 0.5), (0.0, 0.0, 0.0), 3)
    s0 = torch.linspace(1.0, 1e-3, nr, dtype=torch.float64, device="cuda")
    faces = [[torch.empty(BX.box_shape(lo, hi, 1, 3, face_dir=d), dtype=torch.float64, device="cuda")
              for d in range(3)] for lo, hi in boxes]
    nface = sum(f.numel() for bx in faces for f in bx)
    ts = []
    for _ in range(reps + 1):
        torch.cuda.synchronize()
        t = time.perf_counter()
        plan.make_s0mac(s0, sph, faces)
        ts.append(time.perf_counter() - t)
    tm = float(np.median(ts[1:]))
    out["make_s0mac"] = {"us": round(1e6 * tm, 1), "bytes": 8 * nface, "GB_per_s": round(8 * nface / tm / 1e9, 1)}
    return out


def time_helmholtz(plan, n, box, torch, reps=3):
    """mac_solve (mac_multigrid with alpha) on the bench layout: a
    thermal-diffusion-like Helmholtz problem (alpha - D beta G) phi = rh with
    alpha = (0.05/h^2) U[0.5, 2], beta = U[0.5, 2] on faces, x/y periodic, z
    Dirichlet below (values from phi's ghost cells) and Neumann above, rel
    residual 1e-10.  Wall time per solve (device pointers), median of reps,
    FP64 and mixed precision; the same inputs each time."""
    from maestro_amd import mac
    from maestro_amd import workloads as W
    from maestro_amd import boxes as BX
    boxes = W.box_list(n, box)
    h = 1.0 / n[0]
    gen = torch.Generator(device="cuda").manual_seed(11)
    uni = lambda shp, a, b: a + (b - a) * torch.rand(shp, dtype=torch.float64, device="cuda", generator=gen)
    rh = [torch.randn(BX.box_shape(lo, hi, 0, 3), dtype=torch.float64, device="cuda", generator=gen)
          for lo, hi in boxes]
    alpha = [uni(BX.box_shape(lo, hi, 0, 3), 0.05 / h**2 * 0.5, 0.05 / h**2 * 2.0) for lo, hi in boxes]
    beta = [[uni(BX.box_shape(lo, hi, 0, 3, face_dir=d), 0.5, 2.0) for d in range(3)] for lo, hi in boxes]
    phi0 = [uni(BX.box_shape(lo, hi, 1, 3), -1.0, 1.0) for lo, hi in boxes]
    ell = [[mac.ELL_PER, mac.ELL_PER], [mac.ELL_PER, mac.ELL_PER], [mac.ELL_DIR, mac.ELL_NEU]]
    out = {"workload": f"3D {n[0]}^3 Helmholtz (alpha - D beta G) phi = rh, {len(boxes)} boxes, x/y periodic, "
                       "z Dirichlet/Neumann, alpha/(beta/h^2) ~ 0.05, rel residual 1e-10"}
    for name, prec in (("fp64", mac.PREC_FP64), ("mixed", mac.PREC_MIXED)):
        plan.set_opts(precision=prec, device_ptrs=1, inner_cycles=0)
        ts, st = [], None
        for _ in range(reps + 1):
            phi = [a.clone() for a in phi0]
            torch.cuda.synchronize()
            t = time.perf_counter()
            st = plan.solve(rh, phi, beta, alpha, ell_bc=ell, eps_rel=1e-10)
            ts.append(time.perf_counter() - t)
        tm = float(np.median(ts[1:]))
        out[name] = {"ms": round(1e3 * tm, 2), "solves_per_s": round(1.0 / tm, 2), "iters": st["iters"],
                     "res_over_res0": st["res"] / st["res0"]}
    return out


def time_star_512(torch, steps=3, warmup=1):
    """The 512^3 full-star workload (512 boxes of 64^3, n = 1.5 polytrope with
    a density floor, coefficient contrast 1e6, outflow on every side, rel
    residual 1e-10) in mixed precision, V- and F-cycles with Gauss-Seidel and
    with over-relaxed sweeps, plus the FP64 V-cycle.  Wall time per solve (device pointers, fresh input copies
    staged in HBM before each timed solve), median of ``steps``."""
    from maestro_amd import mac
    from maestro_amd import workloads as W
    from maestro_amd import boxes as BX
    n, box = (512,) * 3, 64
    wl = W.make_workload(3, n=n, box=box)
    boxes, umac, phi, rho, _, cart = BX.split_workload(wl)
    cv = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    du0 = [[cv(a) for a in b] for b in umac]
    dp0 = [cv(a) for a in phi]
    dr = [cv(a) for a in rho]
    dc = [[cv(a) for a in b] for b in cart]
    del umac, phi, rho, cart
    lay = mac.Layout.uniform(n, box, wl["dx"], wl["bc"])
    plan = mac.MacPlan(lay, mac.default_opts(device_ptrs=1))
    out = {"workload": "3D 512^3 full star in a box, 512 boxes of 64^3, n = 1.5 polytrope with a density floor "
                       "(contrast 1e6), outflow on every side, rel residual 1e-10"}
    du = [[a.clone() for a in b] for b in du0]
    dp = [a.clone() for a in dp0]
    for name, prec, cyc, w in (("mixed_v", mac.PREC_MIXED, mac.CYCLE_V, 0.0),
                               ("mixed_f", mac.PREC_MIXED, mac.CYCLE_F, 0.0),
                               ("mixed_v_over_relaxed", mac.PREC_MIXED, mac.CYCLE_V, OMEGA_HEADLINE),
                               ("mixed_f_over_relaxed", mac.PREC_MIXED, mac.CYCLE_F, OMEGA_HEADLINE),
                               ("fp64_v", mac.PREC_FP64, mac.CYCLE_V, 0.0)):
        plan.set_opts(precision=prec, cycle_type=cyc, device_ptrs=1, max_iter=200, omega=w)
        tu, tp, tr = mac.PtrTable([a for b in du for a in b]), mac.PtrTable(dp), mac.PtrTable(dr)
        tc = mac.PtrTable([c for b in dc for c in b])
        ts, st = [], None
        for _ in range(warmup + steps):
            for b, b0 in zip(du, du0):
                for a, a0 in zip(b, b0):
                    a.copy_(a0)
            for a, a0 in zip(dp, dp0):
                a.copy_(a0)
            torch.cuda.synchronize()
            t = time.perf_counter()
            st = plan.project(tu, tp, tr, None, None, None, tc)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t)
        tm = float(np.median(ts[warmup:]))
        out[name] = {"ms": round(1e3 * tm, 2), "solves_per_s": round(1.0 / tm, 3), "iters": st["iters"],
                     "fp32_cycles": st["inner_cycles"], "res_over_res0": st["res"] / st["res0"]}
        print(f"# c3 512^3 {name}: {1e3 * tm:.1f} ms, {st['iters']} iterations", file=sys.stderr, flush=True)
    plan.close()
    return out


def time_slab_share(torch, steps=5, warmup=2):
    """One GPU's share of the 256^3 plane-parallel workload split in z over 8
    GPUs: the bottom 256 x 256 x 32 cells of the same inputs (64 boxes of
    32^3), solved on its own (x/y periodic, z walls) to 1e-10.  It
    times the kernels at the slab's real shape, FP64 and the headline mixed
    path, per solve and per finest-level sweep; the halo exchange and the
    all-reduces a sharded solve adds are not run (one GPU)."""
    from maestro_amd import mac
    from maestro_amd import workloads as W
    from maestro_amd import boxes as BX
    n, box, nzs = (256, 256, 32), 32, 32
    # the bottom slab of the 256^3 inputs: cubic cells of 1/256, z walls in
    # place of the slab's halo interfaces
    full = W.make_workload(2, n=(256,) * 3, box=64)
    wl = dict(full)
    wl["n"], wl["box"] = n, box
    wl["faces"] = [np.ascontiguousarray(full["faces"][0][:nzs]), np.ascontiguousarray(full["faces"][1][:nzs]),
                   np.ascontiguousarray(full["faces"][2][:nzs + 1])]
    wl["faces"][2][0] = 0.0
    wl["faces"][2][-1] = 0.0
    wl["rho"] = W._fill_rho_ghosts(full["rho"][1:nzs + 1, 1:-1, 1:-1], full["bc"], 3)
    wl["phi"] = np.zeros((nzs, n[1], n[0]))
    wl["beta0"] = np.ascontiguousarray(full["beta0"][:nzs])
    wl["beta0_edge"] = np.ascontiguousarray(full["beta0_edge"][:nzs + 1])
    del full
    boxes_, umac, phi, rho, _, _ = BX.split_workload(wl)
    cv = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    du0 = [[cv(a) for a in b] for b in umac]
    dp0 = [cv(a) for a in phi]
    dr = [cv(a) for a in rho]
    b0, b0e = cv(wl["beta0"]), cv(wl["beta0_edge"])
    lay = mac.Layout(dim=3, boxes=boxes_, domain_lo=(0, 0, 0), domain_hi=tuple(m - 1 for m in n),
                     dx=tuple(wl["dx"]), phys_bc=wl["bc"])
    plan = mac.MacPlan(lay, mac.default_opts(device_ptrs=1))
    out = {"workload": "one GPU's z-slab of the 256^3 plane-parallel workload over 8 GPUs: 256 x 256 x 32 "
                       "cells of 1/256 (64 boxes of 32^3, the bottom 32 planes of the 256^3 inputs), x/y "
                       "periodic, z walls in place of the halo interfaces, rel residual 1e-10"}
    du = [[a.clone() for a in b] for b in du0]
    dp = [a.clone() for a in dp0]
    for name, prec, w in (("fp64", mac.PREC_FP64, 0.0), ("mixed", mac.PREC_MIXED, OMEGA_HEADLINE)):
        plan.set_opts(precision=prec, device_ptrs=1, omega=w)
        tu, tp, tr = mac.PtrTable([a for b in du for a in b]), mac.PtrTable(dp), mac.PtrTable(dr)
        ts, st = [], None
        for k in range(warmup + steps):
            for b, b0_ in zip(du, du0):
                for a, a0 in zip(b, b0_):
                    a.copy_(a0)
            for a, a0 in zip(dp, dp0):
                a.copy_(a0)
            if k == warmup + steps - 1:
                plan.profile(True)
            torch.cuda.synchronize()
            t = time.perf_counter()
            st = plan.project(tu, tp, tr, None, b0, b0e, None)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t)
        prof = plan.profile_read()
        plan.profile(False)
        cls = "gsrb64" if prec == mac.PREC_FP64 else "gsrb32"
        cnt, ms = prof[cls]
        tm = float(np.median(ts[warmup:]))
        out[name] = {"ms": round(1e3 * tm, 3), "iters": st["iters"], "fp32_cycles": st["inner_cycles"],
                     "finest_sweep_us": round(1e3 * ms / cnt, 2) if cnt else None}
        print(f"# slab 256x256x32 {name}: {1e3 * tm:.2f} ms, {st['iters']} iterations", file=sys.stderr, flush=True)
    plan.close()
    return out


def time_two_level(torch, steps=3, warmup=1):
    """The two-level projection (mac_ml_project): a 128^3 coarse level (64
    boxes of 32^3, x/y periodic, z walls, isothermal beta0 of contrast 1e4)
    with the central 64 x 64 x 64 coarse cells at the bottom wall refined by
    2 (128^3 fine cells in 64 boxes of 32^3), U ~ U[-1,1], rho = beta0 (1 +
    0.01 N(0,1)), to the two-level tolerance min(1e-10 * 10, 1e-8) = 1e-9.
    Device arrays; median of the timed calls."""
    from maestro_amd import mac
    from maestro_amd import workloads as W
    from maestro_amd import boxes as BX
    rng = np.random.default_rng(1234 + 7)
    nc, box = 128, 32
    rlo, rhi = (64, 64, 0), (191, 191, 127)
    nr = [rhi[d] - rlo[d] + 1 for d in range(3)]
    bc = [[W.PERIODIC, W.PERIODIC], [W.PERIODIC, W.PERIODIC], [W.SLIP_WALL, W.SLIP_WALL]]
    lnc = np.log(1e4)
    out = {"workload": "two-level MAC projection: 128^3 coarse (64 boxes of 32^3, x/y periodic, z walls, "
                       "beta0 contrast 1e4) with the central 64x64x64 coarse cells at the bottom wall refined "
                       "by 2 (128^3 fine cells, 64 boxes of 32^3); rel residual 1e-9 (two levels)"}
    levels = []
    for l, (n, org) in enumerate(((nc, (0, 0, 0)), (2 * nc, rlo))):
        h = 1.0 / n
        ext = [nc] * 3 if l == 0 else nr
        zc = (np.arange(n) + 0.5) * h
        ze = np.arange(n + 1) * h
        b0, b0e = np.exp(-lnc * zc), np.exp(-lnc * ze)
        faces = []
        for d in range(3):
            shp = [ext[2], ext[1], ext[0]]
            shp[2 - d] += 1
            u = rng.uniform(-1, 1, shp)
            if d == 2 and org[2] == 0:
                u[0] = 0.0
            if d == 2 and l == 0:
                u[-1] = 0.0
            faces.append(u)
        zg = (np.arange(-1, ext[2] + 1) + org[2] + 0.5) * h
        rho = np.exp(-lnc * np.clip(zg, 0.5 * h, 1 - 0.5 * h))[:, None, None] * \
            (1 + 0.01 * rng.standard_normal((ext[2] + 2, ext[1] + 2, ext[0] + 2)))
        if l == 0:
            rho = W._fill_rho_ghosts(rho[1:-1, 1:-1, 1:-1], bc, 3)
        levels.append((ext, org, faces, rho, b0, b0e))
    cv = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    lays, args = [], []
    for l, (ext, org, faces, rho, b0, b0e) in enumerate(levels):
        rel = W.box_list(ext, box)
        boxes = [([lo[d] + org[d] for d in range(3)], [hi[d] + org[d] for d in range(3)]) for lo, hi in rel]
        sbc = bc if l == 0 else [[W.SLIP_WALL, W.SLIP_WALL]] * 3
        du = [[cv(BX.split_faces(faces[d], d, lo, hi, 1, 3, sbc)) for d in range(3)] for lo, hi in rel]
        dp = [cv(np.zeros(BX.box_shape(lo, hi, 1, 3))) for lo, hi in rel]
        dr = [cv(BX.split_cells(rho, lo, hi, 1, 3, sbc, glob_ghost=1)) for lo, hi in rel]
        n = nc * (1 + l)
        lays.append(mac.Layout(dim=3, boxes=boxes, domain_lo=(0, 0, 0), domain_hi=(n - 1,) * 3,
                               dx=(1.0 / n,) * 3, phys_bc=bc))
        args.append((du, dp, dr, cv(b0), cv(b0e)))
    u0 = [[[a.clone() for a in b] for b in x[0]] for x in args]
    out["cells"] = nc ** 3 + int(np.prod(nr))
    # FP64 inner V-cycles (the headline of this leg), then FP32 inner cycles
    # inside FP64 refinement: the composite residual is FP64 in both
    for name, prec in (("fp64", mac.PREC_FP64), ("mixed", mac.PREC_MIXED)):
        plan = mac.MacMLPlan(lays, mac.default_opts(eps_rel=1e-9, precision=prec))
        ts, st = [], None
        for k in range(warmup + steps):
            for l in range(2):
                for b, b_ in zip(args[l][0], u0[l]):
                    for a, a0 in zip(b, b_):
                        a.copy_(a0)
                for a in args[l][1]:
                    a.zero_()
            torch.cuda.synchronize()
            t = time.perf_counter()
            st = plan.project([x[0] for x in args], [x[1] for x in args], [x[2] for x in args], None,
                              [x[3] for x in args], [x[4] for x in args])
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t)
        plan.close()
        tm = float(np.median(ts[warmup:]))
        r = {"ms": round(1e3 * tm, 3), "composite_steps": st["iters"], "inner_vcycles": st["inner_cycles"],
             "res_over_res0": st["res"] / st["res0"], "setup_ms": round(st["t_setup_ms"], 3),
             "solve_ms": round(st["t_solve_ms"], 3), "correct_ms": round(st["t_correct_ms"], 3)}
        if name == "fp64":
            out.update(r)
        else:
            out["mixed_inner"] = r
        print(f"# two-level 128^3 + 128^3 {name}: {1e3 * tm:.2f} ms, {st['iters']} composite steps",
              file=sys.stderr, flush=True)
    return out


def time_adjoint_256(torch, steps=3, warmup=1):
    """The 256^3 differentiable workload: C1's inputs, L = sum(w phi) +
    1/2 ||U_p||^2 with w ~ N(0, 1), mixed-precision forward (the projection
    through the torch.autograd.Function) and backward (the adjoint solve and
    the per-face gradient kernel), rel residual 1e-10.  The face coefficients
    beta0^2/rho are formed in torch from the workload's rho.  Wall time of
    each, median of ``steps``."""
    from maestro_amd import mac
    from maestro_amd import workloads as W
    from maestro_amd import autograd as AG
    n = (256,) * 3
    wl = W.make_workload(4, n=n, box=64)
    cv = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    rho = cv(wl["rho"])   # one ghost layer
    b0, b0e = cv(wl["beta0"]), cv(wl["beta0_edge"])
    b0f = [b0[:, None, None], b0[:, None, None], b0e[:, None, None]]
    # face beta = 2 / (rho_L + rho_R) beta0^2 on every face (the walls' normal
    # faces carry no flux: their velocity is zero)
    bx = 2.0 / (rho[1:-1, 1:-1, :-1] + rho[1:-1, 1:-1, 1:]) * b0f[0] ** 2
    by = 2.0 / (rho[1:-1, :-1, 1:-1] + rho[1:-1, 1:, 1:-1]) * b0f[1] ** 2
    bz = 2.0 / (rho[:-1, 1:-1, 1:-1] + rho[1:, 1:-1, 1:-1]) * b0f[2] ** 2
    B = [bx.contiguous(), by.contiguous(), bz.contiguous()]
    U = [cv(f) for f in wl["faces"]]
    w = torch.from_numpy(np.random.default_rng(4321).standard_normal((n[2], n[1], n[0]))).cuda()
    plan = AG.level_plan(n, wl["dx"], wl["bc"], precision=mac.PREC_MIXED, eps_rel=1e-10)
    tf, tb = [], []
    for _ in range(warmup + steps):
        Ul = [u.clone().requires_grad_(True) for u in U]
        Bl = [b.clone().requires_grad_(True) for b in B]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        Us = [Ul[d] * b0f[d] for d in range(3)]
        phi, ux, uy, uz = AG.mac_project(plan, Us[0], Us[1], Us[2], Bl[0], Bl[1], Bl[2])
        Up = [ux / b0f[0], uy / b0f[1], uz / b0f[2]]
        L = (w * phi).sum() + 0.5 * sum((u * u).sum() for u in Up)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        L.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        tf.append(t1 - t0)
        tb.append(t2 - t1)
    plan.close()
    f, b = float(np.median(tf[warmup:])), float(np.median(tb[warmup:]))
    print(f"# c4 256^3 adjoint: forward {1e3 * f:.1f} ms, backward {1e3 * b:.1f} ms", file=sys.stderr, flush=True)
    return {"workload": "3D 256^3 differentiable MAC projection (C1 inputs), L = sum(w phi) + 1/2 ||U_p||^2, "
                        "mixed precision, rel residual 1e-10",
            "forward_ms": round(1e3 * f, 2), "backward_ms": round(1e3 * b, 2),
            "forward_plus_backward_per_s": round(1.0 / (f + b), 3)}


def time_host_pointers(plan, pb, wl, prec, omega=0.0):
    """One projection with host (numpy) per-box arrays: the library copies
    them in and out around the solve (pageable host memory over PCIe).
    Returns the split of the second of two calls (the first allocates the
    device mirrors)."""
    from maestro_amd import mac
    boxes, umac, phi, rho, _, _ = pb
    plan.set_opts(precision=prec, device_ptrs=0, omega=omega)
    b0 = np.ascontiguousarray(wl["beta0"])
    b0e = np.ascontiguousarray(wl["beta0_edge"])
    st = None
    for _ in range(2):
        u = [[a.copy() for a in box] for box in umac]
        p = [a.copy() for a in phi]
        t = time.perf_counter()
        st = plan.project(u, p, rho, None, b0, b0e, None)
        wall = time.perf_counter() - t
    nbytes_in = sum(a.nbytes for box in umac for a in box) + sum(a.nbytes for a in phi) + \
        sum(a.nbytes for a in rho)
    nbytes_out = sum(a.nbytes for box in umac for a in box) + sum(a.nbytes for a in phi)
    plan.set_opts(device_ptrs=1, omega=0.0)
    return {"ms_total": round(1e3 * wall, 2), "solves_per_s": round(1.0 / wall, 2),
            "h2d_ms": round(st["t_h2d_ms"], 2), "solve_ms": round(st["t_setup_ms"] + st["t_solve_ms"] +
                                                              st["t_correct_ms"], 2),
            "d2h_ms": round(st["t_d2h_ms"], 2), "h2d_bytes": nbytes_in, "d2h_bytes": nbytes_out}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--box", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-solves", type=int, default=5,
                    help="timed oracle solves of the CPU baseline (median; after one warm-up)")
    ap.add_argument("--no-check", action="store_true", help="skip the accuracy check")
    ap.add_argument("--no-profile", action="store_true",
                    help="time the steps without the per-kernel HIP events (no roofline figures)")
    ap.add_argument("--no-producers", action="store_true",
                    help="skip timing make_macrhs / make_s0mac and the Helmholtz mac_solve on the bench layout")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip timing the 512^3 star (mixed V/F, FP64), the 256^3 adjoint, the slab share "
                         "and the two-level projection")
    ap.add_argument("--headline", default="mixed", choices=("mixed", "mixed_gauss_seidel", "mixed_bf16"),
                    help="the mixed path whose throughput is `value`")
    ap.add_argument("--inner-cycles", type=int, default=0,
                    help="FP32 V-cycles per refinement step of the mixed paths "
                         "(0: the library's default, 3 with FP32 storage, 2 with bfloat16 storage)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy 5 times (the PMC byte-counter "
                         "calibration: 1 GiB read + 1 GiB written per launch)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the sharding across GPUs is a design, not run)", file=sys.stderr)
        sys.exit(2)

    import torch
    if args.calibrate_copy:
        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "copy", "bytes_read": a.numel() * 8,
                          "bytes_written": a.numel() * 8, "launches": 5}))
        return
    from maestro_amd import mac
    from maestro_amd import workloads as W
    from maestro_amd import boxes as BX
    per_box_inputs = BX.split_workload

    n = (args.n,) * 3
    t_gen = time.perf_counter()
    wl = W.make_workload(2, n=n, box=args.box)
    pb = per_box_inputs(wl)
    print(f"# inputs ready in {time.perf_counter() - t_gen:.1f}s", file=sys.stderr, flush=True)
    copies = args.warmup + args.steps
    sets, dr = stage(pb, copies, torch)
    b0 = torch.from_numpy(np.ascontiguousarray(wl["beta0"])).cuda()
    b0e = torch.from_numpy(np.ascontiguousarray(wl["beta0_edge"])).cuda()
    lay = mac.Layout.uniform(n, args.box, wl["dx"], wl["bc"])
    plan = mac.MacPlan(lay, mac.default_opts(device_ptrs=1))
    ncell = n[0] * n[1] * n[2]

    results = {}
    # (name, precision, omega): 0 = the default, Gauss-Seidel red-black
    paths = (("fp64", mac.PREC_FP64, 0.0), ("mixed", mac.PREC_MIXED, OMEGA_HEADLINE),
             ("mixed_gauss_seidel", mac.PREC_MIXED, 0.0), ("mixed_bf16", mac.PREC_MIXED_BF16, 0.0),
             ("fp64_over_relaxed", mac.PREC_FP64, OMEGA_HEADLINE))
    for name, prec, omega in paths:
        # refresh the staged inputs from the host copies (outside timing)
        boxes, umac, phi, _, _, _ = pb
        for du, dp in sets:
            for box_d, box_h in zip(du, umac):
                for a, h in zip(box_d, box_h):
                    a.copy_(torch.from_numpy(h))
            for a, h in zip(dp, phi):
                a.copy_(torch.from_numpy(h))
        torch.cuda.synchronize()
        dt, st, prof = run_precision(plan, sets, dr, b0, b0e, args.warmup, args.steps, prec, torch,
                                     args.inner_cycles, profile=not args.no_profile, omega=omega)
        # host snapshot of the last step's outputs (the staged buffers are reused)
        results[name] = {"seconds": dt, "stats": st, "prof": prof,
                         "umac": [[a.cpu().numpy() for a in box] for box in sets[-1][0]],
                         "phi": [a.cpu().numpy() for a in sets[-1][1]]}
        print(f"# {name}: {args.steps / dt:.2f} solves/s, {1e3 * dt / args.steps:.2f} ms/solve, "
              f"iters {[s['iters'] for s in st]}, res/res0 {st[-1]['res'] / st[-1]['res0']:.2e}",
              file=sys.stderr, flush=True)
    plan.set_opts(omega=0.0)

    def sweep_roofline(path):
        """Roofline of the finest-level sweep of one path (the launches that
        read the old iterate)."""
        s_, ws, b16, rho, cls, sor = {
            "fp64": (8, 8, False, False, "gsrb64", False), "mixed": (4, 4, False, True, "gsrb32", True),
            "mixed_gauss_seidel": (4, 4, False, True, "gsrb32", False),
            "mixed_bf16": (4, 2, True, False, "gsrb32", False),
            "fp64_over_relaxed": (8, 8, False, False, "gsrb64", True)}[path]
        cnt_, ms_ = results[path]["prof"][cls]
        avg = ms_ / cnt_ if cnt_ else float("inf")   # --no-profile: no per-kernel timing
        nbytes = gsrb_launch_bytes(ncell, s_, ws, rho)
        ach = nbytes / (avg * 1e-3) / 1e9
        name = sweep_kernel_name(s_, b16, n[0], rho, sor)
        return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": measured_traffic(name),
                "kernel": name + " (finest level, one full red-black sweep)",
                "avg_launch_us": round(1e3 * avg, 2), "launches": cnt_, "bytes_per_launch": nbytes}

    producers = None if args.no_producers else time_producers(plan, n, args.box, torch)
    helmholtz = None if args.no_producers else time_helmholtz(plan, n, args.box, torch)
    host = None if args.no_producers else time_host_pointers(
        plan, pb, wl, *{"mixed": (mac.PREC_MIXED, OMEGA_HEADLINE), "mixed_gauss_seidel": (mac.PREC_MIXED, 0.0),
                        "mixed_bf16": (mac.PREC_MIXED_BF16, 0.0)}[args.headline])

    c3 = None if args.no_configs else time_star_512(torch)
    c4 = None if args.no_configs else time_adjoint_256(torch)
    slab = None if args.no_configs else time_slab_share(torch)
    two_level = None if args.no_configs else time_two_level(torch)

    head = args.headline
    mixed, fp64 = results[head], results["fp64"]
    value = args.steps / mixed["seconds"]

    # accuracy of the mixed path against the FP64 GPU path and the oracle
    dim = 3
    nn = list(n)
    boxes = pb[0]

    def gather(res):
        faces = [BX.gather_faces([u[d] for u in res["umac"]], boxes, d, 1, nn, dim)
                 for d in range(dim)]
        ph = BX.gather_cells(res["phi"], boxes, 1, nn, dim)
        return faces, ph

    acc = {}
    if not args.no_check:
        f6, p6 = gather(fp64)

        def errs(fa, pa, fb, pb_):
            pa0, pb0 = pa - pa.mean(), pb_ - pb_.mean()
            return {
                "phi_max": float(np.abs(pa0 - pb0).max() / np.abs(pb0).max()),
                "phi_rel": float(np.linalg.norm(pa0 - pb0) / np.linalg.norm(pb0)),
                "u_max": float(max(np.abs(fa[d] - fb[d]).max() / np.abs(fb[d]).max() for d in range(dim))),
                "u_rel": float(max(np.linalg.norm(fa[d] - fb[d]) / np.linalg.norm(fb[d]) for d in range(dim))),
            }
        for path in ("mixed", "mixed_gauss_seidel", "mixed_bf16", "fp64_over_relaxed"):
            fm, pm = gather(results[path])
            acc[f"{path}_vs_fp64_gpu"] = errs(fm, pm, f6, p6)

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        O.build()
        g = O.make_grid(wl["n"], wl["dx"], wl["bc"])
        tcs = []
        for _ in range(1 + args.cpu_solves):   # one warm-up, then the timed solves
            t = time.perf_counter()
            ref = O.macproject(g, wl["faces"], wl["phi"], wl["rho"], beta0=wl["beta0"],
                               beta0_edge=wl["beta0_edge"], opts=O.default_opts())
            tcs.append(time.perf_counter() - t)
        tc = float(np.median(tcs[1:]))
        cpu = {"value": 1.0 / tc, "unit": "solves/s", "cores": O.num_threads(), "kind": "port",
               "sample": f"full 256^3 FP64 solves of the same inputs by the C/OpenMP oracle (the build's "
                         f"CPU restatement of macproject and the FBoxLib-documented V-cycle), "
                         f"{ref['stats'].iters} V-cycles each: median of {args.cpu_solves} after one warm-up, "
                         f"{tc:.2f} s (all: {', '.join('%.2f' % x for x in tcs[1:])} s)"}
        if not args.no_check:
            # against the oracle at the same tolerance, and against its 1e-13
            # solve (the exact discrete solution to rounding)
            exact = O.macproject(g, wl["faces"], wl["phi"], wl["rho"], beta0=wl["beta0"],
                                 beta0_edge=wl["beta0_edge"], opts=O.default_opts(eps_rel=1e-13))
            for which, r in (("oracle", ref), ("oracle_1e-13", exact)):
                fo, po = r["faces"], r["phi"] - r["phi"].mean()
                for path in ("fp64", "mixed", "mixed_gauss_seidel", "mixed_bf16", "fp64_over_relaxed"):
                    fa, pa = gather(results[path])
                    pa0 = pa - pa.mean()
                    acc[f"{path}_vs_{which}"] = {
                        "phi_max": float(np.abs(pa0 - po).max() / np.abs(po).max()),
                        "phi_rel": float(np.linalg.norm(pa0 - po) / np.linalg.norm(po)),
                        "u_max": float(max(np.abs(fa[d] - fo[d]).max() / np.abs(fo[d]).max() for d in range(dim))),
                        "u_rel": float(max(np.linalg.norm(fa[d] - fo[d]) / np.linalg.norm(fo[d]) for d in range(dim))),
                    }
        print(f"# cpu oracle: {tc:.2f} s/solve on {O.num_threads()} threads", file=sys.stderr)

    storage = {"mixed": "FP32 V-cycle with red-black sweeps over-relaxed by 1.25 (mac_opts.omega) inside FP64 IR, "
                        "finest-level FP32 coefficients formed from the FP32 density; against the plain FP64 V-cycle",
               "mixed_gauss_seidel": "FP32 V-cycle (Gauss-Seidel red-black, the library's default) inside FP64 IR, "
                                     "finest-level FP32 coefficients formed from the FP32 density",
               "mixed_bf16": "FP32 V-cycle inside FP64 IR, finest-level weights and rhs stored as bfloat16"}
    line = {
        "metric": "MAC solves/s at 256^3 to 1e-10 resid: FP32-MG+FP64 IR vs FP64; HBM GB/s",
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * mixed["seconds"] / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded isothermal stratification, 1% density noise, U~U[-1,1])",
        "config": {"workload": "3D 256^3 plane-parallel MAC projection, 64 boxes of 64^3, "
                               "x/y periodic, z walls, beta0 contrast 1e4, rel residual 1e-10",
                   "path": storage[head],
                   "cells": ncell, "boxes": len(boxes), "parallelism": "single GPU"},
        "fp64_solves_per_s": round(args.steps / fp64["seconds"], 3),
        "fp64_ms_per_step": round(1e3 * fp64["seconds"] / args.steps, 3),
        "solves_per_s": {p: round(args.steps / results[p]["seconds"], 3) for p in results},
        "mixed_over_fp64": round(fp64["seconds"] / mixed["seconds"], 3),
        "compact_bf16_over_fp64": round(fp64["seconds"] / results["mixed_bf16"]["seconds"], 3),
        # what precision alone gains: mixed against FP64 with the same sweeps
        "mixed_over_fp64_same_smoother": {
            "over_relaxed": round(results["fp64_over_relaxed"]["seconds"] / results["mixed"]["seconds"], 3),
            "gauss_seidel": round(fp64["seconds"] / results["mixed_gauss_seidel"]["seconds"], 3)},
        "iters": {"fp64_vcycles": fp64["stats"][-1]["iters"],
                  "fp64_over_relaxed_vcycles": results["fp64_over_relaxed"]["stats"][-1]["iters"],
                  **{f"{p}_ir_steps": results[p]["stats"][-1]["iters"]
                     for p in ("mixed", "mixed_gauss_seidel", "mixed_bf16")},
                  **{f"{p}_fp32_vcycles": results[p]["stats"][-1]["inner_cycles"]
                     for p in ("mixed", "mixed_gauss_seidel", "mixed_bf16")}},
        "roofline": sweep_roofline(head),
        "roofline_other": {p: sweep_roofline(p) for p in results if p != head},
        "kernel_ms_per_solve": {
            p: {k: round(v[1], 3) for k, v in results[p]["prof"].items()} for p in results},
        "accuracy": acc,
        "producers": producers,
        "helmholtz": helmholtz,
        "host_pointers": host,
        "cpu_baseline": cpu,
        "c3_star_512": c3,
        "c4_adjoint_256": c4,
        "slab_share_256x256x32": slab,
        "two_level_128": two_level,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
