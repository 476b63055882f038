"""GPU parity: the HIP MAC projection (through the C ABI) against the FP64
oracle on the same seeded inputs.

Tolerances (the parity contract): phi (mean-removed where the problem is
singular) and U_proj within 1e-8 relative (2-norm) of the oracle solved to
1e-13; element-wise kernels bitwise or within 1e-14.
"""
import numpy as np
import pytest

from maestro_amd import workloads as W

from .helpers import rel_err, run_gpu, run_oracle

pytestmark = pytest.mark.gpu

CASES = [
    # (workload, size, box, alt, precision)
    (0, (64, 64), 16, 1, 0),
    (0, (64, 64), 16, 0, 0),
    (1, (32, 32, 32), 16, 1, 0),
    (1, (32, 32, 32), 16, 0, 0),
    (2, (32, 32, 32), 16, 1, 1),
    (3, (32, 32, 32), 16, 1, 0),
    (3, (32, 32, 32), 16, 0, 1),
    (1, (64, 64, 32), 32, 1, 1),
    # nx a multiple of 128: the full-row FP32 smoother and the row-march FP64 residual passes
    (1, (128, 64, 64), 32, 1, 0),
    (2, (128, 64, 64), 32, 1, 1),
    (2, (256, 64, 64), 64, 0, 1),
    (3, (128, 64, 64), 32, 0, 1),
    # bfloat16 storage of the finest level's weights and rhs (compact form
    # for nx = 128, 256; FP32 storage for nx = 64 and in 2D)
    (1, (128, 64, 64), 32, 1, 2),
    (2, (256, 64, 64), 64, 1, 2),
    (3, (128, 64, 64), 32, 0, 2),
    (3, (128, 128, 128), 64, 1, 2),
    (2, (64, 64, 32), 32, 1, 2),
    (0, (64, 64), 16, 1, 2),
    # extents that are not powers of two (coarsening stops at an odd extent;
    # the bottom level is larger than two cells per direction)
    (1, (96, 80, 48), 16, 1, 0),
    (3, (96, 80, 48), 16, 0, 1),
    (0, (96, 80), 16, 1, 0),
]


@pytest.mark.parametrize("cfg,n,box,alt,prec", CASES)
def test_macproject_matches_oracle(gpu, cfg, n, box, alt, prec):
    wl = W.make_workload(cfg, n=n, box=box)
    truth = run_oracle(wl, eps=1e-13, alt=alt)
    same = run_oracle(wl, eps=1e-10, alt=alt)      # the oracle at the same tolerance
    out = run_gpu(wl, use_alt_energy_fix=alt, precision=prec)
    st = out["stats"]
    print(f"cfg{cfg} n={n} alt={alt} prec={prec} iters={st['iters']} "
          f"(oracle {same['stats'].iters}) res0={st['res0']:.3e} res={st['res']:.3e}")
    assert st["converged"]
    assert st["res"] <= 1e-10 * st["res0"] * (1 + 1e-12)
    singular = all(b[0] != W.OUTLET for b in wl["bc"])
    dim = wl["dim"]
    e_phi = rel_err(out["phi"], truth["phi"], remove_mean=singular)
    e_u = max(rel_err(out["faces"][d], truth["faces"][d]) for d in range(dim))
    # error of the reference algorithm's own 1e-10 solve against the truth
    e_u_ref = max(rel_err(same["faces"][d], truth["faces"][d]) for d in range(dim))
    e_phi_ref = rel_err(same["phi"], truth["phi"], remove_mean=singular)
    print(f"  vs 1e-13 truth: phi {e_phi:.2e}  U {e_u:.2e} "
          f"(oracle@1e-10: phi {e_phi_ref:.2e} U {e_u_ref:.2e})")
    if prec == 2:
        # bfloat16 weights perturb the inner operator by up to 2^-9, so the
        # refinement stops at a different 1e-10 iterate than the reference;
        # its error is held to that of the reference's own 1e-10 solve, and
        # the extra refinement steps are bounded
        assert e_phi <= max(1e-8, 2.0 * e_phi_ref)
        assert st["iters"] <= 1.5 * same["stats"].iters + 1
    else:
        assert e_phi <= 1e-8
    assert e_u <= max(1e-8, 2.0 * e_u_ref)
    if prec == 0:
        # FP64 path: the same V-cycle as the oracle, so the 1e-10 iterates agree
        # up to the bottom solver's summation order
        e_same_phi = rel_err(out["phi"], same["phi"], remove_mean=singular)
        e_same_u = max(rel_err(out["faces"][d], same["faces"][d]) for d in range(dim))
        print(f"  vs oracle@1e-10: phi {e_same_phi:.2e}  U {e_same_u:.2e}")
        assert abs(st["iters"] - same["stats"].iters) <= 1
        assert e_same_phi <= 1e-9 and e_same_u <= 1e-9


# the second size moves about 195 MiB in and 157 MiB out through the two
# 64 MiB pinned slots of the host-pointer path (four and three slot loads, so
# each slot is reused); the first copies each array directly
@pytest.mark.parametrize("n,box", [((32, 32, 32), 16), ((256, 256, 64), 32)])
def test_host_and_device_pointers_agree(gpu, n, box):
    wl = W.make_workload(1, n=n, box=box)
    a = run_gpu(wl, device=False)
    b = run_gpu(wl, device=True)
    for d in range(3):
        assert np.array_equal(a["faces"][d], b["faces"][d])
    assert np.array_equal(a["phi"], b["phi"])


@pytest.mark.parametrize("cfg,macrhs", [(1, False), (3, False), (2, True)])
def test_pointer_tables_agree_with_lists(gpu, cfg, macrhs):
    """Per-box device arrays passed as prebuilt pointer tables give the
    projection of the same arrays passed as lists, bit for bit."""
    wl = W.make_workload(cfg, n=(32, 32, 32), box=16)
    rhs = np.random.default_rng(5).standard_normal(wl["n"][::-1]) if macrhs else None
    a = run_gpu(wl, device=True, macrhs=rhs)
    b = run_gpu(wl, device=True, macrhs=rhs, tables=True)
    for d in range(3):
        assert np.array_equal(a["faces"][d], b["faces"][d])
    assert np.array_equal(a["phi"], b["phi"])


def test_array_counts_are_checked(gpu):
    """A list or pointer table with the wrong number of arrays for the
    layout's boxes, or a call mixing device and host arrays, is refused
    before the library reads it."""
    from maestro_amd import mac
    from .helpers import layout_for, per_box_inputs
    wl = W.make_workload(1, n=(32, 32, 32), box=16)
    boxes, umac, phi, rho, _, _ = per_box_inputs(wl)
    plan = mac.MacPlan(layout_for(wl, ng_umac=1, ng_macrhs=0), mac.default_opts())
    b0, b0e = np.ascontiguousarray(wl["beta0"]), np.ascontiguousarray(wl["beta0_edge"])
    with pytest.raises(ValueError):
        plan.project(umac[:-1], phi, rho, None, b0, b0e, None)
    with pytest.raises(ValueError):
        plan.project(umac, phi, rho[:-1], None, b0, b0e, None)
    with pytest.raises(ValueError):
        plan.project([box[:2] for box in umac], phi, rho, None, b0, b0e, None)
    with pytest.raises(ValueError):
        plan.project(mac.PtrTable([a for box in umac for a in box][:-1]), phi, rho, None, b0, b0e, None)
    # device faces with host phi, rho and beta0: every pointer of a call is of one kind
    du = [[gpu.from_numpy(a).cuda() for a in box] for box in umac]
    with pytest.raises(ValueError):
        plan.project(du, phi, rho, None, b0, b0e, None)
    with pytest.raises(ValueError):
        plan.project(mac.PtrTable([a for box in du for a in box]), phi, rho, None, b0, b0e, None)
    plan.close()


def test_deterministic(gpu):
    wl = W.make_workload(1, n=(32, 32, 32), box=16)
    a = run_gpu(wl)
    b = run_gpu(wl)
    for d in range(3):
        assert np.array_equal(a["faces"][d], b["faces"][d])
    assert np.array_equal(a["phi"], b["phi"])


@pytest.mark.parametrize("cfg,n,box,prec", [(1, (128, 64, 64), 32, 0), (2, (256, 64, 64), 64, 1),
                                            (3, (128, 64, 64), 32, 0), (3, (128, 64, 64), 32, 1),
                                            (1, (32, 32, 32), 16, 1)])
def test_initial_residual_norm_bitwise(gpu, cfg, n, box, prec):
    """The solver's first residual norm (FP64 in both paths) against the
    oracle's for the same warm start: bit for bit where the problem is not
    singular (setup and the residual pass evaluate in the oracle's operand
    order); within 1e-12 where the mean of rh is removed first (the two
    sides sum it in different orders)."""
    wl = W.make_workload(cfg, n=n, box=box)
    ref = run_oracle(wl, eps=1e-10, alt=1, max_iter=0)
    out = run_gpu(wl, use_alt_energy_fix=1, precision=prec, max_iter=0, check=False)
    got, want = out["stats"]["res0"], ref["stats"].res0
    singular = all(b[0] != W.OUTLET for b in wl["bc"])
    if singular:
        assert abs(got - want) <= 1e-12 * want, (got, want)
    else:
        assert got == want, (got, want)


@pytest.mark.parametrize("ng", [1, 2])
def test_ghost_faces_filled_from_neighbours(gpu, ng):
    """After the call, every ghost face of every box that lies inside the
    domain or across a periodic side holds the valid face of the box that
    owns it (multifab_fill_boundary); ghost faces outside non-periodic sides
    belong to the physical edge-velocity rules and are not checked here."""
    wl = W.make_workload(1, n=(32, 32, 32), box=16)
    out = run_gpu(wl, ng_u=ng)
    n = wl["n"]
    per = [b[0] == W.PERIODIC for b in wl["bc"]]
    checked = 0
    for b, (lo, hi) in enumerate(out["boxes"]):
        for d in range(3):
            arr = out["umac_boxes"][b][d]
            ez, ey, ex = arr.shape
            k, j, i = np.meshgrid(np.arange(ez), np.arange(ey), np.arange(ex), indexing="ij")
            g = [lo[0] - ng + i, lo[1] - ng + j, lo[2] - ng + k]
            valid = np.ones(arr.shape, dtype=bool)
            inside = np.ones(arr.shape, dtype=bool)
            G = []
            for e in range(3):
                top_box = hi[e] + (1 if e == d else 0)
                valid &= (g[e] >= lo[e]) & (g[e] <= top_box)
                top = n[e] - 1 + (1 if e == d else 0)
                out_of = (g[e] < 0) | (g[e] > top)
                if per[e]:
                    G.append(np.mod(g[e], n[e]))
                else:
                    inside &= ~out_of
                    G.append(np.clip(g[e], 0, top))
            sel = ~valid & inside
            want = out["faces"][d][G[2][sel], G[1][sel], G[0][sel]]
            assert np.array_equal(arr[sel], want), (b, d)
            checked += int(sel.sum())
    assert checked > 0


def test_physical_edge_velocity_rules(gpu):
    """The ghost faces outside non-periodic sides after the call follow the
    edge-velocity rules of the reference's sample multifab_physbc_edgevel
    (Exec/SCIENCE/flame/multifab_physbc_edgevel.f90:211-321), one ghost
    layer: inlet / slip / no-slip walls zero the tangential ghost faces and
    leave the normal one (the inlet velocity is problem data); an outlet
    copies the adjacent face (the normal ghost beyond the boundary face at the
    high side); symmetry mirrors the normal component with a sign change and
    copies the tangential ones.  Checked away from edges and corners, where
    the sides' fills overlap in the reference's order."""
    rng = np.random.default_rng(3)
    n, box = (32, 32, 32), 16
    bc = [[W.INLET, W.OUTLET], [W.SYMMETRY, W.NO_SLIP_WALL], [W.SLIP_WALL, W.SYMMETRY]]
    wl = W.make_workload(1, n=n, box=box)
    wl["bc"] = bc
    wl["faces"] = W._random_faces(rng, n, 3, bc)
    rho_in = wl["rho"][1:-1, 1:-1, 1:-1]
    wl["rho"] = W._fill_rho_ghosts(rho_in, bc, 3)
    from maestro_amd import boxes as BX
    _, umac_in, _, _, _, _ = BX.split_workload(wl)
    out = run_gpu(wl)
    assert out["stats"]["converged"]
    checked = 0
    for b, (lo, hi) in enumerate(out["boxes"]):
        for comp in range(3):
            arr, inp = out["umac_boxes"][b][comp], umac_in[b][comp]

            def at(a, g):   # global face index (i, j, k) of component comp -> array element
                return a[g[2] - lo[2] + 1, g[1] - lo[1] + 1, g[0] - lo[0] + 1]

            for dr in range(3):
                for side in range(2):
                    if side == 0 and lo[dr] != 0 or side == 1 and hi[dr] != n[dr] - 1:
                        continue
                    phys, normal = bc[dr][side], comp == dr
                    if side == 0:
                        tgt = lo[dr] - 1
                        src = {W.OUTLET: lo[dr], W.SYMMETRY: lo[dr] + 1 if normal else lo[dr]}.get(phys)
                    else:
                        tgt = hi[dr] + 2 if normal and phys in (W.OUTLET, W.SYMMETRY) else hi[dr] + 1
                        src = {W.OUTLET: hi[dr] + 1 if normal else hi[dr], W.SYMMETRY: hi[dr]}.get(phys)
                    sign = -1.0 if phys == W.SYMMETRY and normal else 1.0
                    t = [e for e in range(3) if e != dr]
                    r0 = range(lo[t[0]], hi[t[0]] + 1 + (t[0] == comp))
                    r1 = range(lo[t[1]], hi[t[1]] + 1 + (t[1] == comp))
                    for a0 in r0:
                        for a1 in r1:
                            g = [0, 0, 0]
                            g[t[0]], g[t[1]], g[dr] = a0, a1, tgt
                            got = at(arr, g)
                            if phys in (W.INLET, W.SLIP_WALL, W.NO_SLIP_WALL):
                                want = at(inp, g) if normal else 0.0
                            else:
                                gs = list(g)
                                gs[dr] = src
                                want = sign * at(arr, gs)
                            assert got == want, (b, comp, dr, side, g)
                            checked += 1
    assert checked > 0


@pytest.mark.parametrize("cycle", [1, 2])
@pytest.mark.parametrize("cfg,n,box,prec", [
    (1, (64, 64, 64), 32, 0),      # three levels of separate launches above the coarse cycle
    (3, (32, 32, 32), 16, 0),
    (0, (64, 64), 16, 0),          # 2D: the two-colour kernels
    (2, (128, 64, 64), 32, 1),
    (3, (128, 64, 64), 32, 2),
    # the star at sizes where the V-cycle diverges (8^3) or crawls (16^3):
    # every level inside the single-workgroup coarse cycle
    (3, (16, 16, 16), 8, 0),
    (3, (8, 8, 8), 8, 0),
])
def test_w_and_f_cycles_match_oracle(gpu, cycle, cfg, n, box, prec):
    """mg_cycle_type 1 (F) and 2 (W): the GPU cycle against the oracle's
    cycle of the same shape (FP64: the same iterates to the bottom solver's
    summation order), and the answer against the 1e-13 truth."""
    wl = W.make_workload(cfg, n=n, box=box)
    # the truth by the same shape: the V-cycle diverges on the 8^3 star
    truth = run_oracle(wl, eps=1e-13, alt=1, cycle_type=cycle)
    assert truth["stats"].converged
    same = run_oracle(wl, eps=1e-10, alt=1, cycle_type=cycle)
    out = run_gpu(wl, use_alt_energy_fix=1, precision=prec, cycle_type=cycle)
    st = out["stats"]
    print(f"cycle {cycle} cfg{cfg} n={n} prec={prec} iters={st['iters']} "
          f"(oracle {same['stats'].iters}) res/res0={st['res'] / st['res0']:.2e}")
    assert st["converged"]
    singular = all(b[0] != W.OUTLET for b in wl["bc"])
    dim = wl["dim"]
    e_phi = rel_err(out["phi"], truth["phi"], remove_mean=singular)
    e_phi_ref = rel_err(same["phi"], truth["phi"], remove_mean=singular)
    e_u = max(rel_err(out["faces"][d], truth["faces"][d]) for d in range(dim))
    e_u_ref = max(rel_err(same["faces"][d], truth["faces"][d]) for d in range(dim))
    assert e_phi <= (max(1e-8, 2.0 * e_phi_ref) if prec == 2 else 1e-8)
    assert e_u <= max(1e-8, 2.0 * e_u_ref)
    if prec == 0:
        assert abs(st["iters"] - same["stats"].iters) <= 1
        assert rel_err(out["phi"], same["phi"], remove_mean=singular) <= 1e-9


def test_unknown_cycle_type_is_refused(gpu):
    from maestro_amd import mac
    wl = W.make_workload(1, n=(32, 32, 32), box=16)
    with pytest.raises(mac.MacError):
        run_gpu(wl, cycle_type=4)


BOTTOM_CASES = [(b, cfg, n, box, prec)
                for cfg, n, box, prec in [(1, (64, 64, 64), 32, 0), (0, (64, 64), 16, 0),
                                          (1, (96, 80, 48), 16, 0),   # 90-cell bottom: block form
                                          (3, (96, 80, 48), 16, 1), (2, (128, 64, 64), 32, 1)]
                for b in (0, 1)
                if not (b == 0 and cfg == 3)]   # smoothing only stalls on the star's contrast


@pytest.mark.parametrize("bottom,cfg,n,box,prec", BOTTOM_CASES)
def test_bottom_solvers_match_oracle(gpu, bottom, cfg, n, box, prec):
    """mg_bottom_solver 0 (smoothing only) and 1 (BiCGStab) against the
    oracle with the same bottom solver (FP64: the same iterates up to the
    bottom solve's summation order), and the answer against the truth."""
    wl = W.make_workload(cfg, n=n, box=box)
    truth = run_oracle(wl, eps=1e-13, alt=1)
    same = run_oracle(wl, eps=1e-10, alt=1, bottom_solver=bottom)
    out = run_gpu(wl, use_alt_energy_fix=1, precision=prec, bottom_solver=bottom)
    st = out["stats"]
    print(f"bottom {bottom} cfg{cfg} n={n} prec={prec} iters={st['iters']} (oracle {same['stats'].iters})")
    assert st["converged"]
    singular = all(b[0] != W.OUTLET for b in wl["bc"])
    dim = wl["dim"]
    assert rel_err(out["phi"], truth["phi"], remove_mean=singular) <= 1e-8
    e_u = max(rel_err(out["faces"][d], truth["faces"][d]) for d in range(dim))
    e_u_ref = max(rel_err(same["faces"][d], truth["faces"][d]) for d in range(dim))
    assert e_u <= max(1e-8, 2.0 * e_u_ref)
    if prec == 0:
        assert abs(st["iters"] - same["stats"].iters) <= 1
        assert rel_err(out["phi"], same["phi"], remove_mean=singular) <= 1e-9


def test_unknown_bottom_solver_is_refused(gpu):
    from maestro_amd import mac
    wl = W.make_workload(1, n=(32, 32, 32), box=16)
    with pytest.raises(mac.MacError):
        run_gpu(wl, bottom_solver=3)


@pytest.mark.parametrize("prec,inner", [(1, 1), (1, 4), (2, 1), (2, 3)])
def test_explicit_inner_cycles(gpu, prec, inner):
    """A fixed count of FP32 V-cycles per refinement step (the default picks
    3 or 2): the same answer to the tolerance, in the expected step count."""
    wl = W.make_workload(2, n=(128, 64, 64), box=32)
    truth = run_oracle(wl, eps=1e-13, alt=1)
    out = run_gpu(wl, use_alt_energy_fix=1, precision=prec, inner_cycles=inner)
    dflt = run_gpu(wl, use_alt_energy_fix=1, precision=prec)
    st, sd = out["stats"], dflt["stats"]
    print(f"prec {prec} inner {inner}: {st['iters']} steps, {st['inner_cycles']} cycles; "
          f"default {sd['iters']} steps, {sd['inner_cycles']} cycles")
    assert st["converged"] and st["inner_cycles"] == inner * st["iters"]
    assert sd["inner_cycles"] == (3 if prec == 1 else 2) * sd["iters"]
    assert rel_err(out["phi"], truth["phi"], remove_mean=True) <= 1e-8


@pytest.mark.parametrize("cfg,n,box", [(0, (128, 128), 32), (2, (96, 64, 64), 32)])
def test_default_inner_cycles_follow_storage(gpu, cfg, n, box):
    """MAC_PREC_MIXED_BF16 where the compact storage does not apply (2D; nx
    neither 128 nor 256) keeps FP32 storage, and so runs the FP32-storage
    default of 3 FP32 cycles per refinement step."""
    wl = W.make_workload(cfg, n=n, box=box)
    truth = run_oracle(wl, eps=1e-13, alt=1)
    out = run_gpu(wl, use_alt_energy_fix=1, precision=2)
    st = out["stats"]
    print(f"n={n}: {st['iters']} steps, {st['inner_cycles']} cycles")
    assert st["converged"] and st["inner_cycles"] == 3 * st["iters"]
    assert rel_err(out["phi"], truth["phi"], remove_mean=True) <= 1e-8


@pytest.mark.parametrize("cfg,n,box,cycle", [(1, (32, 32, 32), 16, 3), (2, (128, 64, 64), 32, 3),
                                             (0, (64, 64), 16, 3), (3, (32, 32, 32), 16, 3),
                                             (3, (64, 64, 64), 32, 1), (1, (64, 64, 32), 32, 2)])
def test_over_relaxed_fp64_matches_oracle(gpu, cfg, n, box, cycle):
    """FP64 with mac_opts.omega = 1.25 (1.2 in 2D) against the oracle's
    over-relaxed cycle of the same shape (V, F, W; the same levels
    over-relaxed: those above the single-workgroup coarse cycle): the same
    cycle count (+-1) and the same 1e-10 iterate up to the bottom solver's
    summation order; fewer cycles than the Gauss-Seidel cycle."""
    wl = W.make_workload(cfg, n=n, box=box)
    w = 1.25 if len(n) == 3 else 1.2
    same = run_oracle(wl, eps=1e-10, alt=1, omega=w, cycle_type=cycle)
    gs = run_oracle(wl, eps=1e-10, alt=1, cycle_type=cycle)
    out = run_gpu(wl, use_alt_energy_fix=1, precision=0, omega=w, cycle_type=cycle)
    st = out["stats"]
    print(f"cfg{cfg} n={n} cycle {cycle} omega={w}: {st['iters']} cycles (oracle {same['stats'].iters}, "
          f"Gauss-Seidel {gs['stats'].iters})")
    assert st["converged"]
    assert abs(st["iters"] - same["stats"].iters) <= 1
    assert same["stats"].iters < gs["stats"].iters
    singular = all(b[0] != W.OUTLET for b in wl["bc"])
    assert rel_err(out["phi"], same["phi"], remove_mean=singular) <= 1e-9
    dim = wl["dim"]
    assert max(rel_err(out["faces"][d], same["faces"][d]) for d in range(dim)) <= 1e-9


@pytest.mark.parametrize("cfg", [1, 2])
def test_mixed_over_relaxed_saves_cycles(gpu, cfg):
    """MAC_PREC_MIXED with mac_opts.omega = 1.25 (over-relaxed FP32 cycles)
    against the default Gauss-Seidel cycles: fewer FP32 cycles to the same
    1e-10 relative residual.  The stopping rule holds both to the residual;
    the over-relaxed iterate leaves more of its error in smooth modes, which
    the residual weighs least: at 64^3 the oracle's FP64 cycle at 1e-10 is
    1.2e-8 from the truth in phi with omega = 1.25 against 4.3e-9 with
    Gauss-Seidel, 7e-10 against 4.5e-10 in U."""
    wl = W.make_workload(cfg, n=(64, 64, 64), box=32)
    truth = run_oracle(wl, eps=1e-13, alt=1)
    same = run_oracle(wl, eps=1e-10, alt=1, omega=1.25)
    e_ref = rel_err(same["phi"], truth["phi"], remove_mean=True)
    e_u_ref = max(rel_err(same["faces"][d], truth["faces"][d]) for d in range(3))
    sor = run_gpu(wl, use_alt_energy_fix=1, precision=1, omega=1.25)
    gs = run_gpu(wl, use_alt_energy_fix=1, precision=1)
    print(f"cfg{cfg}: {sor['stats']['inner_cycles']} FP32 cycles over-relaxed, "
          f"{gs['stats']['inner_cycles']} Gauss-Seidel")
    assert sor["stats"]["converged"] and gs["stats"]["converged"]
    assert sor["stats"]["res"] <= 1e-10 * sor["stats"]["res0"]
    assert sor["stats"]["inner_cycles"] < gs["stats"]["inner_cycles"]
    e = rel_err(sor["phi"], truth["phi"], remove_mean=True)
    e_u = max(rel_err(sor["faces"][d], truth["faces"][d]) for d in range(3))
    print(f"  phi {e:.2e} U {e_u:.2e} (oracle's over-relaxed FP64 cycle: {e_ref:.2e} {e_u_ref:.2e})")
    assert e <= max(1e-8, 2.0 * e_ref) and e_u <= max(1e-8, 2.0 * e_u_ref)


def test_omega_out_of_range_is_refused(gpu):
    from maestro_amd import mac
    wl = W.make_workload(1, n=(32, 32, 32), box=16)
    for w in (-0.5, 2.0):
        with pytest.raises(mac.MacError):
            run_gpu(wl, omega=w)

