"""Kernel-level parity: each solver operator of the HIP library, driven
through the diagnostic C entry points, against the oracle on identical
inputs.  The stencils evaluate in the same order with explicit fma on both
sides, so these comparisons are bitwise unless stated otherwise."""
import numpy as np
import pytest

from maestro_amd import mac
from maestro_amd import workloads as W

from .helpers import bf16_round, ell, fold

pytestmark = pytest.mark.gpu

P, WALL, OUT = W.PERIODIC, W.SLIP_WALL, W.OUTLET

BCS_3D = [
    [[P, P], [P, P], [WALL, WALL]],
    [[OUT, OUT], [OUT, OUT], [OUT, OUT]],
    [[P, P], [WALL, OUT], [OUT, WALL]],
    [[WALL, WALL], [WALL, WALL], [WALL, WALL]],
]
SIZES_3D = [(32, 16, 16), (64, 48, 40), (128, 12, 8), (128, 12, 9)]


def make_plan(n, bc, dx=None, **opts):
    dim = len(n)
    dx = dx or [1.0 / n[0]] * dim
    lay = mac.Layout(dim=dim, boxes=[((0,) * dim, tuple(x - 1 for x in n))],
                     domain_lo=(0,) * dim, domain_hi=tuple(x - 1 for x in n), dx=tuple(dx),
                     phys_bc=bc)
    return mac.MacPlan(lay, mac.default_opts(**opts))


def rand_level(g, n, dim, rng, O):
    b = [rng.uniform(0.5, 2.0, size=O.face_shape(g, d)) for d in range(dim)]
    phi = rng.standard_normal(O.cell_shape(g))
    rhs = rng.standard_normal(O.cell_shape(g))
    return b, phi, rhs


def push_level(plan, level, bc, n, b, phi, rhs, dim):
    for k, v in fold(bc, n, b, dim).items():
        plan.debug_set(level, k, v)
    plan.debug_set(level, "phi", phi.ravel())
    plan.debug_set(level, "rhs", rhs.ravel())


# over-relaxation of the sweeps (mac_opts.omega): 0 = the default,
# Gauss-Seidel; 1.25 = over-relaxed red-black (the oracle's orc_gsrb_omega)
OMEGAS = [0.0, 1.25]


def ref_omega(n, w):
    """The over-relaxation a plan applies to level 0's sweeps: w where level
    0 lies above the single-workgroup coarse cycle (more than 512 cells in
    3D, 1024 in 2D, and a coarser level exists: every extent even and at
    least 4), else none (Gauss-Seidel, also for w = 0, the default)."""
    coarser = all(m % 2 == 0 and m // 2 >= 2 for m in n)
    above = int(np.prod(n)) > (512 if len(n) == 3 else 1024) and coarser
    return w if (w != 0.0 and above) else None


@pytest.mark.parametrize("omega", OMEGAS)
@pytest.mark.parametrize("n", SIZES_3D)
@pytest.mark.parametrize("bc", BCS_3D)
def test_gsrb_sweep_bitwise(gpu, oracle_lib, n, bc, omega):
    O = oracle_lib
    rng = np.random.default_rng(7)
    g = O.make_grid(n, [1.0 / n[0]] * 3, bc)
    b, phi, rhs = rand_level(g, n, 3, rng, O)
    plan = make_plan(n, bc, omega=omega)
    push_level(plan, 0, bc, n, b, phi, rhs, 3)
    plan.debug_op(0, "gsrb", 2)
    got = plan.debug_get(0, "phi").reshape(phi.shape)
    ref = O.gsrb(g, b, phi, rhs, sweeps=2, omega=ref_omega(n, omega))
    assert np.array_equal(got, ref), np.abs(got - ref).max()
    if ref_omega(n, omega) is not None:
        assert not np.array_equal(got, O.gsrb(g, b, phi, rhs, sweeps=2))


@pytest.mark.parametrize("omega", [0.0, 1.2])
@pytest.mark.parametrize("bc", [[[P, P], [WALL, WALL]], [[OUT, WALL], [P, P]]])
def test_gsrb_sweep_bitwise_2d(gpu, oracle_lib, bc, omega):
    O = oracle_lib
    n = (48, 32)
    rng = np.random.default_rng(3)
    g = O.make_grid(n, [1.0 / n[0]] * 2, bc)
    b, phi, rhs = rand_level(g, n, 2, rng, O)
    plan = make_plan(n, bc, omega=omega)
    push_level(plan, 0, bc, n, b, phi, rhs, 2)
    plan.debug_op(0, "gsrb", 1)
    got = plan.debug_get(0, "phi").reshape(phi.shape)
    ref = O.gsrb(g, b, phi, rhs, sweeps=1, omega=ref_omega(n, omega))
    assert np.array_equal(got, ref)


def fp32_reference(O, n, bc, b, phi, rhs, sweeps, bf16=False, omega=None):
    """The oracle's FP32 sweep on the level rounded to FP32 as the library
    rounds it (weights, rhs, phi and 1/h^2 each cast from FP64); bf16: the
    weights and rhs further rounded to bfloat16 (compact storage)."""
    dim = len(n)
    w = {k: np.asarray(v, dtype=np.float32) for k, v in fold(bc, n, b, dim).items()}
    r32 = rhs.astype(np.float32)
    if bf16:
        w = {k: bf16_round(v) for k, v in w.items()}
        r32 = bf16_round(r32).reshape(rhs.shape)
    ih2 = [np.float32(1.0 / (1.0 / n[0]) ** 2)] * dim
    per = [1 if bc[d][0] == P else 0 for d in range(dim)]
    out = O.gsrb_folded_f32(dim, n, per, ih2, w, phi.astype(np.float32), r32, sweeps=sweeps, omega=omega)
    return out.astype(np.float64)


@pytest.mark.parametrize("omega", OMEGAS)
@pytest.mark.parametrize("sweeps", [1, 2])
@pytest.mark.parametrize("n", [(128, 12, 8), (256, 20, 10), (64, 48, 40), (128, 30, 6), (128, 12, 9)])
@pytest.mark.parametrize("bc", BCS_3D)
def test_gsrb32_sweep_bitwise(gpu, oracle_lib, n, bc, sweeps, omega):
    """The FP32 smoother of the mixed path (full-row kernel where nx is a
    multiple of 128, the one-column kernel elsewhere) against the oracle's
    FP32 sweep, bit for bit, plain and over-relaxed."""
    O = oracle_lib
    rng = np.random.default_rng(21)
    g = O.make_grid(n, [1.0 / n[0]] * 3, bc)
    b, phi, rhs = rand_level(g, n, 3, rng, O)
    plan = make_plan(n, bc, omega=omega)
    push_level(plan, 0, bc, n, b, phi, rhs, 3)
    plan.debug_op(0, "gsrb32", sweeps)
    got = plan.debug_get(0, "phi").reshape(phi.shape)
    ref = fp32_reference(O, n, bc, b, phi, rhs, sweeps, omega=ref_omega(n, omega))
    assert np.array_equal(got, ref), np.abs(got - ref).max()


@pytest.mark.parametrize("omega", OMEGAS)
@pytest.mark.parametrize("sweeps", [1, 2])
@pytest.mark.parametrize("n", [(128, 12, 8), (256, 20, 10), (256, 14, 8), (256, 30, 9), (128, 30, 6),
                               (128, 12, 9), (64, 48, 40)])
@pytest.mark.parametrize("bc", BCS_3D)
def test_gsrb32_bf16_sweep_bitwise(gpu, oracle_lib, n, bc, sweeps, omega):
    """The compact-storage smoother (bfloat16 weights and rhs, FP32 iterate
    and arithmetic) against the oracle's FP32 sweep on the same rounded
    values, bit for bit.  Levels without the full-row kernel (nx other than
    128 or 256, or fewer than 8 planes) have no compact form: the plan keeps
    FP32 storage there, and the FP32 reference applies."""
    O = oracle_lib
    rng = np.random.default_rng(29)
    g = O.make_grid(n, [1.0 / n[0]] * 3, bc)
    b, phi, rhs = rand_level(g, n, 3, rng, O)
    plan = make_plan(n, bc, precision=mac.PREC_MIXED_BF16, omega=omega)
    push_level(plan, 0, bc, n, b, phi, rhs, 3)
    plan.debug_op(0, "gsrb32", sweeps)
    got = plan.debug_get(0, "phi").reshape(phi.shape)
    compact = n[0] in (128, 256) and n[2] >= 8
    w = ref_omega(n, omega)
    ref = fp32_reference(O, n, bc, b, phi, rhs, sweeps, bf16=compact, omega=w)
    assert np.array_equal(got, ref), np.abs(got - ref).max()
    if compact:
        # the rounding is visible: the FP32-storage sweep differs
        assert not np.array_equal(got, fp32_reference(O, n, bc, b, phi, rhs, sweeps, omega=w))


@pytest.mark.parametrize("omega", [0.0, 1.2])
@pytest.mark.parametrize("bc", [[[P, P], [WALL, WALL]], [[OUT, WALL], [P, P]]])
def test_gsrb32_sweep_bitwise_2d(gpu, oracle_lib, bc, omega):
    O = oracle_lib
    n = (48, 32)
    rng = np.random.default_rng(23)
    g = O.make_grid(n, [1.0 / n[0]] * 2, bc)
    b, phi, rhs = rand_level(g, n, 2, rng, O)
    plan = make_plan(n, bc, omega=omega)
    push_level(plan, 0, bc, n, b, phi, rhs, 2)
    plan.debug_op(0, "gsrb32", 1)
    got = plan.debug_get(0, "phi").reshape(phi.shape)
    ref = fp32_reference(O, n, bc, b, phi, rhs, 1, omega=ref_omega(n, omega))
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("n", [(32, 16, 16), (64, 48, 40), (128, 12, 8), (256, 20, 12), (128, 36, 10)])
@pytest.mark.parametrize("bc", BCS_3D)
def test_residual_restrict_bitwise(gpu, oracle_lib, n, bc):
    O = oracle_lib
    rng = np.random.default_rng(11)
    g = O.make_grid(n, [1.0 / n[0]] * 3, bc)
    b, phi, rhs = rand_level(g, n, 3, rng, O)
    plan = make_plan(n, bc)
    push_level(plan, 0, bc, n, b, phi, rhs, 3)
    plan.debug_set(1, "phi", np.ones(int(np.prod([x // 2 for x in n]))))
    plan.debug_op(0, "resid_restrict")
    got = plan.debug_get(1, "rhs")
    res, _ = O.residual(g, b, phi, rhs)
    cn = [x // 2 for x in n]
    fine = O.LevelView(g, b, phi, rhs, res=res)
    coarse = O.LevelView(g, [np.zeros(1)] * 3, np.zeros(cn[::-1]), np.zeros(cn[::-1]), n=cn)
    O.level_restrict(g, fine, coarse)
    assert np.array_equal(got, coarse.rhs.ravel())
    assert not plan.debug_get(1, "phi").any()


# nx = 128, 256: the z-marching row prolongation, one coarse plane per wave
# up to (256, 256, 128), where a wave marches four coarse planes
@pytest.mark.parametrize("n", [(32, 16, 16), (64, 48, 40), (128, 12, 8), (256, 20, 12), (128, 36, 10),
                               (256, 128, 128), (256, 256, 128)])
@pytest.mark.parametrize("bc", BCS_3D)
def test_prolong_then_sweep_bitwise(gpu, oracle_lib, n, bc):
    O = oracle_lib
    rng = np.random.default_rng(5)
    g = O.make_grid(n, [1.0 / n[0]] * 3, bc)
    b, phi, rhs = rand_level(g, n, 3, rng, O)
    cn = [x // 2 for x in n]
    e = rng.standard_normal(cn[::-1])
    plan = make_plan(n, bc)
    push_level(plan, 0, bc, n, b, phi, rhs, 3)
    plan.debug_set(1, "phi", e.ravel())
    plan.debug_op(0, "prolong_sweep")
    got = plan.debug_get(0, "phi").reshape(phi.shape)
    fine = O.LevelView(g, b, phi.copy(), rhs)
    coarse = O.LevelView(g, [np.zeros(1)] * 3, e, np.zeros_like(e), n=cn)
    O.level_prolong(g, coarse, fine)
    ref = O.gsrb(g, b, fine.phi, rhs, sweeps=1)
    assert np.array_equal(got, ref), np.abs(got - ref).max()


@pytest.mark.parametrize("bc", BCS_3D)
def test_coarsen_bitwise(gpu, oracle_lib, bc):
    O = oracle_lib
    n = (32, 16, 16)
    rng = np.random.default_rng(2)
    g = O.make_grid(n, [1.0 / n[0]] * 3, bc)
    b, phi, rhs = rand_level(g, n, 3, rng, O)
    plan = make_plan(n, bc)
    push_level(plan, 0, bc, n, b, phi, rhs, 3)
    plan.debug_op(0, "coarsen")
    cn = [x // 2 for x in n]
    gc = O.make_grid(cn, [2.0 / n[0]] * 3, bc)
    fine = O.LevelView(g, b, phi, rhs)
    cb = [np.zeros(O.face_shape(gc, d)) for d in range(3)]
    coarse = O.LevelView(gc, cb, np.zeros(cn[::-1]), np.zeros(cn[::-1]))
    O.level_coarsen(g, fine, coarse)
    want = fold(bc, cn, coarse.b, 3)
    for k, v in want.items():
        assert np.array_equal(plan.debug_get(1, k), v), k


@pytest.mark.parametrize("omega", OMEGAS)
@pytest.mark.parametrize("bc", [BCS_3D[1], BCS_3D[2]])
def test_vcycle_matches_oracle(gpu, oracle_lib, bc, omega):
    """One full V-cycle (non-singular BCs): equal up to the bottom solver's
    reduction order; with over-relaxed sweeps on the levels above the coarse
    cycle (64x32x32 .. 16x8x8), as the oracle's cycle applies them."""
    O = oracle_lib
    n = (64, 32, 32)
    rng = np.random.default_rng(9)
    g = O.make_grid(n, [1.0 / n[0]] * 3, bc)
    b, phi, rhs = rand_level(g, n, 3, rng, O)
    plan = make_plan(n, bc, bottom_eps=1e-15, bottom_max_iter=200, omega=omega)
    push_level(plan, 0, bc, n, b, phi, rhs, 3)
    nlev = 0
    while True:
        try:
            plan.level_dims(nlev + 1)
            plan.debug_op(nlev, "coarsen")
            nlev += 1
        except mac.MacError:
            break
    plan.debug_op(0, "vcycle")
    got = plan.debug_get(0, "phi").reshape(phi.shape)
    ref, _, st, _ = O.mg_solve(g, b, phi, rhs, O.default_opts(max_iter=1, eps_rel=1e-300,
                                                                   bottom_eps=1e-15, bottom_maxit=200,
                                                                   omega=omega or 1.0))
    assert st.iters == 1
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 1e-12, err


@pytest.mark.parametrize("cfg,n,box,alt", [(1, (32, 32, 32), 16, 1), (3, (32, 32, 32), 16, 1),
                                           (0, (64, 64), 16, 0)])
def test_setup_and_correction_match_oracle(gpu, oracle_lib, cfg, n, box, alt):
    """macproject with no V-cycle: the setup kernel's coefficients and rh and
    the face correction with the warm-start phi, against the oracle."""
    O = oracle_lib
    from .helpers import run_gpu
    wl = W.make_workload(cfg, n=n, box=box)
    rng = np.random.default_rng(4)
    wl["phi"] = rng.standard_normal(wl["phi"].shape)
    dim = wl["dim"]
    lay_opts = mac.default_opts(max_iter=0, use_alt_energy_fix=alt)
    from .helpers import layout_for
    plan = mac.MacPlan(layout_for(wl), lay_opts)
    out = run_gpu(wl, plan=plan, check=False)
    g = O.make_grid(wl["n"], wl["dx"], wl["bc"])
    b = O.mk_mac_coeffs(g, wl["rho"])
    faces = [f.copy() for f in wl["faces"]]
    if wl["beta0"] is not None:
        b = O.scale_faces_1d(g, b, wl["beta0"], wl["beta0_edge"], 1)
        if alt:
            b = O.scale_faces_1d(g, b, wl["beta0"], wl["beta0_edge"], 1)
        faces = O.scale_faces_1d(g, faces, wl["beta0"], wl["beta0_edge"], 1)
    else:
        cart = wl["cart"]
        b = [b[d] * cart[d] for d in range(dim)]
        if alt:
            b = [b[d] * cart[d] for d in range(dim)]
        faces = [faces[d] * cart[d] for d in range(dim)]
    nn = list(wl["n"]) + [1] * (3 - dim)
    for k, v in fold(wl["bc"], nn, b, dim).items():
        assert np.array_equal(plan.debug_get(0, k), v), k
    rh = O.divumac(g, *faces)
    singular = all(x[0] != OUT for x in wl["bc"])
    if singular:
        rh = rh - rh.mean()
    got_rh = plan.debug_get(0, "rhs")
    assert np.abs(got_rh - rh.ravel()).max() <= 1e-13 * np.abs(rh).max()
    proj = O.mkumac(g, faces, wl["phi"], b)
    if wl["beta0"] is not None:
        proj = O.scale_faces_1d(g, proj, wl["beta0"], wl["beta0_edge"], 0)
    else:
        proj = [proj[d] / wl["cart"][d] for d in range(dim)]
    for d in range(dim):
        assert np.array_equal(out["faces"][d], proj[d]), d
    plan.close()


@pytest.mark.parametrize("cfg,n,box", [(2, (256, 64, 64), 64), (3, (128, 64, 64), 32), (1, (128, 32, 32), 32)])
def test_compact_level0_and_level1_weights_bitwise(gpu, cfg, n, box):
    """After a MAC_PREC_MIXED_BF16 solve: the compact finest-level weights are
    the bfloat16 rounding of the FP32 rounding of the FP64 weights, and the
    FP32 level-1 weights are the FP32 averages of the FP32-rounded fine faces
    (the one-pass conversion against the same steps done here in numpy)."""
    from .helpers import run_gpu
    wl = W.make_workload(cfg, n=n, box=box)
    from .helpers import layout_for
    plan = mac.MacPlan(layout_for(wl), mac.default_opts(precision=mac.PREC_MIXED_BF16))
    out = run_gpu(wl, plan=plan)
    assert out["stats"]["converged"]
    nx, ny, nz = n
    shapes = {"bx": (nz, ny, nx), "by": (nz, ny, nx), "bz": (nz, ny, nx),
              "bxh": (nz, ny), "byh": (nz, nx), "bzh": (ny, nx)}
    f32 = {}
    for w, shp in shapes.items():
        a = plan.debug_get(0, w).reshape(shp).astype(np.float32)
        f32[w] = a
        got = plan.debug_get(0, w, copy="bf16").reshape(shp)
        assert np.array_equal(got, bf16_round(a).reshape(shp).astype(np.float64)), w
    q4 = np.float32(0.25)
    bx, by, bz = f32["bx"], f32["by"], f32["bz"]
    ref = {
        "bx": q4 * ((bx[0::2, 0::2, 0::2] + bx[0::2, 1::2, 0::2]) + (bx[1::2, 0::2, 0::2] + bx[1::2, 1::2, 0::2])),
        "by": q4 * ((by[0::2, 0::2, 0::2] + by[0::2, 0::2, 1::2]) + (by[1::2, 0::2, 0::2] + by[1::2, 0::2, 1::2])),
        "bz": q4 * ((bz[0::2, 0::2, 0::2] + bz[0::2, 0::2, 1::2]) + (bz[0::2, 1::2, 0::2] + bz[0::2, 1::2, 1::2])),
    }
    for w in ("bxh", "byh", "bzh"):
        h = f32[w]
        ref[w] = q4 * ((h[0::2, 0::2] + h[0::2, 1::2]) + (h[1::2, 0::2] + h[1::2, 1::2]))
    for w, r in ref.items():
        got = plan.debug_get(1, w, copy="f32").reshape(r.shape)
        assert np.array_equal(got, r.astype(np.float64)), w
    plan.close()


@pytest.mark.parametrize("omega", OMEGAS)
@pytest.mark.parametrize("sweeps", [1, 2])
@pytest.mark.parametrize("n,box", [((128, 64, 32), 32), ((256, 16, 24), 8)])
@pytest.mark.parametrize("zbc", [[WALL, WALL], [OUT, WALL], [P, P]])
def test_gsrb32_rho_sweep(gpu, oracle_lib, n, box, zbc, sweeps, omega):
    """The rho form of the FP32 finest-level sweep (face weights formed in the
    sweep from the FP32 density: 2 beta0^2 / (rho_L + rho_R) with the
    hardware reciprocal, 1 ulp) against the oracle's FP32 sweep on the same
    weights formed with IEEE division, and the domain's z faces from the
    folded FP64 weights: within 1e-5 relative (a 1-ulp weight moves a cell
    by ~1e-7 of its value).  The plan's rho and factors come from a
    mixed-precision plane-parallel mac_project on the same inputs."""
    O = oracle_lib
    bc = [[P, P], [P, P], zbc]
    rng = np.random.default_rng(31)
    wl = W.make_workload(1, n=n, box=box)
    wl["bc"] = bc
    wl["rho"] = W._fill_rho_ghosts(wl["rho"][1:-1, 1:-1, 1:-1], bc, 3)
    wl["faces"] = W._random_faces(rng, n, 3, bc)
    if zbc[0] == P:   # one face at z = 0 and z = L: one beta0 there
        wl["beta0_edge"] = wl["beta0_edge"].copy()
        wl["beta0_edge"][-1] = wl["beta0_edge"][0]
    from maestro_amd import boxes as BX
    boxes, umac, phi0, rho, _, _ = BX.split_workload(wl)
    lay = mac.Layout.uniform(n, box, wl["dx"], bc)
    plan = mac.MacPlan(lay, mac.default_opts(precision=mac.PREC_MIXED, max_iter=0, omega=omega))
    plan.project(umac, phi0, rho, None, np.ascontiguousarray(wl["beta0"]),
                 np.ascontiguousarray(wl["beta0_edge"]), None, check=False)
    phi = rng.standard_normal((n[2], n[1], n[0]))
    rhs = rng.standard_normal((n[2], n[1], n[0])) * 1e4
    plan.debug_set(0, "phi", phi.ravel())
    plan.debug_set(0, "rhs", rhs.ravel())
    plan.debug_op(0, "gsrb32_rho", sweeps)
    got = plan.debug_get(0, "phi").reshape(phi.shape)
    plan.close()
    # the weights the rho form stands for, in FP32 with IEEE division
    f32 = np.float32
    rg = wl["rho"]                                  # one ghost layer
    r32 = rg[1:-1, 1:-1, 1:-1].astype(f32)
    b0, b0e = wl["beta0"], wl["beta0_edge"]
    cxy = (2.0 * b0 * b0).astype(f32)[:, None, None]
    cz = (2.0 * b0e * b0e).astype(f32)[:, None, None]
    inv = lambda a, b_: f32(1.0) / (a + b_)
    wx = cxy * inv(np.roll(r32, 1, axis=2), r32)    # low x faces (periodic)
    wy = cxy * inv(np.roll(r32, 1, axis=1), r32)
    wz = np.empty_like(r32)
    wz[1:] = cz[1:-1] * inv(r32[:-1], r32[1:])
    # the domain's z faces: the FP64 coefficient folded by the BC, in FP32
    def zface(k_lo, k_hi, e, side):
        beta = 2.0 / (rg[k_lo, 1:-1, 1:-1] + rg[k_hi, 1:-1, 1:-1]) * b0e[e] * b0e[e]
        kind = ell(bc[2][side])
        return (0.0 * beta if kind == 2 else (2.0 * beta if kind == 1 else beta)).astype(f32)
    wz[0] = zface(0, 1, 0, 0)
    bzh = zface(-2, -1, n[2], 1)
    w = {"bx": wx.ravel(), "by": wy.ravel(), "bz": wz.ravel(), "bxh": wx[:, :, 0].ravel(),
         "byh": wy[:, 0, :].ravel(), "bzh": bzh.ravel()}
    ih2 = [f32(1.0 / wl["dx"][d] ** 2) for d in range(3)]
    per = [1, 1, 1 if zbc[0] == P else 0]
    ref = O.gsrb_folded_f32(3, n, per, ih2, w, phi.astype(f32), rhs.astype(f32), sweeps=sweeps,
                            omega=ref_omega(n, omega)).astype(np.float64)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"n={n} zbc={zbc} sweeps={sweeps}: rel max diff {err:.2e}")
    assert err <= 1e-5
