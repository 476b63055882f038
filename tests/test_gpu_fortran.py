"""The Fortran drop-in (maestro_amd/fortran/macproject_hip.f90) called by a
Fortran program, as MAESTRO's time step calls macproject: same inputs as the
Python binding, so the results must be identical; checked against the oracle
too.  The driver projects twice with the cached plan (predictor/corrector)."""
import os
import subprocess

import numpy as np
import pytest

from maestro_amd import boxes as BX
from maestro_amd import workloads as W

from .helpers import rel_err, run_gpu, run_oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "maestro_amd", "fortran", "macproject_driver")


def write_input(path, wl, alt, prec, ncalls=1, ng=(1, 1, 1), cycle=3, bottom=-1, omega=0.0):
    boxes, umac, phi, rho, _, _ = BX.split_workload(wl, ng_u=ng[0], ng_phi=ng[1], ng_rho=ng[2])
    dim = wl["dim"]
    n = list(wl["n"]) + [1] * (3 - dim)
    has_b0 = wl["beta0"] is not None
    hdr = [dim, len(boxes), ng[0], ng[1], ng[2], alt, int(has_b0), prec, ncalls, cycle, bottom,
           int(round(1000 * omega))]
    dhi = [x - 1 for x in n[:dim]] + [0] * (3 - dim)
    bc = wl["bc"] + [[-1, -1]] * (3 - dim)
    pbc = [bc[d][0] for d in range(3)] + [bc[d][1] for d in range(3)]   # (3,2), d fastest
    with open(path, "wb") as f:
        np.array(hdr + dhi + pbc, dtype=np.int32).tofile(f)
        for lo, hi in boxes:
            l3 = list(lo) + [0] * (3 - dim)
            h3 = list(hi) + [0] * (3 - dim)
            np.array(l3 + h3, dtype=np.int32).tofile(f)
        np.array(list(wl["dx"]) + [1.0] * (3 - dim), dtype=np.float64).tofile(f)
        for b in range(len(boxes)):
            for d in range(dim):
                umac[b][d].astype(np.float64).tofile(f)
            phi[b].tofile(f)
            rho[b].tofile(f)
        if has_b0:
            np.asarray(wl["beta0"], dtype=np.float64).tofile(f)
            np.asarray(wl["beta0_edge"], dtype=np.float64).tofile(f)
    return boxes, umac, phi


def read_output(path, umac, phi):
    raw = np.fromfile(path, dtype=np.uint8)
    off = 0
    out_u, out_p = [], []
    for b in range(len(phi)):
        row = []
        for a in umac[b]:
            nbytes = a.size * 8
            row.append(raw[off:off + nbytes].view(np.float64).reshape(a.shape))
            off += nbytes
        out_u.append(row)
        nbytes = phi[b].size * 8
        out_p.append(raw[off:off + nbytes].view(np.float64).reshape(phi[b].shape))
        off += nbytes
    iters, conv = raw[off:off + 8].view(np.int32)
    res0, res = raw[off + 8:off + 24].view(np.float64)
    return out_u, out_p, int(iters), int(conv), float(res0), float(res)


@pytest.mark.parametrize("cfg,n,box,alt,prec,cycle,bottom,omega", [(1, (32, 32, 32), 16, 1, 0, 3, -1, 0.0),
                                                                   (0, (64, 64), 16, 0, 0, 3, -1, 0.0),
                                                                   (2, (32, 32, 32), 16, 1, 1, 3, -1, 0.0),
                                                                   # mg_cycle_type W, mg_bottom_solver BiCGStab
                                                                   (1, (32, 32, 32), 16, 1, 0, 2, 1, 0.0),
                                                                   # F-cycle, smoothing-only bottom, 2D
                                                                   (0, (64, 64), 16, 0, 1, 1, 0, 0.0),
                                                                   # mac_omega: over-relaxed sweeps
                                                                   (2, (64, 64, 64), 32, 1, 1, 3, -1, 1.25)])
def test_fortran_driver_matches_binding_and_oracle(gpu, tmp_path, cfg, n, box, alt, prec, cycle, bottom, omega):
    assert os.path.exists(DRIVER), "build the Fortran shim first (maestro_amd.build)"
    wl = W.make_workload(cfg, n=n, box=box)
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    boxes, umac, phi = write_input(fin, wl, alt, prec, cycle=cycle, bottom=bottom, omega=omega)
    r = subprocess.run([DRIVER, fin, fout], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    fu, fp, iters, conv, res0, res = read_output(fout, umac, phi)
    assert conv == 1 and res <= 1e-10 * res0
    dim = wl["dim"]
    nn = list(wl["n"]) + [1] * (3 - dim)
    faces = [BX.gather_faces([u[d] for u in fu], boxes, d, 1, nn, dim) for d in range(dim)]
    phig = BX.gather_cells(fp, boxes, 1, nn, dim)
    # identical to the same library called from Python (mg_bottom_solver -1: the library's CG)
    py = run_gpu(wl, use_alt_energy_fix=alt, precision=prec, cycle_type=cycle,
                 bottom_solver=2 if bottom == -1 else bottom, omega=omega)
    for d in range(dim):
        assert np.array_equal(faces[d], py["faces"][d])
    assert np.array_equal(phig, py["phi"])
    # ghost faces too (fill_boundary + physical edge BCs)
    for b in range(len(boxes)):
        for d in range(dim):
            assert np.array_equal(fu[b][d], py["umac_boxes"][b][d])
    truth = run_oracle(wl, eps=1e-13, alt=alt)
    if omega == 0.0:
        assert rel_err(phig, truth["phi"], remove_mean=True) <= 1e-8
    else:
        # over-relaxed cycles: held to the oracle's over-relaxed 1e-10 solve's own error
        same = run_oracle(wl, eps=1e-10, alt=alt, omega=omega)
        e_ref = rel_err(same["phi"], truth["phi"], remove_mean=True)
        assert rel_err(phig, truth["phi"], remove_mean=True) <= max(1e-8, 2.0 * e_ref)


def test_fortran_two_calls_reuse_plan(gpu, tmp_path):
    """Predictor/corrector: the second call projects an already projected
    field with the warm-started phi (Source/advance.f90:441-442,937)."""
    wl = W.make_workload(1, n=(32, 32, 32), box=16)
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    write_input(fin, wl, 1, 0, ncalls=2)
    r = subprocess.run([DRIVER, fin, fout], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0
    lines = [l for l in r.stdout.splitlines() if l.startswith("call")]
    assert len(lines) == 2


SOLVE_DRIVER = os.path.join(ROOT, "maestro_amd", "fortran", "mac_multigrid_driver")


@pytest.mark.parametrize("n,box,pbc,bc_comp,prec", [
    # thermal diffusion: bc_comp = dm + rhoh_comp (a scalar; INLET Dirichlet)
    ((32, 32, 32), 16, [[W.PERIODIC] * 2, [W.SLIP_WALL] * 2, [W.INLET, W.OUTLET]], 5, 0),
    ((64, 32), 16, [[W.INLET, W.SLIP_WALL], [W.PERIODIC] * 2], 4, 1),
    # press_comp (OUTLET Dirichlet) and foextrap_comp (Neumann everywhere)
    ((32, 32, 32), 16, [[W.PERIODIC] * 2, [W.PERIODIC] * 2, [W.SLIP_WALL, W.OUTLET]], 100, 0),
    ((32, 32, 32), 16, [[W.PERIODIC] * 2, [W.PERIODIC] * 2, [W.INLET, W.OUTLET]], 101, 1),
])
def test_fortran_mac_multigrid_matches_binding_and_oracle(gpu, oracle_lib, tmp_path, n, box, pbc, bc_comp,
                                                          prec):
    """The Fortran mac_multigrid (alpha = rho-like, beta on faces, Dirichlet
    values in phi's ghost cells, reference ghost widths: alpha and beta 1)
    equals the Python binding of mac_solve bit for bit and the oracle's
    solve to its tolerance."""
    from maestro_amd import mac
    O = oracle_lib
    assert os.path.exists(SOLVE_DRIVER), "build the Fortran shim first (maestro_amd.build)"
    dim = len(n)
    comp = {100: mac.COMP_PRESS, 101: mac.COMP_EXTRAP}.get(bc_comp, mac.COMP_SCALAR)
    ell = [[mac.ell_bc_of(pbc[d][s], comp) for s in range(2)] for d in range(dim)]
    rng = np.random.default_rng(bc_comp + dim)
    h = 1.0 / n[0]
    g = O.ell_grid(n, [h] * dim, ell)
    b = [rng.uniform(0.5, 2.0, O.face_shape(g, d)) for d in range(dim)]
    for d in range(dim):
        if ell[d][0] == mac.ELL_PER:
            ax = 2 - d
            lo, hi = [slice(None)] * 3, [slice(None)] * 3
            lo[ax], hi[ax] = 0, -1
            b[d][tuple(hi)] = b[d][tuple(lo)]
    alpha = 0.05 / h**2 * rng.uniform(0.5, 2.0, O.cell_shape(g))
    rh = rng.standard_normal(O.cell_shape(g))
    phi_g = rng.uniform(-1.0, 1.0, O.ghost_shape(g))
    boxes = W.box_list(n, box)
    ng_rh, ng_phi, ng_a, ng_b = 0, 1, 1, 1
    rhb = [BX.split_cells(rh, lo_, hi_, ng_rh, dim, pbc) for lo_, hi_ in boxes]
    phb = [BX.split_cells(phi_g, lo_, hi_, ng_phi, dim, pbc, glob_ghost=1) for lo_, hi_ in boxes]
    ab = [BX.split_cells(alpha, lo_, hi_, ng_a, dim, pbc) for lo_, hi_ in boxes]
    bb = [[BX.split_faces(b[d], d, lo_, hi_, ng_b, dim, pbc) for d in range(dim)] for lo_, hi_ in boxes]
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    n3 = list(n) + [1] * (3 - dim)
    bc3 = pbc + [[W.PERIODIC] * 2] * (3 - dim)
    with open(fin, "wb") as f:
        np.array([dim, len(boxes), ng_rh, ng_phi, ng_a, ng_b, prec, bc_comp] + [x - 1 for x in n3[:dim]] +
                 [0] * (3 - dim) + [bc3[d][0] for d in range(3)] + [bc3[d][1] for d in range(3)],
                 dtype=np.int32).tofile(f)
        for lo_, hi_ in boxes:
            np.array(list(lo_) + [0] * (3 - dim) + list(hi_) + [0] * (3 - dim), dtype=np.int32).tofile(f)
        np.array([h] * dim + [1.0] * (3 - dim) + [1e-10, -1.0], dtype=np.float64).tofile(f)
        for q in range(len(boxes)):
            rhb[q].tofile(f)
            phb[q].tofile(f)
            ab[q].tofile(f)
            for d in range(dim):
                bb[q][d].tofile(f)
    r = subprocess.run([SOLVE_DRIVER, fin, fout], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    raw = np.fromfile(fout, dtype=np.uint8)
    off, fp = 0, []
    for a in phb:
        fp.append(raw[off:off + a.nbytes].view(np.float64).reshape(a.shape))
        off += a.nbytes
    iters, conv = raw[off:off + 8].view(np.int32)
    res0, res = raw[off + 8:off + 24].view(np.float64)
    assert conv == 1 and res <= 1e-10 * res0
    # the Python binding of the same entry point on the same inputs
    lay = mac.Layout.uniform(n, box, [h] * dim, pbc, ng_phi=ng_phi)
    plan = mac.MacPlan(lay, mac.default_opts(precision=prec))
    php = [a.copy() for a in phb]
    st = plan.solve(rhb, php, bb, ab, comp=comp, ng_phi=ng_phi, ng_alpha=ng_a, ng_beta=ng_b)
    plan.close()
    assert st["iters"] == iters
    for a, c in zip(fp, php):
        assert np.array_equal(a, c)
    ref, st_o, rc = O.solve(g, alpha, b, rh, phi_g, O.default_opts(eps_rel=1e-10))
    assert rc == 0
    truth, _, rc = O.solve(g, alpha, b, rh, phi_g, O.default_opts(eps_rel=1e-14))
    assert rc == 0
    got = BX.gather_cells(fp, boxes, ng_phi, n3, dim)
    inner = (slice(1, -1),) * dim
    same, truth = ref[inner].reshape(got.shape), truth[inner].reshape(got.shape)
    # a 1e-10 residual from a rough start leaves ~6e-8 in phi here (the oracle's
    # own 1e-10 solve included): the mixed path is held to twice that error,
    # the FP64 path (the oracle's iterates) to the oracle's answer
    e_ref = rel_err(same, truth)
    assert rel_err(got, truth) <= max(1e-8, 2.0 * e_ref)
    if prec == 0:
        assert rel_err(got, same) <= 1e-8


ML_DRIVER = os.path.join(ROOT, "maestro_amd", "fortran", "macproject_ml_driver")


@pytest.mark.parametrize("case", [3, 5, "L"])
def test_fortran_two_level_matches_binding(oracle_lib, tmp_path, case):
    """macproject with nlevs = 2 from Fortran (the level-2 boxes in the
    ml_layout, beta0_1d(2, 0:) per level): the same library call as the
    Python binding's host-array path, so the same bits."""
    from oracle import ml_oracle as M
    from .test_ml_oracle import REGIONS, UNIONS, random_two_level
    from .test_gpu_ml import ml_box_arrays, run_ml_gpu
    O = oracle_lib
    boxes = None
    if case == "L":
        # an L of two fine boxes (level 2 as a union of boxes)
        dim, n, bc, boxes = UNIONS[0]
        lo = [min(b[0][d] for b in boxes) for d in range(dim)]
        hi = [max(b[1][d] for b in boxes) for d in range(dim)]
    else:
        dim, n, bc, lo, hi = REGIONS[case]
    use_b0 = bc[dim - 1][0] != W.PERIODIC
    tl, fc, ff, rc, rf, kw = random_two_level(M, O, dim, n, bc, lo, hi, 41, beta0=use_b0, boxes=boxes)
    box = 4 if dim == 3 else 8
    cb, fb, rel, lv = ml_box_arrays(tl, fc, ff, rc, rf, box, box)
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    b3 = [list(bc[d]) for d in range(dim)] + [[-1, -1]] * (3 - dim)
    hdr = [dim, len(cb), len(fb), 1, 1, 1, 1, int(use_b0), 0, 1]
    dhi = [n - 1] * dim + [0] * (3 - dim)
    pbc = [b3[d][0] for d in range(3)] + [b3[d][1] for d in range(3)]
    with open(fin, "wb") as f:
        np.array(hdr + dhi + pbc, dtype=np.int32).tofile(f)
        for boxes in (cb, fb):
            for l, h in boxes:
                np.array(list(l) + [0] * (3 - dim) + list(h) + [0] * (3 - dim), dtype=np.int32).tofile(f)
        np.array([1.0 / n] * dim + [1.0] * (3 - dim), dtype=np.float64).tofile(f)
        for u, ph, r, _ in lv:
            for b in range(len(ph)):
                for d in range(dim):
                    u[b][d].tofile(f)
                ph[b].tofile(f)
                r[b].tofile(f)
        if use_b0:
            for key in ("beta0", "beta0_edge", "beta0_f", "beta0_edge_f"):
                np.asarray(kw[key], dtype=np.float64).tofile(f)
    r = subprocess.run([ML_DRIVER, fin, fout], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(fout, dtype=np.uint8)
    off = 0
    outs = []
    for u, ph, _, _ in lv:
        lev = []
        for b in range(len(ph)):
            for a in list(u[b]) + [ph[b]]:
                nbytes = a.size * 8
                lev.append(raw[off:off + nbytes].view(np.float64).reshape(a.shape))
                off += nbytes
        outs.append(lev)
    iters, conv = raw[off:off + 8].view(np.int32)
    assert conv == 1
    # the binding's host path on the same arrays
    g = run_ml_gpu(tl, fc, ff, rc, rf, kw if use_b0 else {}, box_c=box, box_f=box, eps=1e-9, host=True)
    for l in range(2):
        ref = []
        nbx = len(lv[l][1])
        for b in range(nbx):
            ref += g["ghosts"][l][b * dim:(b + 1) * dim]
        got_u = [x for b in range(nbx) for x in outs[l][b * (dim + 1):b * (dim + 1) + dim]]
        for x, y in zip(got_u, ref):
            assert np.array_equal(x, y)
    dc, df = M.composite_divergence(tl, g["faces_c"], g["faces_f"], **(kw if use_b0 else {}))
    assert max(np.abs(dc).max(), np.abs(df).max()) < 1e-6
