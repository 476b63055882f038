"""mac_project_noalloc (the projection that allocates nothing; its call
buffers live in a fixed per-GPU array of the library) against mac_project:
the same outputs bit for bit and the same stats, on two calls in a row on
different plans, the second with a smaller input; the refusals (host
pointers, solver storage not yet allocated, a layout whose buffers exceed
the fixed array)."""
import numpy as np
import pytest

from maestro_amd import boxes as BX
from maestro_amd import mac
from maestro_amd import workloads as W

pytestmark = pytest.mark.gpu


def _device_inputs(torch, wl):
    boxes, umac, phi, rho, _, _ = BX.split_workload(wl)
    cv = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    return ([[cv(a) for a in b] for b in umac], [cv(a) for a in phi], [cv(a) for a in rho],
            cv(wl["beta0"]), cv(wl["beta0_edge"]))


def _run(torch, plan, wl, noalloc):
    du, dp, dr, b0, b0e = _device_inputs(torch, wl)
    st = plan.project(du, dp, dr, None, b0, b0e, None, noalloc=noalloc)
    torch.cuda.synchronize()
    return st, [[a.cpu().numpy() for a in b] for b in du], [a.cpu().numpy() for a in dp]


@pytest.mark.parametrize("prec", [mac.PREC_FP64, mac.PREC_MIXED, mac.PREC_MIXED_BF16])
def test_noalloc_matches_project(gpu, prec):
    torch = gpu
    # the second call's input is smaller; nx = 128 puts MIXED_BF16's finest
    # level in compact storage
    cases = [((128, 64, 64), 32), ((64, 64, 64), 32)]
    plans = []
    for n, box in cases:
        wl = W.make_workload(1, n=n, box=box)
        lay = mac.Layout.uniform(n, box, wl["dx"], wl["bc"])
        plan = mac.MacPlan(lay, mac.default_opts(device_ptrs=1, precision=prec))
        if prec == mac.PREC_FP64:   # its storage exists from plan creation on
            st0, _, _ = _run(torch, plan, wl, True)
            assert st0["converged"]
        else:                       # refused until a call in this precision has allocated
            du, dp, dr, b0, b0e = _device_inputs(torch, wl)
            with pytest.raises(mac.MacError) as ei:
                plan.project(du, dp, dr, None, b0, b0e, None, noalloc=True)
            assert ei.value.code == 1
        ref = _run(torch, plan, wl, False)
        plans.append((plan, wl, ref))
    # two calls in a row through the fixed array, the second smaller
    for plan, wl, (st_r, u_r, p_r) in plans:
        st, u, p = _run(torch, plan, wl, True)
        for key in ("iters", "inner_cycles", "res0", "res", "converged", "rhs_norm"):
            assert st[key] == st_r[key], (key, st[key], st_r[key])
        for b in range(len(u)):
            for d in range(3):
                assert np.array_equal(u[b][d], u_r[b][d])
            assert np.array_equal(p[b], p_r[b])
        plan.close()


def test_noalloc_refuses_host_pointers(gpu):
    n, box = (32, 32, 32), 16
    wl = W.make_workload(1, n=n, box=box)
    lay = mac.Layout.uniform(n, box, wl["dx"], wl["bc"])
    plan = mac.MacPlan(lay, mac.default_opts(device_ptrs=0))
    boxes, umac, phi, rho, _, _ = BX.split_workload(wl)
    with pytest.raises(mac.MacError) as ei:
        plan.project(umac, phi, rho, None, np.ascontiguousarray(wl["beta0"]),
                     np.ascontiguousarray(wl["beta0_edge"]), None, noalloc=True)
    plan.close()
    assert ei.value.code == 1


def test_noalloc_refuses_layout_beyond_its_scratch(gpu):
    """32768 boxes of 2^3 (a 64^3 domain): the pointer table alone is 2.4 MB,
    more than the 2 MiB array; the call returns MAC_ERR_TOO_LARGE before
    touching any array (the same placeholder tensor stands in for every box)."""
    torch = gpu
    n, box = (64, 64, 64), 2
    wl = W.make_workload(1, n=(8, 8, 8), box=8)   # only dx, bc and the beta0 profile's shape matter
    lay = mac.Layout.uniform(n, box, [1.0 / 64] * 3, wl["bc"])
    plan = mac.MacPlan(lay, mac.default_opts(device_ptrs=1))
    nb = len(lay.boxes)
    assert nb == 32768
    t = torch.zeros(8, dtype=torch.float64, device="cuda")
    b0 = torch.ones(64, dtype=torch.float64, device="cuda")
    b0e = torch.ones(65, dtype=torch.float64, device="cuda")
    with pytest.raises(mac.MacError) as ei:
        plan.project([[t, t, t]] * nb, [t] * nb, [t] * nb, None, b0, b0e, None, noalloc=True)
    plan.close()
    assert ei.value.code == 7
