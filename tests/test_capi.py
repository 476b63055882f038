"""CPU checks of the C ABI: libmaestro_mac.so loads, exports every function
include/maestro_mac.h declares, and validates layouts and options before any
device work (so these run without a GPU)."""
import ctypes as C
import os
import re

import pytest

from maestro_amd import mac

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "maestro_mac.h")


def header_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mac_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_function():
    L = mac.lib()
    names = header_functions()
    assert len(names) >= 12
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(mac.SYMBOLS)
    assert L.mac_version().startswith(b"maestro_mac")


def test_default_opts_match_reference_settings():
    o = mac.default_opts()
    assert o.eps_rel == 1e-10          # eps_mac, Source/mg_eps.f90:22
    assert o.eps_abs == -1.0           # Source/macproject.f90:209
    assert o.bottom_eps == 1e-3        # Source/mac_multigrid.f90:61
    assert o.nub == 10                 # mg_bottom_nu, Source/_parameters:392
    assert (o.nu1, o.nu2) == (2, 2)
    assert o.use_alt_energy_fix == 1   # Source/_parameters:552-554
    assert o.precision == mac.PREC_FP64


def _create(layout):
    c = layout.to_c()
    h = C.c_void_p()
    rc = mac.lib().mac_plan_create(C.byref(c), C.byref(mac.default_opts()), C.byref(h))
    return rc, h


P, W = mac.PERIODIC, mac.SLIP_WALL


@pytest.mark.parametrize("boxes,code", [
    ([((0, 0, 0), (7, 7, 7))], 2),                             # does not cover the domain
    ([((0, 0, 0), (15, 15, 15)), ((0, 0, 0), (7, 7, 7))], 2),   # overlap
    ([((0, 0, 0), (16, 15, 15))], 2),                          # outside the domain
])
def test_bad_layouts_rejected(boxes, code):
    lay = mac.Layout(dim=3, boxes=boxes, domain_lo=(0, 0, 0), domain_hi=(15, 15, 15),
                     dx=(1 / 16,) * 3, phys_bc=[[P, P], [P, P], [W, W]])
    rc, h = _create(lay)
    assert rc == code and not h.value


def test_bad_bc_and_ghosts_rejected():
    box = [((0, 0, 0), (15, 15, 15))]
    lay = mac.Layout(dim=3, boxes=box, domain_lo=(0, 0, 0), domain_hi=(15, 15, 15),
                     dx=(1 / 16,) * 3, phys_bc=[[P, W], [P, P], [W, W]])   # half-periodic
    assert _create(lay)[0] == 1      # a side periodic only on one end
    lay = mac.Layout(dim=3, boxes=box, domain_lo=(0, 0, 0), domain_hi=(15, 15, 15),
                     dx=(1 / 16,) * 3, phys_bc=[[P, P], [P, P], [W, W]], ng_rho=0)
    assert _create(lay)[0] == 2      # rho needs a ghost layer for the boundary faces
    odd = [((0, 0, 0), (14, 15, 15))]
    lay = mac.Layout(dim=3, boxes=odd, domain_lo=(0, 0, 0), domain_hi=(14, 15, 15),
                     dx=(1 / 15,) * 3, phys_bc=[[P, P], [P, P], [W, W]])
    assert _create(lay)[0] == 2      # odd periodic extent breaks red-black colouring


def test_dim_validation():
    lay = mac.Layout(dim=1, boxes=[((0,), (7,))], domain_lo=(0,), domain_hi=(7,), dx=(0.1,),
                     phys_bc=[[P, P]])
    assert _create(lay)[0] == 1


def test_python_mirror_raises_on_invalid_combination():
    """macproject with both beta0 forms is an error, as in the reference
    (Source/macproject.f90:99-101)."""
    lay = mac.Layout.uniform((16, 16, 16), 8, (1 / 16,) * 3, [[P, P], [P, P], [W, W]])
    with pytest.raises(mac.MacError) as e:
        mac.macproject(lay, [], [], [], beta0_1d=object(), beta0_cart_edge=object())
    assert e.value.code == 1


def test_pointer_table_holds_checked_pointers():
    """mac.PtrTable: the arrays' data pointers in order, one kind of array per
    table, float64 and contiguous only."""
    import numpy as np
    import torch
    xs = [np.zeros((3, 4)), np.ones(5)]
    t = mac.PtrTable(xs)
    assert len(t) == 2 and not t.device
    assert [t.arr[i] for i in range(2)] == [a.ctypes.data for a in xs]
    ts = [torch.zeros(4, dtype=torch.float64), torch.ones(2, 3, dtype=torch.float64)]
    tt = mac.PtrTable(ts)
    assert tt.device and [tt.arr[i] for i in range(2)] == [a.data_ptr() for a in ts]
    with pytest.raises(ValueError):
        mac.PtrTable([np.zeros(3), torch.zeros(3, dtype=torch.float64)])
    with pytest.raises(ValueError):
        mac.PtrTable([np.zeros(3, dtype=np.float32)])
    with pytest.raises(ValueError):
        mac.PtrTable([np.zeros((4, 4))[:, ::2]])
    with pytest.raises(ValueError):
        mac.PtrTable([torch.zeros(4, 4, dtype=torch.float64).t()])


def _ml_layouts(fine_boxes, dhi_c=(15, 15), bc_f=None, dx_f=None):
    P = mac.PERIODIC
    bc = [[P, P], [P, P]]
    c = mac.Layout(dim=2, boxes=mac.Layout.uniform((16, 16), 8, (1 / 16, 1 / 16), bc).boxes, domain_lo=(0, 0),
                   domain_hi=dhi_c, dx=(1 / 16, 1 / 16), phys_bc=bc)
    f = mac.Layout(dim=2, boxes=fine_boxes, domain_lo=(0, 0), domain_hi=(31, 31),
                   dx=dx_f or (1 / 32, 1 / 32), phys_bc=bc_f or bc)
    return c, f


@pytest.mark.parametrize("case,code", [
    ("region_not_whole_coarse_cells", 2),   # the refined rectangle starts at an odd fine index
    ("fine_domain_not_refined", 2),         # the fine domain is not the coarse one refined by 2
    ("bc_mismatch", 2),                     # the two levels' physical BCs differ
    ("dx_not_halved", 2),                   # the fine cell size is not half the coarse one
    ("three_levels", 1),                    # nlevs != 2
])
def test_two_level_layouts_rejected(case, code):
    """mac_ml_plan_create refuses layouts the composite solve does not take,
    before any device work (so on the CPU too)."""
    W = mac.SLIP_WALL
    boxes = [([8, 8], [15, 15]), ([16, 8], [23, 15])]
    kw = {}
    if case == "region_not_whole_coarse_cells":
        boxes = [([9, 8], [16, 15])]
    elif case == "fine_domain_not_refined":
        kw["dhi_c"] = (14, 15)
    elif case == "bc_mismatch":
        kw["bc_f"] = [[W, W], [mac.PERIODIC, mac.PERIODIC]]
    elif case == "dx_not_halved":
        kw["dx_f"] = (1 / 16, 1 / 16)
    c, f = _ml_layouts(boxes, **kw)
    arr = (mac.CLayout * 2)(c.to_c(), f.to_c())
    h = C.c_void_p()
    nlev = 3 if case == "three_levels" else 2
    rc = mac.lib().mac_ml_plan_create(nlev, arr, C.byref(mac.default_opts()), C.byref(h))
    assert rc == code
    assert not h.value
