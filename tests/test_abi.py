"""CPU checks of the C ABI: the library loads, exports every function that
include/xrd_tomo.h declares, the descriptor layouts match, and argument
validation rejects bad calls before any device work (the reference gateway
performs no checks at all, Forward_XRD.cu:76-119)."""
import ctypes
import math
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "xrd_tomo.h")


@pytest.fixture(scope="module")
def L():
    from xrd_tomography_amd import _lib
    return _lib.lib()


def _header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(xrd_[a-z0-9_]+)\s*\(", text)))


def test_exports_every_declared_symbol(L):
    from xrd_tomography_amd import _lib
    declared = _header_functions()
    assert declared, "no declarations parsed"
    for name in declared:
        assert hasattr(L, name), f"{name} declared in xrd_tomo.h but not exported"
    assert sorted(_lib.EXPORTS) == declared


def test_struct_layout_matches_header():
    """Field order/size of the ctypes mirrors vs the C structs (parsed from the header)."""
    from xrd_tomography_amd.descriptors import Discretization, Geometry
    text = open(HEADER).read()
    for cls, tag in ((Geometry, "xrd_geometry"), (Discretization, "xrd_discretization")):
        body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (tag, tag), text, re.S).group(1)
        body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
        names = []
        for decl in body.split(";"):
            decl = decl.strip()
            if not decl:
                continue
            parts = decl.split(None, 1)
            names += [n.strip() for n in parts[1].split(",")]
        assert [f for f, _ in cls._fields_] == names


def test_reference_presets(L):
    from xrd_tomography_amd.descriptors import Discretization, Geometry
    g = Geometry()
    assert L.xrd_geometry_reference(100, 100, 32, 1024, 1, ctypes.byref(g)) == 0
    ref = Geometry.reference(100, 100, 32, 1024, 1)
    for f, _ in Geometry._fields_:
        assert getattr(g, f) == pytest.approx(getattr(ref, f), rel=1e-15, abs=0), f
    # integrate_diffraction.cu:81-128 at NX = 100: mm_over_vox = 1
    assert (g.r_src, g.r_det, g.center_x, g.col_pitch, g.det_z0) == (150, 170, 100, 0.5, 10)
    d = Discretization()
    assert L.xrd_discretization_reference(64, 256, 768, 3, ctypes.byref(d)) == 0
    r = Discretization.reference(64, 256, 768, 3)
    for f, _ in Discretization._fields_:
        assert getattr(d, f) == pytest.approx(getattr(r, f), rel=1e-15, abs=0), f
    assert d.e_width == 72.0 / 64 and d.energy_halfwidth == 5 and d.ray_step == 0.25
    assert L.xrd_num_rows(ctypes.byref(g), ctypes.byref(d)) == 32 * 1024 * 64
    assert L.xrd_num_cols(ctypes.byref(d)) == 256 * 3


def test_error_paths_without_device(L):
    from xrd_tomography_amd import _lib
    from xrd_tomography_amd.descriptors import Discretization, Geometry
    g = Geometry.reference(16, 16, 4, 8)
    d = Discretization.reference(16, 8, 24, 2)
    assert L.xrd_forward(None, ctypes.byref(g), ctypes.byref(d), *([None] * 6), 0.0, 1.0, 0.0,
                         1.0, None, None, None, None) == -2
    assert b"context" in L.xrd_last_error_string()
    assert L.xrd_context_create(0, None) == -2
    assert L.xrd_geometry_reference(0, 16, 4, 8, 1, ctypes.byref(g)) == -3
    assert L.xrd_discretization_reference(16, 8, 24, 2, None) == -2
    assert L.xrd_project(None, None, 0, 10, 8, None, None, None) == -2
    assert L.xrd_mlem_plan_create(None, None, 0, 10, 8, None, None, None, None, 0, None,
                                  ctypes.byref(ctypes.c_void_p())) == -2
    assert L.xrd_mlem_plan_iterate(None, 1, 0, None) == -2
    assert L.xrd_mlem_iterate_static(None, None, 0, 10, 8, None, None, None, None, 1, None) == -2
    assert L.xrd_forward_set_launch(None, 16, 0) == -2
    assert L.xrd_forward_bf16(None, ctypes.byref(g), ctypes.byref(d), None, None, None, 1, 0,
                              None, 0.0, 1.0, 0.0, 1.0, None, None, None, None) == -2
    assert L.xrd_trace_rays(None, 4, 4, 0.25, 1, None, None, None, None, 0, None, 0, None,
                            None, None) == -2
    assert L.xrd_version() >= 100
    with pytest.raises(_lib.XrdError) as ei:
        _lib.check(L.xrd_context_create(0, None), "xrd_context_create")
    assert ei.value.code == -2


def test_num_rows_matches_layout():
    from xrd_tomography_amd.synthetic import make_problem
    from xrd_tomography_amd.descriptors import num_rows
    p = make_problem("fan256")
    assert num_rows(p.geometry, p.disc) == 360 * 256 * 61 == p.n_rows
    assert p.n_cols == 1024


# ---------------------------------------------------------------------------
# host-side argument checks of the Python mirror: every bad tensor is refused with
# TypeError/ValueError before a context is created or a pointer reaches the library
# (on this CPU-only host a context cannot even be created, so reaching that point
# would raise RuntimeError instead)

def _small():
    import torch
    from xrd_tomography_amd.synthetic import make_problem
    p = make_problem("small")
    t = lambda name, dt: torch.from_numpy(p.flat(name)).to(dt)  # noqa: E731
    args = [t("spectrum", torch.float32), t("ifine", torch.float32),
            t("segmentation", torch.int32), t("a1", torch.float32), t("a2", torch.float32),
            t("s2_table", torch.float32)]
    return p, args


@pytest.mark.parametrize("which,bad", [
    (0, "short"), (1, "short"), (2, "dtype"), (3, "short"), (4, "dtype"), (5, "short"),
    (0, "host"), (2, "not_tensor")])
def test_forward_rejects_bad_inputs_before_launch(which, bad):
    import torch
    from xrd_tomography_amd import api
    p, args = _small()
    if bad == "short":
        args[which] = args[which][:-1]
    elif bad == "dtype":
        args[which] = args[which].to(torch.float64)
    elif bad == "not_tensor":
        args[which] = args[which].numpy()
    with pytest.raises((TypeError, ValueError)) as ei:
        api.Forward_XRD(*args, *p.lut, geometry=p.geometry, disc=p.disc)
    if bad == "host":  # right sizes and types: the device check refuses host memory
        assert "GPU tensors only" in str(ei.value)


def test_forward_rejects_bad_outputs_and_slices():
    import torch
    from xrd_tomography_amd import api
    p, args = _small()
    rows = p.n_rows
    good = (torch.empty(rows), torch.empty(rows), torch.empty(rows * p.n_cols))
    for out in ((torch.empty(rows - 1), good[1], good[2]),
                (good[0], torch.empty(rows, dtype=torch.float64), good[2]),
                (good[0], good[1], torch.empty(rows * p.n_cols - 4))):
        with pytest.raises((TypeError, ValueError)):
            api.Forward_XRD(*args, *p.lut, geometry=p.geometry, disc=p.disc, out=out)
    for n_slices, sl in ((2, 2), (2, -1), (3, 0)):  # out of range; volume sizes short
        with pytest.raises(ValueError):
            api.forward_slice(*args, *p.lut, n_slices, sl, geometry=p.geometry, disc=p.disc)


def test_matrix_operators_reject_bad_inputs():
    import torch
    from xrd_tomography_amd import api
    A = torch.zeros(10, 8)
    for call in (lambda: api.project(A, torch.zeros(7)),
                 lambda: api.project(A, torch.zeros(8, dtype=torch.float64)),
                 lambda: api.backproject(A, torch.zeros(9)),
                 lambda: api.project(A.double(), torch.zeros(8)),
                 lambda: api.project(A[:, 0], torch.zeros(8)),
                 lambda: api.convert_matrix(A.double(), torch.float16),
                 lambda: api.convert_matrix_scaled(A, torch.float32),
                 lambda: api.MlemPlan(A, torch.zeros(10), torch.zeros(9), torch.zeros(8)),
                 lambda: api.MlemPlan(A, torch.zeros(10), torch.zeros(10), torch.zeros(8),
                                      col_scale=torch.zeros(7)),
                 lambda: api.mlem_iterate_static(A, torch.zeros(10), torch.zeros(10),
                                                 torch.zeros(8, dtype=torch.float16), 1),
                 lambda: api.MlemBatch([(A, torch.zeros(10), torch.zeros(10), torch.zeros(9))]),
                 lambda: api.project(A, torch.zeros(8))):  # right sizes, host memory
        with pytest.raises((TypeError, ValueError)):
            call()


def test_coefficient_and_ct_wrappers_reject_bad_inputs():
    import torch
    from xrd_tomography_amd import api
    from xrd_tomography_amd.synthetic import make_coeff_problem, make_problem
    cp = make_coeff_problem(make_problem("small"), K=4, full_mu=False)
    g, d = cp.base.geometry, cp.base.disc
    f = lambda a: torch.from_numpy(a.copy())  # noqa: E731
    sp = f(cp.base.flat("spectrum"))
    seg = f(cp.base.flat("segmentation"))
    good = [sp, seg, f(cp.coef), f(cp.basis), f(cp.maps), f(cp.beta)]
    bads = {2: f(cp.coef[:-1]), 3: f(cp.basis[:, :-1]), 4: f(cp.maps[:, :-1]),
            5: f(cp.beta[:1]), 1: seg.to(torch.int64)}
    for i, b in bads.items():
        args = list(good)
        args[i] = b
        with pytest.raises((TypeError, ValueError)):
            api.forward_coeffs(*args, geometry=g, disc=d)
    with pytest.raises(ValueError):
        api.vjp_coeffs(*good, torch.zeros(cp.base.n_rows - 1), geometry=g, disc=d)
    with pytest.raises(ValueError):
        api.forward_ct(sp, f(cp.maps), f(cp.beta[:, :-1]), geometry=g, disc=d)
    with pytest.raises(ValueError):
        api.vjp_ct(sp, f(cp.maps), f(cp.beta), torch.zeros(3), geometry=g, disc=d)
    with pytest.raises(ValueError):
        api.poisson_nll(torch.zeros(5), torch.zeros(4))
    with pytest.raises(ValueError):
        api.trace_rays(torch.zeros(3, 3, dtype=torch.float64), 4, 4)
