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
