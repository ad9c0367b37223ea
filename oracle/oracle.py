"""ctypes wrapper of the FP64 oracle (oracle/libxrd_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker or the timed CPU baseline.  The
product path (xrd_tomography_amd) never imports it.
"""
from __future__ import annotations

import contextlib
import ctypes
import os
import subprocess

import numpy as np

from xrd_tomography_amd.descriptors import Discretization, Geometry

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libxrd_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, FP64, no contraction)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE, "libxrd_oracle.so"],
                       check=True, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = _configure(ctypes.CDLL(_LIB_PATH))
    return _lib


@contextlib.contextmanager
def using_library(path: str):
    """Route every call of this module to another build of the oracle (for example
    one compiled with a planted mistake, to show that a pin detects it)."""
    global _lib
    saved = _lib
    _lib = _configure(ctypes.CDLL(os.path.abspath(path)))
    try:
        yield _lib
    finally:
        _lib = saved


def _configure(L):
    """Declare the argument and result types of the oracle's exports."""
    L.xo_get_intersection.restype = ctypes.c_int
    L.xo_get_intersection.argtypes = [ctypes.c_double] * 4 + [ctypes.c_int] * 2 + [_dp, _dp]
    L.xo_trace_ray.restype = ctypes.c_int
    L.xo_trace_ray.argtypes = ([ctypes.c_double] * 4 + [ctypes.c_int] * 2 +
                               [ctypes.c_double, _dp, _dp, _ip, ctypes.c_int32, _ip])
    L.xo_line_integrals.restype = None
    L.xo_line_integrals.argtypes = ([_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_double] + [ctypes.c_double] * 4 + [_dp])
    L.xo_bilinear.restype = ctypes.c_double
    L.xo_bilinear.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                              ctypes.c_double]
    L.xo_find_src_nrg_range.restype = None
    L.xo_find_src_nrg_range.argtypes = [ctypes.c_int] * 3 + [_ip, _ip]
    L.xo_wvt_value.restype = None
    L.xo_wvt_value.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 3 + [_dp, _dp]
    L.xo_scatter_geometry.restype = None
    L.xo_scatter_geometry.argtypes = [ctypes.POINTER(Geometry)] + [ctypes.c_int] * 5 + [_dp]
    L.xo_forward.restype = ctypes.c_int
    L.xo_forward.argtypes = [ctypes.POINTER(Geometry), ctypes.POINTER(Discretization),
                             _dp, _dp, _ip, _dp, _dp, _dp] + [ctypes.c_double] * 4 + \
                            [_ip, ctypes.c_int, _ip, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int]
    for name in ("xo_project", "xo_backproject"):
        f = getattr(L, name)
        f.restype = None
        f.argtypes = [_dp, ctypes.c_int64, ctypes.c_int, _dp, _dp, ctypes.c_int]
    L.xo_sensitivity.restype = None
    L.xo_sensitivity.argtypes = [_dp, ctypes.c_int64, ctypes.c_int, _dp, ctypes.c_int]
    L.xo_mlem.restype = None
    L.xo_mlem.argtypes = [_dp, ctypes.c_int64, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int,
                          ctypes.c_int]
    L.xo_mlem_update.restype = None
    L.xo_mlem_update.argtypes = [_dp, ctypes.c_int64, ctypes.c_int, _dp, _dp, _dp, _dp,
                                 ctypes.c_int]
    L.xo_loglik.restype = ctypes.c_double
    L.xo_loglik.argtypes = [_dp, ctypes.c_int64, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int]
    L.xo_max_threads.restype = ctypes.c_int
    L.xo_channel_match.restype = ctypes.c_double
    L.xo_channel_match.argtypes = [ctypes.c_double] * 5
    return L


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


def get_intersection(x1, y1, x2, y2, nx, ny):
    rx, ry = ctypes.c_double(), ctypes.c_double()
    side = lib().xo_get_intersection(x1, y1, x2, y2, nx, ny, ctypes.byref(rx), ctypes.byref(ry))
    return side, rx.value, ry.value


def trace_ray(x1, y1, x2, y2, nx, ny, step, max_samples=0):
    ex, ey, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int32()
    tex = np.zeros(2 * max(max_samples, 1), dtype=np.int32)
    side = lib().xo_trace_ray(x1, y1, x2, y2, nx, ny, step, ctypes.byref(ex), ctypes.byref(ey),
                              ctypes.byref(n), max_samples,
                              tex.ctypes.data_as(_ip) if max_samples > 0 else None)
    return side, ex.value, ey.value, n.value, tex[:2 * max_samples].reshape(-1, 2)


def line_integrals(maps, nx, ny, step, x1, y1, x2, y2):
    """maps: [nb, nx*ny] (x fastest within a map)."""
    m, mp = _d(np.asarray(maps).reshape(-1))
    nb = m.size // (nx * ny)
    out = np.zeros(nb)
    lib().xo_line_integrals(mp, nb, nx, ny, step, x1, y1, x2, y2, out.ctypes.data_as(_dp))
    return out


def bilinear(map_, nx, ny, x, y):
    m, mp = _d(np.asarray(map_).reshape(-1))
    return lib().xo_bilinear(mp, nx, ny, x, y)


def find_src_nrg_range(nrg, n_energy, halfwidth):
    lo, hi = ctypes.c_int32(), ctypes.c_int32()
    lib().xo_find_src_nrg_range(nrg, n_energy, halfwidth, ctypes.byref(lo), ctypes.byref(hi))
    return lo.value, hi.value


def wvt_value(n, n_q, q_min, q_max, dq0):
    q, dq = ctypes.c_double(), ctypes.c_double()
    lib().xo_wvt_value(n, n_q, q_min, q_max, dq0, ctypes.byref(q), ctypes.byref(dq))
    return q.value, dq.value


GEOM_FIELDS = ("src_x", "src_y", "det_x", "det_y", "cos_theta", "sin_half", "anorm2",
               "bnorm2", "res2", "geom", "rayleigh", "vox_z", "det_z")


def scatter_geometry(g: Geometry, view, col, row, ix, iy):
    out = np.zeros(len(GEOM_FIELDS))
    lib().xo_scatter_geometry(ctypes.byref(g), view, col, row, ix, iy, out.ctypes.data_as(_dp))
    return dict(zip(GEOM_FIELDS, out))


def _col_list(g, cols):
    if cols is None:
        return None, 0, g.n_cols
    cl = np.ascontiguousarray(cols, np.int32)
    if cl.size == 0 or cl.min() < 0 or cl.max() >= g.n_cols:
        raise ValueError("column list out of range")
    return cl, cl.size, cl.size


def forward(prob, views=None, with_A=True, nthreads=0, disc=None, cols=None):
    """Oracle forward model of a synthetic.Problem for the listed views and detector
    columns.  Returns (N, C, A) float64 with SRC replaced by len(views) and COL by
    len(cols); A is [rows, cols]."""
    g = prob.geometry
    d = disc if disc is not None else prob.disc
    vl = np.arange(g.n_views, dtype=np.int32) if views is None else np.asarray(views, np.int32)
    nvl = len(vl)
    cl, ncl, ncol = _col_list(g, cols)
    rows = nvl * ncol * g.n_rows * d.n_energy
    N = np.zeros(rows)
    C = np.zeros(rows)
    A = np.zeros(rows * d.n_q * d.n_mat) if with_A else None
    sp, spp = _d(prob.flat("spectrum"))
    fi, fip = _d(prob.flat("ifine"))
    sg, sgp = _i(prob.flat("segmentation"))
    a1, a1p = _d(prob.flat("a1"))
    a2, a2p = _d(prob.flat("a2"))
    s2, s2p = _d(prob.flat("s2_table"))
    vv, vvp = _i(vl)
    rc = lib().xo_forward(ctypes.byref(g), ctypes.byref(d), spp, fip, sgp, a1p, a2p, s2p,
                          *[float(v) for v in prob.lut], vvp, nvl,
                          cl.ctypes.data_as(_ip) if cl is not None else None, ncl,
                          N.ctypes.data_as(_dp),
                          C.ctypes.data_as(_dp),
                          A.ctypes.data_as(_dp) if with_A else None, nthreads)
    if rc != 0:
        raise RuntimeError(f"xo_forward failed: {rc}")
    if with_A:
        A = A.reshape(rows, d.n_q * d.n_mat)
    return N, C, A


def project(A, x, nthreads=0):
    A, Ap = _d(A)
    x, xp = _d(x)
    out = np.zeros(A.shape[0])
    lib().xo_project(Ap, A.shape[0], A.shape[1], xp, out.ctypes.data_as(_dp), nthreads)
    return out


def backproject(A, w, nthreads=0):
    A, Ap = _d(A)
    w, wp = _d(w)
    out = np.zeros(A.shape[1])
    lib().xo_backproject(Ap, A.shape[0], A.shape[1], wp, out.ctypes.data_as(_dp), nthreads)
    return out


def sensitivity(A, nthreads=0):
    A, Ap = _d(A)
    out = np.zeros(A.shape[1])
    lib().xo_sensitivity(Ap, A.shape[0], A.shape[1], out.ctypes.data_as(_dp), nthreads)
    return out


def mlem(A, y, c, x0, n_iter, nthreads=0):
    A, Ap = _d(A)
    y, yp = _d(y)
    c, cp = _d(c)
    x = np.array(x0, dtype=np.float64, copy=True)
    lib().xo_mlem(Ap, A.shape[0], A.shape[1], yp, cp, x.ctypes.data_as(_dp), n_iter, nthreads)
    return x


def mlem_update(A, y, c, s, x, nthreads=0):
    """One MLEM update of x (float64 array) in place given s = sum(A, 1)'."""
    A, Ap = _d(A)
    y, yp = _d(y)
    c, cp = _d(c)
    s, sp = _d(s)
    assert x.dtype == np.float64 and x.flags.c_contiguous and x.shape == (A.shape[1],)
    lib().xo_mlem_update(Ap, A.shape[0], A.shape[1], yp, cp, sp, x.ctypes.data_as(_dp), nthreads)
    return x


def loglik(A, y, c, x, nthreads=0):
    A, Ap = _d(A)
    y, yp = _d(y)
    c, cp = _d(c)
    x, xp = _d(x)
    return lib().xo_loglik(Ap, A.shape[0], A.shape[1], yp, cp, xp, nthreads)


def channel_match(es1, es2, eratio, ed1, ed2):
    return lib().xo_channel_match(es1, es2, eratio, ed1, ed2)


def max_threads():
    return lib().xo_max_threads()


def _cm_args(cp):
    p = cp.base
    sp = _d(p.flat("spectrum"))
    sg = _i(p.flat("segmentation"))
    co = _d(cp.coef.reshape(-1))
    ba = _d(cp.basis.reshape(-1))
    mp = _d(cp.maps.reshape(-1))
    be = _d(cp.beta.reshape(-1))
    return (sp, sg, co, ba, mp, be)


def forward_coeffs(cp, nthreads=0, coef=None, maps=None):
    """Counts of the coefficient model (no background), float64 [rows]."""
    L = lib()
    L.xo_forward_coeffs.restype = ctypes.c_int
    L.xo_forward_coeffs.argtypes = [ctypes.POINTER(Geometry), ctypes.POINTER(Discretization),
                                    _dp, _ip, _dp, ctypes.c_int, _dp, ctypes.c_int, _dp, _dp,
                                    _dp, ctypes.c_int]
    p = cp.base
    sp, sg, co, ba, mp, be = _cm_args(cp)
    if coef is not None:
        co = _d(np.asarray(coef).reshape(-1))
    if maps is not None:
        mp = _d(np.asarray(maps).reshape(-1))
    out = np.zeros(p.n_rows)
    rc = L.xo_forward_coeffs(ctypes.byref(p.geometry), ctypes.byref(p.disc), sp[1], sg[1], co[1],
                             cp.K, ba[1], cp.B, mp[1], be[1], out.ctypes.data_as(_dp), nthreads)
    if rc:
        raise RuntimeError(f"xo_forward_coeffs failed: {rc}")
    return out


def vjp_coeffs(cp, w, nthreads=0, want_coef=True, want_maps=True):
    """(grad_coef [npix, K], grad_maps [B, npix]) for upstream gradient w [rows]."""
    L = lib()
    L.xo_vjp_coeffs.restype = ctypes.c_int
    L.xo_vjp_coeffs.argtypes = [ctypes.POINTER(Geometry), ctypes.POINTER(Discretization),
                                _dp, _ip, _dp, ctypes.c_int, _dp, ctypes.c_int, _dp, _dp, _dp,
                                _dp, _dp, ctypes.c_int]
    p = cp.base
    sp, sg, co, ba, mp, be = _cm_args(cp)
    ww = _d(w)
    npix = p.geometry.nx * p.geometry.ny
    gc = np.zeros(npix * cp.K) if want_coef else None
    gm = np.zeros(cp.B * npix) if want_maps else None
    rc = L.xo_vjp_coeffs(ctypes.byref(p.geometry), ctypes.byref(p.disc), sp[1], sg[1], co[1],
                         cp.K, ba[1], cp.B, mp[1], be[1], ww[1],
                         gc.ctypes.data_as(_dp) if want_coef else None,
                         gm.ctypes.data_as(_dp) if want_maps else None, nthreads)
    if rc:
        raise RuntimeError(f"xo_vjp_coeffs failed: {rc}")
    return (gc.reshape(npix, cp.K) if want_coef else None,
            gm.reshape(cp.B, npix) if want_maps else None)


def ct_chord(sx, sy, dx, dy, nx, ny):
    """(hit, entry_x, entry_y, exit_x, exit_y) of the segment clipped to the grid."""
    L = lib()
    L.xo_ct_chord.restype = ctypes.c_int
    L.xo_ct_chord.argtypes = [ctypes.c_double] * 4 + [ctypes.c_int] * 2 + [_dp] * 4
    r = [ctypes.c_double() for _ in range(4)]
    hit = L.xo_ct_chord(sx, sy, dx, dy, nx, ny, *[ctypes.byref(x) for x in r])
    return bool(hit), r[0].value, r[1].value, r[2].value, r[3].value


def forward_ct(prob, maps, beta, nthreads=0, spectrum=None):
    """CT transmission T [V*COL*NRG] (v fastest) of maps [B, nx*ny] with beta [B, NRG]."""
    L = lib()
    L.xo_forward_ct.restype = ctypes.c_int
    L.xo_forward_ct.argtypes = [ctypes.POINTER(Geometry), ctypes.POINTER(Discretization), _dp,
                                ctypes.c_int, _dp, _dp, _dp, ctypes.c_int]
    g, d = prob.geometry, prob.disc
    sp = _d(prob.flat("spectrum") if spectrum is None else np.asarray(spectrum).reshape(-1))
    maps = np.asarray(maps, np.float64)
    B = maps.shape[0]
    mp, be = _d(maps.reshape(-1)), _d(np.asarray(beta).reshape(-1))
    out = np.zeros(g.n_views * g.n_cols * d.n_energy)
    rc = L.xo_forward_ct(ctypes.byref(g), ctypes.byref(d), sp[1], B, mp[1], be[1],
                         out.ctypes.data_as(_dp), nthreads)
    if rc:
        raise RuntimeError(f"xo_forward_ct failed: {rc}")
    return out


def vjp_ct(prob, maps, beta, w, nthreads=0):
    """grad_maps [B, nx*ny] of <w, T(maps)> (xo_vjp_ct)."""
    L = lib()
    L.xo_vjp_ct.restype = ctypes.c_int
    L.xo_vjp_ct.argtypes = [ctypes.POINTER(Geometry), ctypes.POINTER(Discretization), _dp,
                            ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_int]
    g, d = prob.geometry, prob.disc
    sp = _d(prob.flat("spectrum"))
    maps = np.asarray(maps, np.float64)
    B = maps.shape[0]
    mp, be, ww = _d(maps.reshape(-1)), _d(np.asarray(beta).reshape(-1)), _d(w)
    gm = np.zeros(B * g.nx * g.ny)
    rc = L.xo_vjp_ct(ctypes.byref(g), ctypes.byref(d), sp[1], B, mp[1], be[1], ww[1],
                     gm.ctypes.data_as(_dp), nthreads)
    if rc:
        raise RuntimeError(f"xo_vjp_ct failed: {rc}")
    return gm.reshape(B, g.nx * g.ny)


def line_integrals_3d(vols, nx, ny, nz, step, x1, y1, z1, x2, y2, z2):
    """Line integrals of volumes [nb, nx*ny*nz] (x fastest) along a 3-D segment."""
    L = lib()
    L.xo_line_integrals_3d.restype = None
    L.xo_line_integrals_3d.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_double] + [ctypes.c_double] * 6 + [_dp]
    v = np.ascontiguousarray(np.asarray(vols, np.float64).reshape(-1))
    nb = v.size // (nx * ny * nz)
    out = np.zeros(nb)
    L.xo_line_integrals_3d(v.ctypes.data_as(_dp), nb, nx, ny, nz, step, x1, y1, z1, x2, y2, z2,
                           out.ctypes.data_as(_dp))
    return out


def forward_slice(vp, n_slices, slice_, views=None, with_A=True, nthreads=0, cols=None):
    """Oracle forward model of slice slice_ of a volume problem (segmentation, a1, a2
    [nx, ny, n_slices]).  Returns (N, C, A) like forward()."""
    g, d = vp.geometry, vp.disc
    vl = np.arange(g.n_views, dtype=np.int32) if views is None else np.asarray(views, np.int32)
    nvl = len(vl)
    cl, ncl, ncol = _col_list(g, cols)
    rows = nvl * ncol * g.n_rows * d.n_energy
    N = np.zeros(rows)
    C = np.zeros(rows)
    A = np.zeros(rows * d.n_q * d.n_mat) if with_A else None
    sp, spp = _d(vp.flat("spectrum"))
    fi, fip = _d(vp.flat("ifine"))
    sg, sgp = _i(vp.flat("segmentation"))
    a1, a1p = _d(vp.flat("a1"))
    a2, a2p = _d(vp.flat("a2"))
    s2, s2p = _d(vp.flat("s2_table"))
    vv, vvp = _i(vl)
    L = lib()
    L.xo_forward_slice.restype = ctypes.c_int
    rc = L.xo_forward_slice(ctypes.byref(g), ctypes.byref(d), spp, fip, sgp, a1p, a2p,
                            ctypes.c_int(n_slices), ctypes.c_int(slice_), s2p,
                            *[ctypes.c_double(float(v)) for v in vp.lut], vvp, ctypes.c_int(nvl),
                            cl.ctypes.data_as(_ip) if cl is not None else None, ctypes.c_int(ncl),
                            N.ctypes.data_as(_dp), C.ctypes.data_as(_dp),
                            A.ctypes.data_as(_dp) if with_A else None, ctypes.c_int(nthreads))
    if rc != 0:
        raise RuntimeError(f"xo_forward_slice failed: {rc}")
    if with_A:
        A = A.reshape(rows, d.n_q * d.n_mat)
    return N, C, A
