"""Host-side mirror of the reference's interface for the forward projector and MLEM.

``Forward_XRD`` keeps the argument order and meaning of the MATLAB MEX gateway
``[Nphotons, Cphotons, Amatrix] = Forward_XRD(spectrum, Ifine, segmentation,
a1, a2, S2_grid, q_min, q_max, a_min, a_max)`` (Forward_XRD.cu:76-302,
called at main.m:56-64) and returns the same three arrays in the same memory
layout (first index fastest): torch tensors shaped in reverse index order,
i.e. Nphotons as [NRG, ROW, COL, SRC] and Amatrix as
[NRG, ROW, COL, SRC, MAT, NWT].  The geometry and discretisation that the
reference hard-codes are passed as ``geometry`` / ``disc``.

``mlem`` is the iterative solver of main.m:153-176.  torch is used for device
memory and streams only; all arithmetic runs in libxrd_tomo.so.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Tuple

import numpy as np
import torch  # loaded before libxrd_tomo.so so both share one HIP runtime

from . import _lib
from ._lib import (DTYPE_BF16, DTYPE_BF24, DTYPE_F16, DTYPE_F32, MLEM_FUSED, MLEM_UNFUSED,
                   XrdError, check, lib)
from .descriptors import Bf16TablesStruct, Discretization, Geometry

_DTYPES = {torch.float32: DTYPE_F32, torch.bfloat16: DTYPE_BF16, torch.float16: DTYPE_F16}

#: dtype argument of ``convert_matrix`` selecting the 3-byte BF24 storage.
BF24 = "bf24"


class Bf24Matrix:
    """A [rows x cols] matrix in BF24 storage (XRD_DTYPE_BF24): FP32 rounded to nearest
    even at 24 bits, kept as one uint8 buffer of 3*rows*cols bytes -- the uint16 plane
    of bits 31..16, then the uint8 plane of bits 15..8.  Accepted wherever the matrix
    operators, MlemPlan and MlemBatch take A."""

    dtype = BF24

    def __init__(self, data: torch.Tensor, n_rows: int, n_cols: int):
        if data.dtype != torch.uint8 or data.dim() != 1 or data.numel() != 3 * n_rows * n_cols:
            raise ValueError("BF24 storage is a flat uint8 tensor of 3*rows*cols bytes")
        self.data = data
        self.shape = (int(n_rows), int(n_cols))
        self.device = data.device

    def data_ptr(self) -> int:
        return self.data.data_ptr()

    def numel(self) -> int:
        return self.shape[0] * self.shape[1]

    def float(self) -> torch.Tensor:
        """The stored values as an FP32 tensor (exact)."""
        n = self.numel()
        hi = self.data[:2 * n].view(torch.int16).to(torch.int32) & 0xFFFF
        lo = self.data[2 * n:].to(torch.int32)
        return ((hi << 16) | (lo << 8)).view(torch.float32).view(self.shape)


def _ptr(t: Optional[torch.Tensor]):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _require_gpu(*ts):
    """Every tensor given (None entries skipped) is a contiguous tensor on one GPU."""
    dev = None
    for t in ts:
        if t is None:
            continue
        if not isinstance(t, torch.Tensor):
            raise TypeError(f"expected a torch.Tensor, got {type(t).__name__}")
        if not t.is_cuda:
            raise ValueError("xrd_tomography_amd operates on GPU tensors only")
        if not t.is_contiguous():
            raise ValueError("tensors must be contiguous")
        if dev is None:
            dev = t.device
        elif t.device != dev:
            raise ValueError(f"tensors on different devices: {dev} and {t.device}")


def _check(t, name: str, dtype, numel: Optional[int] = None, optional: bool = False):
    """Type, dtype and element count of one argument, checked on the host before the
    library sees its pointer (the C ABI cannot see tensor sizes)."""
    if t is None:
        if optional:
            return
        raise ValueError(f"{name} is required")
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name}: expected a torch.Tensor, got {type(t).__name__}")
    if t.dtype != dtype:
        raise TypeError(f"{name}: expected {dtype}, got {t.dtype}")
    if numel is not None and t.numel() != numel:
        raise ValueError(f"{name}: expected {numel} elements, got {t.numel()}")


class Context:
    """Owns an xrd_context (device workspace).  One context per host thread."""

    def __init__(self, device: int = 0):
        if not torch.cuda.is_available():
            raise RuntimeError("xrd_tomography_amd needs a ROCm GPU (torch.cuda unavailable)")
        torch.cuda.init()
        h = ctypes.c_void_p()
        check(lib().xrd_context_create(device, ctypes.byref(h)), "xrd_context_create")
        self.handle = h
        self.device = device

    def close(self):
        if self.handle:
            torch.cuda.synchronize(self.device)
            lib().xrd_context_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def workspace_bytes(self) -> int:
        return int(lib().xrd_context_workspace_bytes(self.handle))


_default_ctx = {}


def default_context(device: int = 0) -> Context:
    if device not in _default_ctx:
        _default_ctx[device] = Context(device)
    return _default_ctx[device]


def num_rows(g: Geometry, d: Discretization) -> int:
    return int(lib().xrd_num_rows(ctypes.byref(g), ctypes.byref(d)))


def Forward_XRD(spectrum: torch.Tensor, Ifine: torch.Tensor, segmentation: torch.Tensor,
                a1: torch.Tensor, a2: torch.Tensor, S2_grid: torch.Tensor,
                q_min: float, q_max: float, a_min: float, a_max: float, *,
                geometry: Geometry, disc: Discretization, with_A: bool = True,
                ctx: Optional[Context] = None, stream=None,
                out: Optional[Tuple[torch.Tensor, torch.Tensor, Optional[torch.Tensor]]] = None):
    """Forward projector (Forward_XRD.cu:76-302).  Inputs are flat device tensors in
    the reference's layouts (spectrum [NRG*NRG] float32, Ifine [NQF*MAT] float32,
    segmentation [NX*NY] int32, a1/a2 [NX*NY] float32 in 1/voxel, S2_grid [NQ*NA]).
    Returns (Nphotons, Cphotons, Amatrix); Amatrix is None if with_A is False."""
    g, d = geometry, disc
    _check_forward_inputs(g, d, spectrum, Ifine, segmentation, a1, a2, S2_grid, 1)
    shape_n = (d.n_energy, g.n_rows, g.n_cols, g.n_views)
    rows = g.n_views * g.n_cols * g.n_rows * d.n_energy
    if out is None:
        _require_gpu(spectrum, Ifine, segmentation, a1, a2, S2_grid)
        dev = spectrum.device
        N = torch.empty(shape_n, dtype=torch.float32, device=dev)
        C = torch.empty(shape_n, dtype=torch.float32, device=dev)
        A = torch.empty(shape_n + (d.n_mat, d.n_q), dtype=torch.float32, device=dev) if with_A else None
    else:
        N, C, A = out
        _check(N, "out[0] (Nphotons)", torch.float32, rows)
        _check(C, "out[1] (Cphotons)", torch.float32, rows)
        _check(A, "out[2] (Amatrix)", torch.float32, rows * d.n_q * d.n_mat, optional=True)
        _require_gpu(spectrum, Ifine, segmentation, a1, a2, S2_grid, N, C, A)
    ctx = ctx or default_context(spectrum.device.index or 0)
    rc = lib().xrd_forward(ctx.handle, ctypes.byref(g), ctypes.byref(d), _ptr(spectrum),
                           _ptr(Ifine), _ptr(segmentation), _ptr(a1), _ptr(a2), _ptr(S2_grid),
                           float(q_min), float(q_max), float(a_min), float(a_max),
                           _ptr(N), _ptr(C), _ptr(A), _stream(stream))
    check(rc, "xrd_forward")
    return N, C, A


def _check_forward_inputs(g, d, spectrum, Ifine, segmentation, a1, a2, S2_grid, n_slices):
    npix = g.nx * g.ny * n_slices
    _check(spectrum, "spectrum", torch.float32, d.n_energy * d.n_energy)
    _check(Ifine, "Ifine", torch.float32, d.n_qf * d.n_mat)
    _check(segmentation, "segmentation", torch.int32, npix)
    _check(a1, "a1", torch.float32, npix)
    _check(a2, "a2", torch.float32, npix)
    _check(S2_grid, "S2_grid", torch.float32, d.lut_nq * d.lut_na)


def forward_slice(spectrum, Ifine, segmentation, a1, a2, S2_grid, q_min, q_max, a_min, a_max,
                  n_slices: int, slice_: int, *, geometry: Geometry, disc: Discretization,
                  with_A: bool = True, ctx: Optional[Context] = None, stream=None):
    """xrd_forward_slice: slice slice_ of a volume (segmentation, a1, a2 flat
    [NX*NY*n_slices] device tensors, x fastest); returns (N, C, A) of that slice."""
    g, d = geometry, disc
    if not 0 <= int(slice_) < int(n_slices):
        raise ValueError(f"slice {slice_} outside [0, {n_slices})")
    _check_forward_inputs(g, d, spectrum, Ifine, segmentation, a1, a2, S2_grid, int(n_slices))
    _require_gpu(spectrum, Ifine, segmentation, a1, a2, S2_grid)
    ctx = ctx or default_context(spectrum.device.index or 0)
    shape_n = (d.n_energy, g.n_rows, g.n_cols, g.n_views)
    dev = spectrum.device
    N = torch.empty(shape_n, dtype=torch.float32, device=dev)
    C = torch.empty(shape_n, dtype=torch.float32, device=dev)
    A = torch.empty(shape_n + (d.n_mat, d.n_q), dtype=torch.float32, device=dev) if with_A else None
    rc = lib().xrd_forward_slice(ctx.handle, ctypes.byref(g), ctypes.byref(d), _ptr(spectrum),
                                 _ptr(Ifine), _ptr(segmentation), _ptr(a1), _ptr(a2), n_slices,
                                 slice_, _ptr(S2_grid), float(q_min), float(q_max), float(a_min),
                                 float(a_max), _ptr(N), _ptr(C), _ptr(A), _stream(stream))
    check(rc, "xrd_forward_slice")
    return N, C, A


def Forward_XRD_bf16(spectrum, Ifine_bf16, Ifine_scale, segmentation, a1_bf16, a2_bf16,
                     mu_scale, S2_grid, q_min, q_max, a_min, a_max, n_slices: int = 1,
                     slice_: int = 0, *, geometry: Geometry, disc: Discretization,
                     with_A: bool = True, ctx: Optional[Context] = None, stream=None):
    """xrd_forward_bf16: Forward_XRD (n_slices = 1) or forward_slice with the form
    factors and the attenuation maps stored as bfloat16 (torch.bfloat16 tensors) with
    per-material FP32 scales, decoded on load by the kernels.  Ifine_scale is [MAT],
    mu_scale [2*(MAT+1)] (a1 scales by segmentation id, then a2's)."""
    g, d = geometry, disc
    if not 0 <= int(slice_) < int(n_slices):
        raise ValueError(f"slice {slice_} outside [0, {n_slices})")
    npix = g.nx * g.ny * int(n_slices)
    _check(spectrum, "spectrum", torch.float32, d.n_energy * d.n_energy)
    _check(Ifine_bf16, "Ifine_bf16", torch.bfloat16, d.n_qf * d.n_mat)
    _check(Ifine_scale, "Ifine_scale", torch.float32, d.n_mat)
    _check(segmentation, "segmentation", torch.int32, npix)
    _check(a1_bf16, "a1_bf16", torch.bfloat16, npix)
    _check(a2_bf16, "a2_bf16", torch.bfloat16, npix)
    _check(mu_scale, "mu_scale", torch.float32, 2 * (d.n_mat + 1))
    _check(S2_grid, "S2_grid", torch.float32, d.lut_nq * d.lut_na)
    _require_gpu(spectrum, Ifine_bf16, Ifine_scale, segmentation, a1_bf16, a2_bf16, mu_scale,
                 S2_grid)
    ctx = ctx or default_context(spectrum.device.index or 0)
    shape_n = (d.n_energy, g.n_rows, g.n_cols, g.n_views)
    dev = spectrum.device
    N = torch.empty(shape_n, dtype=torch.float32, device=dev)
    C = torch.empty(shape_n, dtype=torch.float32, device=dev)
    A = torch.empty(shape_n + (d.n_mat, d.n_q), dtype=torch.float32, device=dev) if with_A else None
    tab = Bf16TablesStruct(Ifine_bf16.data_ptr(), Ifine_scale.data_ptr(), a1_bf16.data_ptr(),
                           a2_bf16.data_ptr(), mu_scale.data_ptr())
    rc = lib().xrd_forward_bf16(ctx.handle, ctypes.byref(g), ctypes.byref(d), _ptr(spectrum),
                                ctypes.byref(tab), _ptr(segmentation), int(n_slices), int(slice_),
                                _ptr(S2_grid), float(q_min), float(q_max), float(a_min),
                                float(a_max), _ptr(N), _ptr(C), _ptr(A), _stream(stream))
    check(rc, "xrd_forward_bf16")
    return N, C, A


def forward_problem_bf16(prob, n_slices: int = 1, slice_: int = 0, device="cuda",
                         with_A=True, ctx=None):
    """Forward_XRD_bf16 on a synthetic.Problem (2-D, or a volume with n_slices) whose
    tables are encoded by synthetic.encode_bf16_tables."""
    from .synthetic import encode_bf16_tables
    t = encode_bf16_tables(prob)
    f32 = lambda a, dt=torch.float32: torch.from_numpy(np.ascontiguousarray(a)).to(device=device, dtype=dt)  # noqa: E731
    b16 = lambda a: torch.from_numpy(np.ascontiguousarray(a).view(np.int16)).to(device).view(torch.bfloat16)  # noqa: E731
    return Forward_XRD_bf16(f32(prob.flat("spectrum")), b16(t.ifine), f32(t.ifine_scale),
                            f32(prob.flat("segmentation"), torch.int32), b16(t.a1), b16(t.a2),
                            f32(t.mu_scale), f32(prob.flat("s2_table")), *prob.lut, n_slices,
                            slice_, geometry=prob.geometry, disc=prob.disc, with_A=with_A, ctx=ctx)


STREAM_SOURCES = ("csrc/matrix_stream.hpp", "csrc/mlem.hip")


def stream_kernels_tag() -> str:
    """First 16 hex digits of the SHA-256 of the streaming kernels' sources shipped with
    the package ("nosrc" without them): tags launch shapes tuned for, and traffic
    measured on, those kernels."""
    import hashlib
    import os
    here = os.path.dirname(os.path.abspath(__file__))
    h = hashlib.sha256()
    for rel in STREAM_SOURCES:
        path = os.path.join(here, rel)
        if not os.path.exists(path):
            return "nosrc"
        with open(path, "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def _cache_table(cache: Optional[str]) -> dict:
    import json
    import os
    if cache and os.path.exists(cache):
        with open(cache) as f:
            return json.load(f)
    return {}


def _cache_store(cache: Optional[str], key: str, value: dict):
    import json
    import os
    if not cache:
        return
    table = _cache_table(cache)
    table[key] = value
    os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
    with open(cache, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)


def set_forward_launch(voxels_per_chunk: int = 0, views_per_batch: int = 0, ctx=None):
    """xrd_forward_set_launch on the (default) context: the projector's voxel chunk and
    view batch (0 = default) for the forward calls that follow."""
    ctx = ctx or default_context(0)
    check(lib().xrd_forward_set_launch(ctx.handle, int(voxels_per_chunk), int(views_per_batch)),
          "xrd_forward_set_launch")


def autotune_forward(prob, sample_views: int = 32, cache: Optional[str] = None, device="cuda",
                     ctx=None) -> dict:
    """Pick the projector's launch shape for a synthetic.Problem (xrd_forward_autotune on
    its first sample_views views) and keep it on the context for later forward calls.
    With `cache` (a JSON file) an earlier result for the same device and problem shape
    is reused instead of re-timing, and new results are added to it."""
    g, d = prob.geometry, prob.disc
    ctx = ctx or default_context(0)
    nvox = int(((prob.segmentation > 0) & (prob.segmentation <= d.n_mat)).sum())
    key = (f"forward|{torch.cuda.get_device_name(ctx.device)}|{g.nx}x{g.ny}|cols {g.n_cols}"
           f"|rows {g.n_rows}|E {d.n_energy}|Q {d.n_q}|QF {d.n_qf}|M {d.n_mat}|vox {nvox}")
    hit = _cache_table(cache).get(key)
    if hit is not None:
        set_forward_launch(hit["voxels_per_chunk"], hit["views_per_batch"], ctx=ctx)
        return dict(hit, cached=True)

    def dev(name, dt):
        return torch.from_numpy(prob.flat(name)).to(device=device, dtype=dt)
    sp, fi, sg = dev("spectrum", torch.float32), dev("ifine", torch.float32), dev("segmentation", torch.int32)
    a1, a2, s2 = dev("a1", torch.float32), dev("a2", torch.float32), dev("s2_table", torch.float32)
    kc, vpb, ms = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_float()
    check(lib().xrd_forward_autotune(ctx.handle, ctypes.byref(g), ctypes.byref(d), _ptr(sp),
                                     _ptr(fi), _ptr(sg), _ptr(a1), _ptr(a2), _ptr(s2),
                                     *[float(v) for v in prob.lut], int(sample_views),
                                     _stream(None), ctypes.byref(kc), ctypes.byref(vpb),
                                     ctypes.byref(ms)), "xrd_forward_autotune")
    res = {"voxels_per_chunk": kc.value, "views_per_batch": vpb.value,
           "ms_sample": round(ms.value, 3), "sample_views": min(int(sample_views), g.n_views)}
    _cache_store(cache, key, res)
    return dict(res, cached=False)


def forward_volume_slice(vp, n_slices: int, slice_: int, device="cuda", with_A=True, ctx=None):
    """forward_slice on a volume Problem (segmentation, a1, a2 [nx, ny, n_slices])."""
    def dev(name, dt):
        return torch.from_numpy(vp.flat(name)).to(device=device, dtype=dt)
    return forward_slice(dev("spectrum", torch.float32), dev("ifine", torch.float32),
                         dev("segmentation", torch.int32), dev("a1", torch.float32),
                         dev("a2", torch.float32), dev("s2_table", torch.float32), *vp.lut,
                         n_slices, slice_, geometry=vp.geometry, disc=vp.disc, with_A=with_A,
                         ctx=ctx)


def forward_problem(prob, device="cuda", with_A=True, ctx=None, disc=None):
    """Run Forward_XRD on a synthetic.Problem; returns (N, C, A) device tensors."""
    def dev(name, dt):
        return torch.from_numpy(prob.flat(name)).to(device=device, dtype=dt)
    return Forward_XRD(dev("spectrum", torch.float32), dev("ifine", torch.float32),
                       dev("segmentation", torch.int32), dev("a1", torch.float32),
                       dev("a2", torch.float32), dev("s2_table", torch.float32), *prob.lut,
                       geometry=prob.geometry, disc=disc if disc is not None else prob.disc,
                       with_A=with_A, ctx=ctx)


def trace_rays(segments: torch.Tensor, nx: int, ny: int, step: float = 0.25,
               max_samples: int = 0, maps: Optional[torch.Tensor] = None,
               ctx: Optional[Context] = None, stream=None):
    """GetIntersection + attenuator sampling for n segments [n, 4] float64 (x1,y1,x2,y2).
    Returns (side int32 [n], entry float64 [n,2], count int32 [n], texels int32
    [n, max_samples, 2] or None, integrals float32 [n_basis, n] or None)."""
    _check(segments, "segments", torch.float64)
    if segments.dim() != 2 or segments.shape[1] != 4:
        raise ValueError("segments must be [n, 4] (x1, y1, x2, y2)")
    if maps is not None:
        _check(maps, "maps", torch.float32)
        if maps.numel() == 0 or maps.numel() % (nx * ny) != 0:
            raise ValueError("maps must hold n_basis whole [nx*ny] maps")
    if max_samples < 0:
        raise ValueError("max_samples < 0")
    _require_gpu(segments, maps)
    ctx = ctx or default_context(segments.device.index or 0)
    n = segments.shape[0]
    dev = segments.device
    side = torch.empty(n, dtype=torch.int32, device=dev)
    entry = torch.empty((n, 2), dtype=torch.float64, device=dev)
    count = torch.empty(n, dtype=torch.int32, device=dev)
    tex = torch.empty((n, max_samples, 2), dtype=torch.int32, device=dev) if max_samples > 0 else None
    nb = 0 if maps is None else maps.numel() // (nx * ny)
    ints = torch.empty((nb, n), dtype=torch.float32, device=dev) if maps is not None else None
    rc = lib().xrd_trace_rays(ctx.handle, nx, ny, float(step), n, _ptr(segments), _ptr(side),
                              _ptr(entry), _ptr(count), max_samples, _ptr(tex), nb, _ptr(maps),
                              _ptr(ints), _stream(stream))
    check(rc, "xrd_trace_rays")
    return side, entry, count, tex, ints


def _mat(A):
    """(dtype code, rows, cols) of a model matrix on the GPU."""
    if isinstance(A, Bf24Matrix):
        _require_gpu(A.data)
        return DTYPE_BF24, A.shape[0], A.shape[1]
    if not isinstance(A, torch.Tensor):
        raise TypeError(f"A: expected a torch.Tensor or Bf24Matrix, got {type(A).__name__}")
    if A.dim() != 2:
        raise ValueError("A must be 2-D [rows, cols]")
    if A.dtype not in _DTYPES:
        raise TypeError(f"unsupported A dtype {A.dtype}")
    _require_gpu(A)
    return _DTYPES[A.dtype], A.shape[0], A.shape[1]


def _mat_tensor(A):
    return A.data if isinstance(A, Bf24Matrix) else A


def project(A: torch.Tensor, x: torch.Tensor, ctx=None, stream=None) -> torch.Tensor:
    """A x (FP32 output)."""
    dt, nr, nc = _mat(A)
    _check(x, "x", torch.float32, nc)
    _require_gpu(_mat_tensor(A), x)
    ctx = ctx or default_context(A.device.index or 0)
    out = torch.empty(nr, dtype=torch.float32, device=A.device)
    check(lib().xrd_project(ctx.handle, _ptr(A), dt, nr, nc, _ptr(x), _ptr(out),
                            _stream(stream)), "xrd_project")
    return out


def backproject(A: torch.Tensor, w: torch.Tensor, ctx=None, stream=None) -> torch.Tensor:
    """A' w with FP64 column sums."""
    dt, nr, nc = _mat(A)
    _check(w, "w", torch.float32, nr)
    _require_gpu(_mat_tensor(A), w)
    ctx = ctx or default_context(A.device.index or 0)
    out = torch.empty(nc, dtype=torch.float64, device=A.device)
    check(lib().xrd_backproject(ctx.handle, _ptr(A), dt, nr, nc, _ptr(w), _ptr(out),
                                _stream(stream)), "xrd_backproject")
    return out


def sensitivity(A: torch.Tensor, ctx=None, stream=None) -> torch.Tensor:
    """sum(A, 1)' with FP64 column sums."""
    dt, nr, nc = _mat(A)
    ctx = ctx or default_context(A.device.index or 0)
    out = torch.empty(nc, dtype=torch.float64, device=A.device)
    check(lib().xrd_sensitivity(ctx.handle, _ptr(A), dt, nr, nc, _ptr(out), _stream(stream)),
          "xrd_sensitivity")
    return out


def convert_matrix(A: torch.Tensor, dtype, ctx=None, stream=None):
    """FP32 -> BF16/F16 (torch tensor) or BF24 (``dtype=BF24``: a Bf24Matrix) storage,
    round to nearest even, with the library's kernel."""
    _check(A, "A", torch.float32)
    _require_gpu(A)
    ctx = ctx or default_context(A.device.index or 0)
    if dtype == BF24:
        if A.dim() != 2:
            raise ValueError("A must be 2-D [rows, cols]")
        out = torch.empty(3 * A.numel(), dtype=torch.uint8, device=A.device)
        check(lib().xrd_convert_matrix(ctx.handle, _ptr(A), A.numel(), DTYPE_BF24,
                                       _ptr(out), _stream(stream)), "xrd_convert_matrix")
        return Bf24Matrix(out, A.shape[0], A.shape[1])
    out = torch.empty(A.shape, dtype=dtype, device=A.device)
    check(lib().xrd_convert_matrix(ctx.handle, _ptr(A), A.numel(), _DTYPES[dtype], _ptr(out),
                                   _stream(stream)), "xrd_convert_matrix")
    return out


class MlemPlan:
    """MLEM of main.m:153-176 on an explicit model matrix resident on the GPU.
    x is updated in place; y, c and A must outlive the plan."""

    def __init__(self, A: torch.Tensor, y: torch.Tensor, c: torch.Tensor, x: torch.Tensor,
                 fused: bool = True, col_scale: Optional[torch.Tensor] = None, ctx=None,
                 stream=None):
        dt, nr, nc = _mat(A)
        _check_mlem_vectors(nr, nc, y, c, x, col_scale)
        _require_gpu(_mat_tensor(A), y, c, x, col_scale)
        self.ctx = ctx or default_context(A.device.index or 0)
        self._keep = (A, y, c, x, col_scale)
        self.x = x
        h = ctypes.c_void_p()
        check(lib().xrd_mlem_plan_create(self.ctx.handle, _ptr(A), dt, nr, nc, _ptr(col_scale),
                                         _ptr(y), _ptr(c), _ptr(x),
                                         MLEM_FUSED if fused else MLEM_UNFUSED,
                                         _stream(stream), ctypes.byref(h)),
              "xrd_mlem_plan_create")
        self.handle = h
        self.n_cols = nc

    def iterate(self, n_iter: int, use_graph: bool = False, stream=None):
        check(lib().xrd_mlem_plan_iterate(self.handle, int(n_iter), int(bool(use_graph)),
                                          _stream(stream)), "xrd_mlem_plan_iterate")
        return self.x

    def autotune(self, reps: int = 3, cache: Optional[str] = None, stream=None) -> dict:
        """Pick the launch shape of the fused pass (xrd_mlem_plan_autotune).  With
        `cache` (a JSON file path) an earlier result for the same matrix shape, dtype,
        device and streaming-kernel sources is reused instead of re-timing."""
        key = f"mlem|{torch.cuda.get_device_name(self.x.device)}|{self._keep[0].shape[0]}x" \
              f"{self.n_cols}|{self._keep[0].dtype}|kernels {stream_kernels_tag()}"
        hit = _cache_table(cache).get(key)
        if hit is not None:
            b, r = hit["blocks_per_cu"], hit["rows_per_wave"]
            check(lib().xrd_mlem_plan_set_launch(self.handle, b, r), "xrd_mlem_plan_set_launch")
            return dict(hit, cached=True)
        b, r, ms = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_float()
        check(lib().xrd_mlem_plan_autotune(self.handle, int(reps), _stream(stream),
                                           ctypes.byref(b), ctypes.byref(r), ctypes.byref(ms)),
              "xrd_mlem_plan_autotune")
        res = {"blocks_per_cu": b.value, "rows_per_wave": r.value,
               "ms_fused_pass": round(ms.value, 4)}
        _cache_store(cache, key, res)
        return dict(res, cached=False)

    def set_launch(self, blocks_per_cu: int = 0, rows_per_wave: int = 0):
        check(lib().xrd_mlem_plan_set_launch(self.handle, blocks_per_cu, rows_per_wave),
              "xrd_mlem_plan_set_launch")

    def time_pass(self, reps: int = 10, stream=None) -> float:
        """Average ms of the streaming pass over A alone (HIP events on the stream)."""
        ms = ctypes.c_float()
        check(lib().xrd_mlem_plan_time_pass(self.handle, int(reps), _stream(stream),
                                            ctypes.byref(ms)), "xrd_mlem_plan_time_pass")
        return ms.value

    def loglik(self, stream=None) -> float:
        v = ctypes.c_double()
        check(lib().xrd_mlem_plan_loglik(self.handle, ctypes.byref(v), _stream(stream)),
              "xrd_mlem_plan_loglik")
        return v.value

    def sensitivity(self) -> torch.Tensor:
        """Copy of the FP64 sensitivity vector computed at plan creation."""
        p = lib().xrd_mlem_plan_sensitivity(self.handle)
        return _wrap_device_ptr(p, self.n_cols, torch.float64, self.x.device).clone()

    def close(self):
        if getattr(self, "handle", None):
            lib().xrd_mlem_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _check_mlem_vectors(nr, nc, y, c, x, col_scale):
    _check(y, "y", torch.float32, nr)
    _check(c, "c", torch.float32, nr)
    _check(x, "x", torch.float32, nc)
    _check(col_scale, "col_scale", torch.float32, nc, optional=True)


def _wrap_device_ptr(ptr: int, n: int, dtype, device):
    """Non-owning torch view of n elements at a device pointer owned by the library."""
    class _Holder:
        pass
    h = _Holder()
    h.__cuda_array_interface__ = {
        "shape": (n,), "typestr": "<f8" if dtype == torch.float64 else "<f4",
        "data": (ptr, False), "version": 2,
    }
    return torch.as_tensor(h, device=device)


def convert_matrix_scaled(A: torch.Tensor, dtype: torch.dtype, ctx=None, stream=None):
    """(A16, col_scale): BF16/F16 storage of A*col_scale with power-of-two per-column
    scales chosen by the library (xrd_convert_matrix_scaled)."""
    _check(A, "A", torch.float32)
    if A.dim() != 2:
        raise ValueError("A must be 2-D [rows, cols]")
    if dtype not in (torch.bfloat16, torch.float16):
        raise TypeError("dtype must be torch.bfloat16 or torch.float16")
    _require_gpu(A)
    ctx = ctx or default_context(A.device.index or 0)
    out = torch.empty(A.shape, dtype=dtype, device=A.device)
    sc = torch.empty(A.shape[1], dtype=torch.float32, device=A.device)
    check(lib().xrd_convert_matrix_scaled(ctx.handle, _ptr(A), A.shape[0], A.shape[1],
                                          _DTYPES[dtype], _ptr(out), _ptr(sc), _stream(stream)),
          "xrd_convert_matrix_scaled")
    return out, sc


def mlem_iterate_static(A, y: torch.Tensor, c: torch.Tensor, x: torch.Tensor, n_iter: int,
                        col_scale: Optional[torch.Tensor] = None, ctx=None, stream=None):
    """n_iter MLEM updates of x in place through xrd_mlem_iterate_static: the fused plan's
    kernels and default launch shape with every buffer in the library's fixed device
    scratch, so the call allocates nothing (usable under graph capture).  Calls that
    overlap on one GPU race on that scratch."""
    dt, nr, nc = _mat(A)
    _check_mlem_vectors(nr, nc, y, c, x, col_scale)
    _require_gpu(_mat_tensor(A), y, c, x, col_scale)
    ctx = ctx or default_context(x.device.index or 0)
    check(lib().xrd_mlem_iterate_static(ctx.handle, _ptr(A), dt, nr, nc, _ptr(col_scale),
                                        _ptr(y), _ptr(c), _ptr(x), int(n_iter),
                                        _stream(stream)), "xrd_mlem_iterate_static")
    return x


def mlem(A: torch.Tensor, y: torch.Tensor, c: torch.Tensor, n_iter: int, x0: float = 10.0,
         fused: bool = True, use_graph: bool = False, col_scale=None, ctx=None) -> torch.Tensor:
    """n_iter MLEM updates from x = x0 (main.m:154) on device; returns x (float32)."""
    x = torch.full((A.shape[1],), float(x0), dtype=torch.float32, device=A.device)
    plan = MlemPlan(A, y, c, x, fused=fused, col_scale=col_scale, ctx=ctx)
    plan.iterate(n_iter, use_graph=use_graph)
    torch.cuda.current_stream().synchronize()
    plan.close()
    return x


# ---- differentiable coefficient model ------------------------------------------------

def forward_coeffs(spectrum, segmentation, coef, basis, maps, beta, *, geometry: Geometry,
                   disc: Discretization, ctx=None, stream=None) -> torch.Tensor:
    """Counts of the coefficient model (xrd_forward_coeffs): coef [nx*ny, K],
    basis [K, NWT], maps [B, nx*ny], beta [B, NRG], float32 device tensors."""
    g, d = geometry, disc
    _check_coeff_inputs(g, d, spectrum, segmentation, coef, basis, maps, beta)
    _require_gpu(spectrum, segmentation, coef, basis, maps, beta)
    ctx = ctx or default_context(spectrum.device.index or 0)
    out = torch.empty((d.n_energy, g.n_rows, g.n_cols, g.n_views), dtype=torch.float32,
                      device=spectrum.device)
    check(lib().xrd_forward_coeffs(ctx.handle, ctypes.byref(g), ctypes.byref(d), _ptr(spectrum),
                                   _ptr(segmentation), _ptr(coef), coef.shape[1], _ptr(basis),
                                   maps.shape[0], _ptr(maps), _ptr(beta), _ptr(out),
                                   _stream(stream)), "xrd_forward_coeffs")
    return out


def _check_coeff_inputs(g, d, spectrum, segmentation, coef, basis, maps, beta):
    npix = g.nx * g.ny
    _check(spectrum, "spectrum", torch.float32, d.n_energy * d.n_energy)
    _check(segmentation, "segmentation", torch.int32, npix)
    for t, name in ((coef, "coef"), (basis, "basis"), (maps, "maps"), (beta, "beta")):
        _check(t, name, torch.float32)
        if t.dim() != 2:
            raise ValueError(f"{name} must be 2-D")
    K, B = coef.shape[1], maps.shape[0]
    if coef.shape[0] != npix:
        raise ValueError(f"coef must be [nx*ny = {npix}, K]")
    if tuple(basis.shape) != (K, d.n_q):
        raise ValueError(f"basis must be [K = {K}, NWT = {d.n_q}]")
    if maps.shape[1] != npix:
        raise ValueError(f"maps must be [B, nx*ny = {npix}]")
    if tuple(beta.shape) != (B, d.n_energy):
        raise ValueError(f"beta must be [B = {B}, NRG = {d.n_energy}]")


def vjp_coeffs(spectrum, segmentation, coef, basis, maps, beta, grad_counts, *,
               geometry: Geometry, disc: Discretization, want_coef=True, want_maps=True,
               ctx=None, stream=None):
    """(grad_coef [nx*ny, K] float64, grad_maps [B, nx*ny] float64) for the upstream
    gradient grad_counts [rows] (xrd_vjp_coeffs)."""
    g, d = geometry, disc
    _check_coeff_inputs(g, d, spectrum, segmentation, coef, basis, maps, beta)
    _check(grad_counts, "grad_counts", torch.float32,
           g.n_views * g.n_cols * g.n_rows * d.n_energy)
    _require_gpu(spectrum, segmentation, coef, basis, maps, beta, grad_counts)
    ctx = ctx or default_context(spectrum.device.index or 0)
    dev = spectrum.device
    gc = torch.empty(coef.shape, dtype=torch.float64, device=dev) if want_coef else None
    gm = torch.empty(maps.shape, dtype=torch.float64, device=dev) if want_maps else None
    check(lib().xrd_vjp_coeffs(ctx.handle, ctypes.byref(g), ctypes.byref(d), _ptr(spectrum),
                               _ptr(segmentation), _ptr(coef), coef.shape[1], _ptr(basis),
                               maps.shape[0], _ptr(maps), _ptr(beta), _ptr(grad_counts),
                               _ptr(gc), _ptr(gm), _stream(stream)), "xrd_vjp_coeffs")
    return gc, gm


def forward_ct(spectrum, maps, beta, *, geometry: Geometry, disc: Discretization, ctx=None,
               stream=None) -> torch.Tensor:
    """CT transmission [V*COL*NRG] (view fastest) of the attenuation maps [B, nx*ny]
    with energy coefficients beta [B, NRG] (xrd_forward_ct)."""
    g, d = geometry, disc
    _check_ct_inputs(g, d, spectrum, maps, beta)
    _require_gpu(spectrum, maps, beta)
    ctx = ctx or default_context(spectrum.device.index or 0)
    out = torch.empty(g.n_views * g.n_cols * d.n_energy, dtype=torch.float32,
                      device=spectrum.device)
    check(lib().xrd_forward_ct(ctx.handle, ctypes.byref(g), ctypes.byref(d), _ptr(spectrum),
                               maps.shape[0], _ptr(maps), _ptr(beta), _ptr(out),
                               _stream(stream)), "xrd_forward_ct")
    return out


def _check_ct_inputs(g, d, spectrum, maps, beta):
    _check(spectrum, "spectrum", torch.float32, d.n_energy * d.n_energy)
    _check(maps, "maps", torch.float32)
    _check(beta, "beta", torch.float32)
    if maps.dim() != 2 or maps.shape[1] != g.nx * g.ny:
        raise ValueError(f"maps must be [B, nx*ny = {g.nx * g.ny}]")
    if tuple(beta.shape) != (maps.shape[0], d.n_energy):
        raise ValueError(f"beta must be [B = {maps.shape[0]}, NRG = {d.n_energy}]")


def vjp_ct(spectrum, maps, beta, grad_t, *, geometry: Geometry, disc: Discretization, ctx=None,
           stream=None) -> torch.Tensor:
    """grad_maps [B, nx*ny] float64 of <grad_t, forward_ct(maps)> (xrd_vjp_ct)."""
    g, d = geometry, disc
    _check_ct_inputs(g, d, spectrum, maps, beta)
    _check(grad_t, "grad_t", torch.float32, g.n_views * g.n_cols * d.n_energy)
    _require_gpu(spectrum, maps, beta, grad_t)
    ctx = ctx or default_context(spectrum.device.index or 0)
    gm = torch.empty(maps.shape, dtype=torch.float64, device=spectrum.device)
    check(lib().xrd_vjp_ct(ctx.handle, ctypes.byref(g), ctypes.byref(d), _ptr(spectrum),
                           maps.shape[0], _ptr(maps), _ptr(beta), _ptr(grad_t), _ptr(gm),
                           _stream(stream)), "xrd_vjp_ct")
    return gm


def poisson_nll(counts, y, background=None, want_grad=True, ctx=None, stream=None):
    """(loss: float, grad: float32 tensor 1 - y/f or None), f = counts + background."""
    _check(counts, "counts", torch.float32)
    _check(y, "y", torch.float32, counts.numel())
    _check(background, "background", torch.float32, counts.numel(), optional=True)
    _require_gpu(counts, y, background)
    ctx = ctx or default_context(counts.device.index or 0)
    grad = torch.empty_like(counts) if want_grad else None
    v = ctypes.c_double()
    check(lib().xrd_poisson_nll(ctx.handle, counts.numel(), _ptr(counts), _ptr(background),
                                _ptr(y), _ptr(grad), ctypes.byref(v), _stream(stream)),
          "xrd_poisson_nll")
    return v.value, grad


def coeff_problem_tensors(cp, device="cuda"):
    """Device tensors (spectrum, segmentation, coef, basis, maps, beta) of a
    synthetic.CoeffProblem."""
    p = cp.base
    t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(device=device, dtype=dt)  # noqa: E731
    return (t(p.flat("spectrum"), torch.float32), t(p.flat("segmentation"), torch.int32),
            t(cp.coef, torch.float32), t(cp.basis, torch.float32), t(cp.maps, torch.float32),
            t(cp.beta, torch.float32))


class MlemBatch:
    """MLEM of main.m:153-176 for many independent problems with one persistent launch
    per pass (xrd_mlem_batch_*).  problems: list of (A, y, c, x[, col_scale]); all A share
    one dtype; every x is updated in place."""

    def __init__(self, problems, ctx=None, stream=None):
        if not problems:
            raise ValueError("no problems")
        dts = {p[0].dtype for p in problems}
        if len(dts) != 1:
            raise TypeError("all matrices of a batch must share one dtype")
        n = len(problems)
        VP = ctypes.c_void_p * n
        A, y, c, x, sc = VP(), VP(), VP(), VP(), VP()
        rows, cols = (ctypes.c_int64 * n)(), (ctypes.c_int32 * n)()
        for k, p in enumerate(problems):
            Ak, yk, ck, xk = p[:4]
            s_k = p[4] if len(p) > 4 else None
            dt, nr, nc = _mat(Ak)
            _check_mlem_vectors(nr, nc, yk, ck, xk, s_k)
            _require_gpu(_mat_tensor(Ak), yk, ck, xk, s_k)
            A[k], y[k], c[k], x[k] = Ak.data_ptr(), yk.data_ptr(), ck.data_ptr(), xk.data_ptr()
            sc[k] = s_k.data_ptr() if s_k is not None else None
            rows[k], cols[k] = nr, nc
        self._keep = problems
        self.ctx = ctx or default_context(problems[0][0].device.index or 0)
        h = ctypes.c_void_p()
        dt = dts.pop()
        check(lib().xrd_mlem_batch_create(self.ctx.handle, n,
                                          DTYPE_BF24 if dt == BF24 else _DTYPES[dt], A, rows, cols,
                                          sc, y, c, x, 0, _stream(stream),
                                          ctypes.byref(h)), "xrd_mlem_batch_create")
        self.handle = h

    def iterate(self, n_iter: int, use_graph: bool = False, stream=None):
        check(lib().xrd_mlem_batch_iterate(self.handle, int(n_iter), int(bool(use_graph)),
                                           _stream(stream)), "xrd_mlem_batch_iterate")

    def set_launch(self, blocks_per_cu: int = 0, rows_per_wave: int = 0):
        check(lib().xrd_mlem_batch_set_launch(self.handle, blocks_per_cu, rows_per_wave),
              "xrd_mlem_batch_set_launch")

    def autotune(self, reps: int = 2, stream=None) -> dict:
        """Pick the fastest launch shape of each row-span group's streaming launch (x is
        not modified); returns the chosen groups and the pass time."""
        ms = ctypes.c_float()
        check(lib().xrd_mlem_batch_autotune(self.handle, int(reps), _stream(stream), None, None,
                                            ctypes.byref(ms)), "xrd_mlem_batch_autotune")
        return {"groups": self.groups(), "ms_pass": ms.value}

    def groups(self) -> list:
        """[{span, grid, rows_per_wave, chunks}] of the row-span groups, one launch each
        (chunks == grid: one fixed chunk per workgroup; more: guided chunks)."""
        n = 8
        sp, gr, rw, ch = ((ctypes.c_int32 * n)() for _ in range(4))
        k = lib().xrd_mlem_batch_groups(self.handle, n, sp, gr, rw, ch)
        if k < 0:
            check(k, "xrd_mlem_batch_groups")
        return [{"span": sp[i], "grid": gr[i], "rows_per_wave": rw[i], "chunks": ch[i]}
                for i in range(min(k, n))]

    def time_pass(self, reps: int = 10, stream=None) -> float:
        """Mean ms of one streaming pass over all problems, HIP events on the stream."""
        ms = ctypes.c_float()
        check(lib().xrd_mlem_batch_time_pass(self.handle, int(reps), _stream(stream),
                                             ctypes.byref(ms)), "xrd_mlem_batch_time_pass")
        return ms.value

    def close(self):
        if getattr(self, "handle", None):
            lib().xrd_mlem_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
