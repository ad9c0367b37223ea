"""ctypes mirrors of the descriptors in include/xrd_tomo.h.

``Geometry`` and ``Discretization`` carry what the reference hard-codes as
literals (Forward_XRD.cu:17-53, integrate_diffraction.cu:81-142,
wvt_value.cu:8-23).  ``Geometry.reference`` / ``Discretization.reference``
reproduce those literals.
"""
from __future__ import annotations

import ctypes
import math

QUIRK_STALE_COMPTON_VOXEL = 0x1
QUIRK_SWAPPED_LEGS = 0x2
QUIRKS_REFERENCE = QUIRK_STALE_COMPTON_VOXEL | QUIRK_SWAPPED_LEGS


class Geometry(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
        ("n_views", ctypes.c_int32), ("n_cols", ctypes.c_int32),
        ("n_rows", ctypes.c_int32), ("reserved0", ctypes.c_int32),
        ("center_x", ctypes.c_double), ("center_y", ctypes.c_double),
        ("r_src", ctypes.c_double), ("r_det", ctypes.c_double),
        ("view_angle0", ctypes.c_double), ("view_angle_step", ctypes.c_double),
        ("col_pitch", ctypes.c_double), ("row_pitch", ctypes.c_double),
        ("det_z0", ctypes.c_double), ("target_angle", ctypes.c_double),
        ("focal_spot", ctypes.c_double), ("wedge_angle", ctypes.c_double),
        ("voxel_size", ctypes.c_double),
    ]

    @classmethod
    def reference(cls, nx: int, ny: int, n_views: int = 32, n_cols: int = 1024,
                  n_rows: int = 1) -> "Geometry":
        """The scanner literals of integrate_diffraction.cu:81-128 (mm_over_vox = NX/100)."""
        mmv = float(nx) / 100.0
        return cls(nx=nx, ny=ny, n_views=n_views, n_cols=n_cols, n_rows=n_rows,
                   center_x=100.0 * mmv, center_y=100.0 * mmv,
                   r_src=150.0 * mmv, r_det=170.0 * mmv,
                   view_angle0=0.0, view_angle_step=2.0 * math.pi / n_views,
                   col_pitch=0.5 * mmv, row_pitch=1.0 * mmv, det_z0=10.0 * mmv,
                   target_angle=math.pi / 6.0, focal_spot=0.5 * mmv,
                   wedge_angle=0.5 / 180.0 * math.pi, voxel_size=1.0)

    @property
    def n_det(self) -> int:
        return self.n_cols * self.n_rows

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class Discretization(ctypes.Structure):
    _fields_ = [
        ("n_energy", ctypes.c_int32), ("energy_halfwidth", ctypes.c_int32),
        ("n_q", ctypes.c_int32), ("n_qf", ctypes.c_int32), ("n_mat", ctypes.c_int32),
        ("lut_nq", ctypes.c_int32), ("lut_na", ctypes.c_int32), ("flags", ctypes.c_uint32),
        ("e_min", ctypes.c_double), ("e_width", ctypes.c_double),
        ("q_min", ctypes.c_double), ("q_max", ctypes.c_double), ("q_dq0", ctypes.c_double),
        ("ray_step", ctypes.c_double), ("hbarc", ctypes.c_double), ("mc2", ctypes.c_double),
        ("kn0", ctypes.c_double), ("kn1", ctypes.c_double), ("kn2", ctypes.c_double),
    ]

    @classmethod
    def reference(cls, n_energy: int = 64, n_q: int = 256, n_qf: int = 768,
                  n_mat: int = 3, flags: int = 0) -> "Discretization":
        """Literals of Forward_XRD.cu:21-53 and integrate_diffraction.cu:138-140."""
        return cls(n_energy=n_energy, energy_halfwidth=5, n_q=n_q, n_qf=n_qf, n_mat=n_mat,
                   lut_nq=64, lut_na=64, flags=flags,
                   e_min=8.0, e_width=72.0 / n_energy,
                   q_min=0.5, q_max=6.0, q_dq0=0.01, ray_step=0.25,
                   hbarc=1.973269804, mc2=510.99895,
                   kn0=0.751691286166726, kn1=1.4599683207703, kn2=-0.64597532690936)

    @property
    def n_cols_matrix(self) -> int:
        return self.n_q * self.n_mat

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class Bf16TablesStruct(ctypes.Structure):
    """ctypes mirror of xrd_bf16_tables (device pointers)."""
    _fields_ = [("ifine", ctypes.c_void_p), ("ifine_scale", ctypes.c_void_p),
                ("a1", ctypes.c_void_p), ("a2", ctypes.c_void_p), ("mu_scale", ctypes.c_void_p)]


def num_rows(g: Geometry, d: Discretization) -> int:
    """Rows of the model matrix: SRC*COL*ROW*NRG (Forward_XRD.cu:143)."""
    return g.n_views * g.n_cols * g.n_rows * d.n_energy


def q_axis(d: Discretization):
    """Coarse parabolic axis centres/widths (wvt_value.cu:8-23) and the fine uniform
    axis centres (integrate_diffraction.cu:302), as float64 lists."""
    nw = float(d.n_q)
    c = (d.q_max - d.q_min - d.q_dq0 * nw) / (nw * nw)
    edges = [d.q_min + n * d.q_dq0 + c * n * n for n in range(d.n_q + 1)]
    q = [0.5 * (edges[n] + edges[n + 1]) for n in range(d.n_q)]
    dq = [edges[n + 1] - edges[n] for n in range(d.n_q)]
    dqf = (d.q_max - d.q_min) / d.n_qf
    qf = [d.q_min + dqf * (0.5 + n) for n in range(d.n_qf)]
    return q, dq, qf, dqf
