"""Seeded synthetic inputs for the forward projector and MLEM.

The reference builds its inputs in MATLAB from measured data
(spectrum_builder.m, XSection_builder.m, phantom_generator.m, XRD_builder.m).
Those builders are replaced here by seeded generators with the same output
semantics, so every workload is reproducible without MATLAB or HDF5:

* attenuation: the reference's two-basis model mu(E) = a1/e^3 + a2/f2(e),
  e = E/mc^2, f2 = 1/KN quadratic (integrate_diffraction.cu:278,291-296), with
  a1 = rho*a, a2 = rho*b per material (phantom_generator.m:31-36 uses the same
  split), in 1/voxel as the kernel expects (main.m:60-61);
* spectrum: Kramers (80 - E)/E behind 2.5 mm Al, smeared with the detector
  response sigma(E) = 0.5*(1.61 + 0.025 E) keV into the NRG x NRG matrix
  Ndet(e_src, e_det) that Forward_XRD takes as its first input
  (spectrum_builder.m:65-95);
* XRD form factors: 3-6 Gaussian Bragg peaks plus an amorphous hump per
  material on the fine q axis (XRD_builder.m builds the same object from
  measured patterns);
* Compton table: a smooth incoherent-scattering surface on the
  (log2 q, log2 a1/a2) grid, values 0..255 (XSection_builder.m:93-128).

Geometry of the scaled workloads: fan beam, source-isocentre 600 mm,
source-detector 1000 mm, isocentre at the centre of the voxel grid, one
detector row at a height and column pitch chosen so that isocentre rays
scatter between 2 and 8 degrees.
"""
from __future__ import annotations

import dataclasses
import math
import os
from typing import Optional

import numpy as np

from .descriptors import Discretization, Geometry, q_axis

_DATA = os.path.join(os.path.dirname(__file__), "data")

MC2 = 510.99895
RHO_AL = 2.699


@dataclasses.dataclass
class Problem:
    """Everything xrd_forward takes, as host numpy arrays (float32/int32)."""
    name: str
    geometry: Geometry
    disc: Discretization
    spectrum: np.ndarray      # [E x E] float32, spectrum[e_src + E*e_det]
    ifine: np.ndarray         # [QF x M] float32 (q fastest), flattened order
    segmentation: np.ndarray  # [NX x NY] int32, x fastest
    a1: np.ndarray            # [NX x NY] float32, 1/voxel
    a2: np.ndarray            # [NX x NY] float32, 1/voxel
    s2_table: np.ndarray      # [NQ x NA] float32, 0..255
    lut: tuple                # (q_min, q_max, a_min, a_max)
    x_true: np.ndarray        # [Q x M] float64 form factor sampled at coarse centres
    photons_per_view: float   # 0: the spectrum is in absolute photons (no rescaling)
    mlem_iterations: int
    voxel_mm: float
    compton_scale: float = 0.0  # C output -> photons (main.m:68); 0 when not calibrated

    @property
    def n_rows(self) -> int:
        g, d = self.geometry, self.disc
        return g.n_views * g.n_cols * g.n_rows * d.n_energy

    @property
    def n_cols(self) -> int:
        return self.disc.n_q * self.disc.n_mat

    def flat(self, name: str) -> np.ndarray:
        """Array in the ABI's memory order (first index fastest), 1-D."""
        a = getattr(self, name)
        return np.ascontiguousarray(a.reshape(-1, order="F") if a.ndim > 1 else a)


# ---------------------------------------------------------------------------
# physics helpers

def f2_kn(e, d: Discretization):
    """1/KN quadratic of Forward_XRD.cu:51-53 at reduced energy e = E/mc^2."""
    return d.kn0 + d.kn1 * e + d.kn2 * e * e


def mu_total(a1, a2, energy_kev, d: Discretization):
    """mu(E) = a1/e^3 + a2/f2(e) (integrate_diffraction.cu:291-296)."""
    e = energy_kev / d.mc2
    return a1 / e ** 3 + a2 / f2_kn(e, d)


def al_mass_attenuation(energy_kev):
    """Log-log interpolation of the Al mu/rho table (cm^2/g)."""
    tab = np.loadtxt(os.path.join(_DATA, "al_mass_attenuation.csv"), delimiter=",",
                     comments="#")
    e_mev = np.asarray(energy_kev, dtype=np.float64) / 1000.0
    return np.exp(np.interp(np.log(e_mev), np.log(tab[:, 0]), np.log(tab[:, 1])))


def _norm_cdf(x):
    from math import erf
    return 0.5 * (1.0 + np.vectorize(erf)(x / math.sqrt(2.0)))


def kramers_response_matrix(d: Discretization, kvp: float = 80.0, al_mm: float = 2.5):
    """Ndet(e_src, e_det) [E x E], src fastest (spectrum_builder.m:65-95 semantics):
    photons emitted in source bin e_src times the probability that the detector
    response sigma(E) = 0.5*(1.61 + 0.025 E) puts them in detector bin e_det."""
    E = d.n_energy
    edges = d.e_min + d.e_width * np.arange(E + 1)
    centres = 0.5 * (edges[:-1] + edges[1:])
    shape = np.clip(kvp - centres, 0.0, None) / centres
    shape *= np.exp(-al_mass_attenuation(centres) * RHO_AL * al_mm / 10.0)
    shape *= d.e_width
    sigma = 0.5 * (1.61 + 0.025 * centres)
    # m[s, t]: probability that a photon of source bin s lands in detector bin t
    z2 = (edges[None, 1:] - centres[:, None]) / sigma[:, None]
    z1 = (edges[None, :-1] - centres[:, None]) / sigma[:, None]
    prob = _norm_cdf(z2) - _norm_cdf(z1)
    return (shape[:, None] * prob).astype(np.float64)  # [s, t]


def form_factor(q, rng, n_peaks_range=(3, 6)):
    """Gaussian Bragg peaks + amorphous hump, evaluated at q (1/Angstrom)."""
    npk = int(rng.integers(n_peaks_range[0], n_peaks_range[1] + 1))
    centres = rng.uniform(0.6, 3.4, npk)
    widths = rng.uniform(0.02, 0.05, npk)
    heights = rng.lognormal(0.0, 1.0, npk)
    hc, hw = rng.uniform(1.0, 2.0), rng.uniform(0.3, 0.6)
    hh = rng.lognormal(0.0, 0.5)
    q = np.asarray(q, dtype=np.float64)
    f = hh * np.exp(-0.5 * ((q - hc) / hw) ** 2)
    for c, w, h in zip(centres, widths, heights):
        f = f + h * np.exp(-0.5 * ((q - c) / w) ** 2)
    return f, dict(centres=centres, widths=widths, heights=heights, hump=(hc, hw, hh))


def compton_table(nq: int, na: int, lut):
    """Smooth incoherent-scattering surface S2(log2 q, log2 a1/a2) in 0..255.
    S rises from 0 to 1 with q (scaled by an a-dependent onset), times 255."""
    qmin, qmax, amin, amax = lut
    lq = qmin + (np.arange(nq) + 0.5) * (qmax - qmin) / nq
    la = amin + (np.arange(na) + 0.5) * (amax - amin) / na
    Lq, La = np.meshgrid(lq, la, indexing="ij")  # [nq, na]
    onset = -8.0 + 0.25 * (La - 0.5 * (amin + amax))
    s = 1.0 / (1.0 + np.exp(-(Lq - onset) * 1.3))
    return (255.0 * s).astype(np.float64)  # [nq, na]


# ---------------------------------------------------------------------------
# geometry presets

def fan_geometry(nx: int, ny: int, voxel_mm: float, n_views: int, n_cols: int,
                 sid_mm: float = 600.0, sdd_mm: float = 1000.0,
                 theta_min_deg: float = 2.0, theta_max_deg: float = 8.0,
                 n_rows: int = 1) -> Geometry:
    """Fan-beam scanner with the XRD panel opposite the source (the reference's
    arrangement, integrate_diffraction.cu:94-121) at source-isocentre sid_mm and
    source-detector sdd_mm; one row at height z0 and COL columns spanning the
    lateral offsets for which isocentre rays scatter in [theta_min, theta_max];
    further rows (n_rows > 1) stack upwards at a 1 mm pitch, as the reference places
    its rows (integrate_diffraction.cu:100-109)."""
    iso_det = sdd_mm - sid_mm
    z0 = iso_det * math.tan(math.radians(theta_min_deg))
    rho_max = iso_det * math.tan(math.radians(theta_max_deg))
    r_max = math.sqrt(rho_max ** 2 - z0 ** 2)
    col_pitch = 2.0 * r_max / n_cols
    v = voxel_mm
    return Geometry(nx=nx, ny=ny, n_views=n_views, n_cols=n_cols, n_rows=n_rows,
                    center_x=nx / 2.0, center_y=ny / 2.0,
                    r_src=sid_mm / v, r_det=iso_det / v,
                    view_angle0=0.0, view_angle_step=2.0 * math.pi / n_views,
                    col_pitch=col_pitch / v, row_pitch=1.0 / v, det_z0=z0 / v,
                    target_angle=math.pi / 6.0, focal_spot=0.5 / v,
                    wedge_angle=math.radians(0.5), voxel_size=1.0)


def scaled_discretization(n_mat: int, n_energy: int = 61, e_lo: float = 20.0,
                          e_hi: float = 80.0, n_q: int = 128, n_qf: int = 384,
                          q_min: float = 0.5, q_max: float = 3.5,
                          flags: int = 0) -> Discretization:
    """Energy bins centred on e_lo..e_hi (1 keV for 61 bins), parabolic q axis."""
    d = Discretization.reference(n_energy=n_energy, n_q=n_q, n_qf=n_qf, n_mat=n_mat,
                                 flags=flags)
    d.e_width = (e_hi - e_lo) / (n_energy - 1)
    d.e_min = e_lo - 0.5 * d.e_width
    d.q_min, d.q_max = q_min, q_max
    return d


# ---------------------------------------------------------------------------
# phantoms

def _paint_phantom(nx, ny, voxel_mm, n_mat, rng, disc_radius_mm, axis_range_mm,
                   allow_rect):
    seg = np.zeros((nx, ny), dtype=np.int32)  # seg[ix, iy]
    x = (np.arange(nx) + 0.5 - nx / 2.0) * voxel_mm
    y = (np.arange(ny) + 0.5 - ny / 2.0) * voxel_mm
    X, Y = np.meshgrid(x, y, indexing="ij")
    for m in range(1, n_mat + 1):
        a = rng.uniform(*axis_range_mm)
        b = rng.uniform(*axis_range_mm)
        r = rng.uniform(0.0, max(disc_radius_mm - max(a, b), 0.0))
        t = rng.uniform(0.0, 2.0 * math.pi)
        cx, cy = r * math.cos(t), r * math.sin(t)
        ang = rng.uniform(0.0, math.pi)
        ca, sa = math.cos(ang), math.sin(ang)
        u = (X - cx) * ca + (Y - cy) * sa
        w = -(X - cx) * sa + (Y - cy) * ca
        if allow_rect and rng.uniform() < 0.5:
            mask = (np.abs(u) <= a) & (np.abs(w) <= b)
        else:
            mask = (u / a) ** 2 + (w / b) ** 2 <= 1.0
        mask &= X ** 2 + Y ** 2 <= disc_radius_mm ** 2
        seg[mask] = m  # later materials win (painted in order)
    return seg


def _materials(n_mat, rng):
    rho = rng.uniform(0.9, 2.7, n_mat)
    a = np.exp(rng.uniform(math.log(2e-5), math.log(2e-4), n_mat))
    b = rng.uniform(0.13, 0.17, n_mat)
    return rho, a, b


def make_problem(name: str = "small", seed: Optional[int] = None, **overrides) -> Problem:
    """Build one of the named workloads.

    ``fan64``  : 64x64 voxels at 4 mm, 4 ellipse materials, 60 views, 64 pixels,
                 61 energies, 128 q-bins; 1e4 photons/view, 20 MLEM iterations.
    ``fan256`` : 256x256 voxels at 2 mm, 8 materials in ellipses and rectangles,
                 360 views, 256 pixels, 61 energies, 128 q-bins; 1e5 photons/view,
                 50 MLEM iterations.
    ``small``  : 24x24 voxels, 4 views, 6 pixels, 16 energies, 24 q-bins, 2 materials
                 (tests).
    ``reference`` : the reference's own scanner literals (integrate_diffraction.cu:81-140)
                 on a 40x40 grid with 3 materials, 4 views, 16 pixels, 16 energies.
    ``paper100``: the problem main.m builds from the reference's own inputs (100x100
                 phantom, 3 measured materials, 32 views, 1024 pixels, 64 energies,
                 256 q-bins, 768 fine q-bins), stored in data/paper100 by
                 ``python -m xrd_tomography_amd.builders``; n_views / n_cols overrides
                 keep the inputs and change the scanner's view and column counts.
    Keyword overrides replace sizes (nx, n_views, n_cols, n_rows, n_energy, n_q, n_qf,
    n_mat, flags, voxel_mm, photons_per_view, mlem_iterations).
    """
    presets = {
        "fan64": dict(nx=64, voxel_mm=4.0, n_mat=4, n_views=60, n_cols=64, n_energy=61,
                      n_q=128, n_qf=384, photons_per_view=1e4, mlem_iterations=20,
                      allow_rect=False, axis=(12.0, 40.0), disc=110.0, seed=0),
        "fan256": dict(nx=256, voxel_mm=2.0, n_mat=8, n_views=360, n_cols=256, n_energy=61,
                       n_q=128, n_qf=384, photons_per_view=1e5, mlem_iterations=50,
                       allow_rect=True, axis=(20.0, 80.0), disc=240.0, seed=1),
        "small": dict(nx=24, voxel_mm=8.0, n_mat=2, n_views=4, n_cols=6, n_energy=16,
                      n_q=24, n_qf=72, photons_per_view=1e4, mlem_iterations=10,
                      allow_rect=True, axis=(20.0, 60.0), disc=90.0, seed=2),
        "reference": dict(nx=40, voxel_mm=2.0, n_mat=3, n_views=4, n_cols=16, n_energy=16,
                          n_q=32, n_qf=96, photons_per_view=1e4, mlem_iterations=10,
                          allow_rect=False, axis=(8.0, 20.0), disc=35.0, seed=3),
    }
    if name == "paper100":
        return _paper100(**overrides)
    if name not in presets:
        raise ValueError(f"unknown workload {name!r}; choose from {sorted(presets)}")
    p = dict(presets[name])
    p.update({k: v for k, v in overrides.items() if k not in ("flags",)})
    flags = int(overrides.get("flags", 0))
    rng = np.random.Generator(np.random.PCG64(p["seed"] if seed is None else seed))
    nx = ny = int(p["nx"])
    vmm = float(p["voxel_mm"])
    M = int(p["n_mat"])

    if name == "reference":
        g = Geometry.reference(nx, ny, n_views=int(p["n_views"]), n_cols=int(p["n_cols"]),
                               n_rows=int(p.get("n_rows", 1)))
        d = Discretization.reference(n_energy=int(p["n_energy"]), n_q=int(p["n_q"]),
                                     n_qf=int(p["n_qf"]), n_mat=M, flags=flags)
        # the reference phantom occupies [0,NX]^2 while its rotation centre is at
        # (NX,NY) (integrate_diffraction.cu:88-95); keep inclusions inside the grid.
    else:
        g = fan_geometry(nx, ny, vmm, int(p["n_views"]), int(p["n_cols"]),
                         n_rows=int(p.get("n_rows", 1)))
        d = scaled_discretization(M, n_energy=int(p["n_energy"]), n_q=int(p["n_q"]),
                                  n_qf=int(p["n_qf"]), flags=flags)

    seg = _paint_phantom(nx, ny, vmm, M, rng, float(p["disc"]), p["axis"], p["allow_rect"])
    rho, a, b = _materials(M, rng)
    vox_cm = vmm / 10.0
    a1 = np.zeros((nx, ny)); a2 = np.zeros((nx, ny))
    for m in range(M):
        a1[seg == m + 1] = rho[m] * a[m] * vox_cm
        a2[seg == m + 1] = rho[m] * b[m] * vox_cm

    q, dq, qf, dqf = q_axis(d)
    ifine = np.zeros((d.n_qf, M))
    x_true = np.zeros((d.n_q, M))
    for m in range(M):
        f_fine, _ = form_factor(qf, rng)
        # reference magnitudes: 1e-3 .. 1 per cm (XRD_builder.m:52,101,152)
        scale = math.exp(rng.uniform(math.log(1e-2), math.log(1.0)))
        ifine[:, m] = scale * f_fine / f_fine.max()
        x_true[:, m] = np.interp(q, qf, ifine[:, m])

    spec = kramers_response_matrix(d)
    lut = (-12.0, -4.0, -14.0, -8.0)
    s2 = compton_table(d.lut_nq, d.lut_na, lut)

    return Problem(name=name, geometry=g, disc=d,
                   spectrum=spec.astype(np.float32), ifine=ifine.astype(np.float32),
                   segmentation=seg, a1=a1.astype(np.float32), a2=a2.astype(np.float32),
                   s2_table=s2.astype(np.float32), lut=lut, x_true=x_true,
                   photons_per_view=float(p["photons_per_view"]),
                   mlem_iterations=int(p["mlem_iterations"]), voxel_mm=vmm)


def _paper100(**overrides) -> Problem:
    from .problem_io import load_problem
    prob, _ = load_problem(os.path.join(_DATA, "paper100"))
    extra = set(overrides) - {"n_views", "n_cols"}
    if extra:
        raise ValueError(f"paper100 takes only n_views / n_cols overrides, not {sorted(extra)}")
    if overrides:
        g = prob.geometry
        prob.geometry = Geometry.reference(g.nx, g.ny,
                                           n_views=int(overrides.get("n_views", g.n_views)),
                                           n_cols=int(overrides.get("n_cols", g.n_cols)))
    return prob


def bf16_round(x: np.ndarray) -> np.ndarray:
    """float32 values rounded to the nearest bfloat16 (ties to even), as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(np.shape(x))


def bf16_bits(x: np.ndarray) -> np.ndarray:
    """bfloat16 bits (uint16) of float32 values, rounded to nearest even."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16).reshape(np.shape(x))


def bf16_float(b: np.ndarray) -> np.ndarray:
    """float32 values of bfloat16 bits (exact)."""
    return (np.asarray(b, np.uint16).astype(np.uint32) << 16).view(np.float32)


@dataclasses.dataclass
class Bf16Tables:
    """A problem's form factors and attenuation maps in BF16 storage with per-material
    FP32 scales: the fields of xrd_bf16_tables, as host arrays in the ABI's order."""
    ifine: np.ndarray        # uint16 [QF*M], q fastest
    ifine_scale: np.ndarray  # float32 [M]
    a1: np.ndarray           # uint16 [nx*ny(*n_slices)], x fastest
    a2: np.ndarray
    mu_scale: np.ndarray     # float32 [2*(M+1)]: a1 scales by segmentation id, then a2's


def _ids(prob: Problem) -> np.ndarray:
    """Segmentation id of every voxel in ABI order, ids outside [0, M] mapped to 0 (air)."""
    seg = prob.flat("segmentation").astype(np.int64)
    seg[(seg < 0) | (seg > prob.disc.n_mat)] = 0
    return seg


def encode_bf16_tables(prob: Problem) -> Bf16Tables:
    """BF16 storage of prob's form-factor table and attenuation maps: each form-factor
    column m and each map's voxels of segmentation id s are divided by a scale (the
    column's or the group's largest magnitude, 1 if that is 0) and rounded to
    bfloat16.  Round-to-nearest keeps 8 significant bits, so decoding moves every
    entry by at most 2^-8 of itself; a material of constant (a1, a2) decodes exactly."""
    M = prob.disc.n_mat
    ifine = np.asarray(prob.flat("ifine"), np.float32).reshape(M, -1)  # [M, QF]
    isc = np.array([np.abs(col).max() for col in ifine], np.float32)
    isc[~(isc > 0)] = 1.0
    ifine_b = bf16_bits(ifine / isc[:, None]).reshape(-1)
    ids = _ids(prob)
    maps, scales = [], []
    for name in ("a1", "a2"):
        a = np.asarray(prob.flat(name), np.float32)
        sc = np.ones(M + 1, np.float32)
        for i in range(M + 1):
            sel = ids == i
            if sel.any():
                mx = np.float32(np.abs(a[sel]).max())
                if mx > 0:
                    sc[i] = mx
        maps.append(bf16_bits(a / sc[ids]))
        scales.append(sc)
    return Bf16Tables(ifine=ifine_b, ifine_scale=isc, a1=maps[0], a2=maps[1],
                      mu_scale=np.concatenate(scales).astype(np.float32))


def decode_bf16_tables(prob: Problem, t: Bf16Tables) -> Problem:
    """A copy of prob whose FP32 tables hold exactly the values the kernels decode from
    t (bf16 times scale, an FP32 product): the FP32 problem that xrd_forward_bf16 on
    t computes, for the oracle."""
    M = prob.disc.n_mat
    ifine = (bf16_float(t.ifine).reshape(M, -1) * t.ifine_scale[:, None]).astype(np.float32)
    ids = _ids(prob)
    shape = prob.segmentation.shape
    a1 = (bf16_float(t.a1) * t.mu_scale[:M + 1][ids]).astype(np.float32)
    a2 = (bf16_float(t.a2) * t.mu_scale[M + 1:][ids]).astype(np.float32)
    unflat = lambda v: v.reshape(shape, order="F")  # noqa: E731
    return dataclasses.replace(prob, ifine=np.ascontiguousarray(ifine.T),
                               a1=unflat(a1), a2=unflat(a2))


def decoded_bf16_tables(prob: Problem) -> Problem:
    """decode_bf16_tables(prob, encode_bf16_tables(prob))."""
    return decode_bf16_tables(prob, encode_bf16_tables(prob))


def make_volume_problem(name: str = "small", n_slices: int = 4, seed: Optional[int] = None,
                        **overrides) -> Problem:
    """A volume of n_slices slices (one voxel apart) for xrd_forward_slice: the
    geometry, discretisation, spectrum, form factors and Compton table of workload
    `name`; every slice is painted with its own seeded phantom, and the materials'
    attenuation is one table for the whole volume.  segmentation, a1 and a2 are
    [nx, ny, n_slices]."""
    base = make_problem(name, seed=seed, **overrides)
    M = base.disc.n_mat
    s0 = 0 if seed is None else int(seed)
    rng = np.random.Generator(np.random.PCG64(1_000_003 + s0))
    rho, a, b = _materials(M, rng)
    vox_cm = base.voxel_mm / 10.0
    a1m, a2m = rho * a * vox_cm, rho * b * vox_cm
    nx, ny = base.geometry.nx, base.geometry.ny
    seg = np.zeros((nx, ny, n_slices), np.int32)
    a1 = np.zeros((nx, ny, n_slices))
    a2 = np.zeros((nx, ny, n_slices))
    for k in range(n_slices):
        sk = make_problem(name, seed=s0 * 1000 + 101 + k, **overrides).segmentation
        seg[:, :, k] = sk
        for m in range(M):
            a1[:, :, k][sk == m + 1] = a1m[m]
            a2[:, :, k][sk == m + 1] = a2m[m]
    return dataclasses.replace(base, name=f"{name}-volume{n_slices}", segmentation=seg,
                               a1=a1.astype(np.float32), a2=a2.astype(np.float32))


def poisson_counts(expected: np.ndarray, seed: int) -> np.ndarray:
    """Poisson noise (the poissrnd of main.m:141), seeded."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.poisson(np.clip(expected, 0.0, None)).astype(np.float32)


@dataclasses.dataclass
class CoeffProblem:
    """Inputs of the differentiable coefficient model (per-voxel form factors on a
    Gaussian-bump basis of the coarse q axis, attenuation from B basis maps)."""
    base: Problem
    coef: np.ndarray   # [NX*NY, K] float32 (k fastest), zero on air voxels
    basis: np.ndarray  # [K, NWT] float32
    maps: np.ndarray   # [B, NX*NY] float32, 1/voxel
    beta: np.ndarray   # [B, NRG] float32

    @property
    def K(self) -> int:
        return self.coef.shape[1]

    @property
    def B(self) -> int:
        return self.maps.shape[0]


def gaussian_bump_basis(d: Discretization, K: int, rng) -> np.ndarray:
    """K seeded Gaussian bumps over the coarse q axis [K, NWT]."""
    q, dq, _, _ = q_axis(d)
    q = np.asarray(q)
    centres = np.linspace(d.q_min, d.q_max, K) + rng.uniform(-0.2, 0.2, K) * (d.q_max - d.q_min) / K
    width = (d.q_max - d.q_min) / K * rng.uniform(0.6, 1.2, K)
    return np.exp(-0.5 * ((q[None, :] - centres[:, None]) / width[:, None]) ** 2)


def make_coeff_problem(prob: Problem, K: int = 16, full_mu: bool = True,
                       seed: int = 21) -> CoeffProblem:
    """Coefficient model on top of `prob`: per-voxel coefficients = material profile
    + 20% seeded noise; attenuation either as the full per-energy map (B = NRG,
    beta = identity) or the reference's two-basis form (B = 2)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    d = prob.disc
    basis = gaussian_bump_basis(d, K, rng)
    npix = prob.geometry.nx * prob.geometry.ny
    seg = prob.flat("segmentation")
    mat_coef = rng.lognormal(0.0, 1.0, (d.n_mat, K)) * 0.05
    coef = np.zeros((npix, K))
    live = seg > 0
    coef[live] = mat_coef[seg[live] - 1] * rng.uniform(0.8, 1.2, (int(live.sum()), K))
    E = d.n_energy
    centres = d.e_min + d.e_width * (np.arange(E) + 0.5)
    e = centres / d.mc2
    b1, b2 = 1.0 / e ** 3, 1.0 / f2_kn(e, d)
    a1, a2 = prob.flat("a1").astype(np.float64), prob.flat("a2").astype(np.float64)
    if full_mu:
        maps = b1[:, None] * a1[None, :] + b2[:, None] * a2[None, :]
        beta = np.eye(E)
    else:
        maps = np.stack([a1, a2])
        beta = np.stack([b1, b2])
    return CoeffProblem(base=prob, coef=coef.astype(np.float32), basis=basis.astype(np.float32),
                        maps=maps.astype(np.float32), beta=beta.astype(np.float32))


def make_batch_problems(n: int = 48, seed: int = 4, n_cols: int = 64):
    """The batched workload: n phantoms with sizes drawn uniformly from
    {64, 96, 128, 192, 256}^2 voxels over a 512 mm field of view, view counts uniform in
    90..360, 2..10 materials each, 61 energies, 128 q-bins, n_cols detector pixels.
    Every phantom has its own seed (seed*1000 + k) for its materials and shapes."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = rng.choice([64, 96, 128, 192, 256], n)
    views = rng.integers(90, 361, n)
    mats = rng.integers(2, 11, n)
    out = []
    for k in range(n):
        nx = int(sizes[k])
        out.append(make_problem("fan256", seed=seed * 1000 + k, nx=nx, voxel_mm=512.0 / nx,
                                n_views=int(views[k]), n_cols=n_cols, n_mat=int(mats[k]),
                                photons_per_view=1e5 * n_cols / 256.0, mlem_iterations=30))
    return out
