"""Input builders restated from the reference's MATLAB builders.

Host-side, run once per scanner/material set; their outputs are inputs of
xrd_forward (the spectrum response matrix, the Compton S2 table and its grid
edges, the form factors, the phantom).  They read the reference's input files:
the physics tables (MATLAB v5: attenuation_data.mat, Hubbell_data.mat,
physical_constants.mat), the SpekCalc source spectrum and the measured
diffraction patterns (MATLAB v7.3, read by xrd_tomography_amd.matfile).

* ``fit_basis``: photoelectric + Compton basis fit of mu/rho per element
  (XSection_builder.m:15-60, least squares without intercept).
* ``xsection_tables``: coherent and incoherent cross-sections over the
  (log2 q, log2 a1/a2) grid, scaled to 0..255 (XSection_builder.m:62-128).
* ``spectrum_response``: the detector-response-smeared spectrum matrix
  Ndet[e_src, e_det] (spectrum_builder.m:8-95); ``load_spectrum_hd`` reads its
  0.05 keV input, Spectrum_80kVp/spectral_data_HD.mat (MAT v7.3).
* ``pattern_q`` and ``xrd_builder``: measured 2-theta diffraction patterns
  (open_source_data/*.mat, MAT v7.3) converted to calibrated form factors on the
  fine q axis, with the Rayleigh + Compton theory outside the measured range
  (XRD_builder.m).
* ``phantom_generator``: the paper's 100 x 100 phantom of cellulose, aluminium
  alloy and ammonium nitrate (phantom_generator.m).
* ``paper_problem``: the problem main.m:12-63 assembles from all of the above.
"""
from __future__ import annotations

import os
from typing import Optional, Sequence

import numpy as np

MEV_TO_KEV = 1000.0


def reference_dir() -> Optional[str]:
    """Directory holding the reference's Physics_inputs/ and Spectrum_80kVp/
    (environment variable XRD_REFERENCE_DIR), or None when absent."""
    d = os.environ.get("XRD_REFERENCE_DIR", "")
    return d if d and os.path.isdir(os.path.join(d, "Physics_inputs")) else None


def load_physics(physics_dir: str) -> dict:
    """The tables XSection_builder.m:22-24 loads: per-element mu/rho vs energy
    (attenuation_coef{Z}: MeV, cm^2/g), Hubbell 1995 coherent form factors and
    incoherent scattering functions on the wvt_list momentum-transfer axis, and
    the physical constants (r_e cm, N_A, M g/mol, hbarc keV*Angstrom, mc2 keV)."""
    from scipy.io import loadmat
    att = loadmat(os.path.join(physics_dir, "attenuation_data.mat"))["attenuation_coef"]
    hub = loadmat(os.path.join(physics_dir, "Hubbell_data.mat"))
    ph = loadmat(os.path.join(physics_dir, "physical_constants.mat"))["phys"][0, 0]
    scalar = lambda k: float(np.asarray(ph[k], dtype=np.float64).ravel()[0])  # noqa: E731
    return {
        "attenuation": [np.asarray(att[0, z], dtype=np.float64) for z in range(att.shape[1])],
        "wvt_list": np.asarray(hub["wvt_list"], dtype=np.float64).ravel(),
        "Fcoh": np.asarray(hub["Fcoh"], dtype=np.float64),
        "Sincoh": np.asarray(hub["Sincoh"], dtype=np.float64),
        "r_e": scalar("r_e"), "N_A": scalar("N_A"), "hbarc": scalar("hbarc"),
        "mc2": scalar("mc2"), "M": np.asarray(ph["M"], dtype=np.float64).ravel(),
    }


def basis_f1(e):
    """Photoelectric basis e^-3 at reduced energy e = E/mc^2 (XSection_builder.m:31)."""
    return np.asarray(e, dtype=np.float64) ** -3.0


def basis_f2(e):
    """Klein-Nishina total cross-section basis (XSection_builder.m:32)."""
    e = np.asarray(e, dtype=np.float64)
    l2 = np.log1p(2.0 * e)
    return ((1.0 + e) / e ** 2 * (2.0 * (1.0 + e) / (1.0 + 2.0 * e) - l2 / e)
            + l2 / (2.0 * e) - (1.0 + 3.0 * e) / (1.0 + 2.0 * e) ** 2)


def fit_basis(attenuation: Sequence[np.ndarray], z_range=range(2, 28), mc2_mev: float = 0.5110):
    """(a1/rho, a2/rho) in cm^2/g per atomic number in z_range: the least-squares
    fit mu/rho(E) = a1 f1(E/mc2) + a2 f2(E/mc2) without intercept
    (XSection_builder.m:35-60; attenuation[Z-1] is [energy MeV, mu/rho])."""
    a1, a2 = [], []
    for z in z_range:
        ac = attenuation[z - 1]
        e = ac[:, 0] / mc2_mev
        X = np.stack([basis_f1(e), basis_f2(e)], axis=1)
        coef, *_ = np.linalg.lstsq(X, ac[:, 1], rcond=None)
        a1.append(coef[0])
        a2.append(coef[1])
    return np.array(a1), np.array(a2)


def makima2(x: np.ndarray, y: np.ndarray, v: np.ndarray, xq: np.ndarray, yq: np.ndarray):
    """Tensor-product modified-Akima interpolation of v[x, y] at the grid xq x yq
    (the 'makima' method of XSection_builder.m:118-119): 1-D makima along x for
    every y node, then along y."""
    from scipy.interpolate import Akima1DInterpolator
    # the regular grids end on the data's end points up to rounding: clamp onto them
    xq = np.clip(xq, x[0], x[-1])
    yq = np.clip(yq, y[0], y[-1])
    t = Akima1DInterpolator(x, v, axis=0, method="makima")(xq)          # [nxq, ny]
    return Akima1DInterpolator(y, t, axis=1, method="makima")(yq)       # [nxq, nyq]


def xsection_tables(phys: dict, nq_bins: int = 64, na_bins: int = 64,
                    q_range=range(2, 39), z_range=range(2, 28)):
    """(S1_grid, S2_grid, grid_edges) of XSection_builder.m:62-128.

    S1 = r_e^2 N_A Fcoh^2 / M and S2 = r_e^2 N_A Sincoh / M per element (q_range and
    z_range are the 1-based MATLAB index ranges nqrange = 2:38, nzrange = 2:27),
    divided by the element's a2, tabulated on q_axis = log2(wvt hbarc / mc2 / 2) and
    a_axis = log2(a1/a2), scaled so that the maximum is 255, and regridded onto
    nq_bins x na_bins regular points spanning both axes.  grid_edges =
    (q_min, q_max, a_min, a_max, S1_max, S2_max)."""
    a1, a2 = fit_basis(phys["attenuation"], z_range)
    qi = np.array(list(q_range)) - 1
    zi = np.array(list(z_range)) - 1
    pref = phys["r_e"] ** 2 * phys["N_A"] / phys["M"][zi]
    S1 = pref[None, :] * phys["Fcoh"][np.ix_(qi, zi)] ** 2
    S2 = pref[None, :] * phys["Sincoh"][np.ix_(qi, zi)]
    q_axis = np.log2(phys["wvt_list"][qi] * phys["hbarc"] / phys["mc2"] / 2.0)
    a_axis = np.log2(a1 / a2)
    if not (np.all(np.diff(q_axis) > 0) and np.all(np.diff(a_axis) > 0)):
        raise ValueError("the q and log2(a1/a2) axes must increase (interpn needs grid vectors)")
    s1 = S1 / a2[None, :]
    s2 = S2 / a2[None, :]
    s1_max, s2_max = float(s1.max()), float(s2.max())
    s1 = s1 * 255.0 / s1_max
    s2 = s2 * 255.0 / s2_max
    q_min, q_max, a_min, a_max = q_axis[0], q_axis[-1], a_axis[0], a_axis[-1]
    q_grid = q_min + np.arange(nq_bins) * (q_max - q_min) / (nq_bins - 1)
    a_grid = a_min + np.arange(na_bins) * (a_max - a_min) / (na_bins - 1)
    S1_grid = makima2(q_axis, a_axis, s1, q_grid, a_grid)
    S2_grid = makima2(q_axis, a_axis, s2, q_grid, a_grid)
    return S1_grid, S2_grid, (q_min, q_max, a_min, a_max, s1_max, s2_max)


def read_spekcalc(path: str):
    """(energy keV, N per keV cm^2 mAs at 1 m) from a SpekCalc text export."""
    rows = []
    with open(path) as f:
        started = False
        for line in f:
            if line.startswith("Energy[keV]"):
                started = True
                continue
            if started:
                parts = line.split()
                if len(parts) >= 2:
                    rows.append((float(parts[0]), float(parts[1])))
    if not rows:
        raise ValueError(f"{path}: no spectrum table found")
    a = np.array(rows)
    return a[:, 0], a[:, 1]


def spectrum_response(n_energy: int, energy: np.ndarray, counts: np.ndarray,
                      e_min: float = 8.0, e_max: float = 80.0,
                      exposure_s: float = 0.100e-3, current_ma: float = 10.0):
    """Ndet[e_src, e_det] [n_energy x n_energy] and the bin edges (spectrum_builder.m:8-95).

    The tabulated spectrum (energy = bin edges, counts[k] = photons per keV cm^2 mAs
    at 1 m in bin [energy[k], energy[k+1]]) is scaled by the exposure (mAs) and
    100^2 cm^2; with detector resolution sigma(E) = 0.5 (1.61 + 0.025 E) keV,
      Ndet[s, d] = sum_{i,j} wt_src[s, i] wt_det[d, j] N_i G(E_i - E_j; sigma_i),
    wt the overlap (keV) of table bin i (or j) with coarse bin s (or d) and G the
    normalised Gaussian: a double integral over source and detector energy on the
    table's grid."""
    energy = np.asarray(energy, dtype=np.float64)
    counts = np.asarray(counts, dtype=np.float64) * exposure_s * current_ma * 100.0 ** 2
    e_left, e_right = energy[:-1], energy[1:]
    e_hd = 0.5 * (e_left + e_right)
    n_hd = counts[:-1]
    edges = e_min + np.arange(n_energy + 1) * (e_max - e_min) / n_energy
    sigma = 0.5 * (1.61 + 0.025 * e_hd)
    # window[j, i]: detector energy e_hd[j] from source energy e_hd[i]
    window = (n_hd[None, :] * np.exp(-(e_hd[None, :] - e_hd[:, None]) ** 2 / (2.0 * sigma[None, :] ** 2))
              / np.sqrt(2.0 * np.pi) / sigma[None, :])
    lo, hi = edges[:-1], edges[1:]
    wt = np.maximum(0.0, np.minimum(hi[:, None], e_right[None, :]) -
                    np.maximum(lo[:, None], e_left[None, :]))             # [bin, table bin]
    ndet = wt @ window.T @ wt.T                                           # [s, d]
    return ndet, edges


def load_spectrum_hd(path: str):
    """(energy keV, N per keV cm^2 mAs at 1 m) of the SpekCalc spectrum struct in a
    MAT-file (spectral_data_HD.mat: 1441 points 8..80 keV at 0.05 keV,
    spectrum_builder.m:9)."""
    from .matfile import load_mat
    sp = load_mat(path, ["spectrum"])["spectrum"]
    return (np.asarray(sp["energy"], np.float64).ravel(),
            np.asarray(sp["counts"], np.float64).ravel())


# ---------------------------------------------------------------------------
# XRD_builder.m

E_CU_KALPHA = 8.04            # keV, the diffractometer's source (XRD_builder.m:10)
HBARC_XRD = 1.973269804       # keV Angstrom (XRD_builder.m:11)
MC2_XRD = 510.99895           # keV (XRD_builder.m:12)

# (file, calibration factor, polarization-corrected) per material, in the order of
# XRD_builder.m:36-52 (cellulose), :86-101 (aluminium alloy), :136-152 (ammonium nitrate)
PAPER_PATTERNS = (("cellulose.mat", 0.00003, False),
                  ("aluminum_alloy.mat", 0.005, True),
                  ("ammoniumnitrate.mat", 0.0007, False))


def pattern_q(D: np.ndarray, e_src_kev: float = E_CU_KALPHA, hbarc: float = HBARC_XRD,
              polarization: bool = False):
    """(q 1/Angstrom, intensity) of a measured pattern D = [2theta deg, counts]
    (XRD_builder.m:40-50): zero-count points are dropped, q = 2 E sin(theta/2) / hbarc
    and, with polarization, the counts are divided by (1 + cos^2 theta) / 2.  The
    points are returned sorted by q (stable), as interp1 sorts unsorted sample
    points; aluminum_alloy.mat has two 2theta values out of order."""
    D = np.asarray(D, np.float64)
    D = D[D[:, 1] > 0]
    theta = D[:, 0] / 180.0 * np.pi
    q = np.sin(theta / 2.0) * (2.0 * e_src_kev) / hbarc
    inten = D[:, 1].copy()
    if polarization:
        inten = inten / ((1.0 + np.cos(theta) ** 2) / 2.0)
    order = np.argsort(q, kind="stable")
    if np.any(np.diff(q[order]) == 0):
        raise ValueError("pattern has repeated 2-theta values (interp1 needs distinct points)")
    return q[order], inten[order]


def makima1(x: np.ndarray, v: np.ndarray, xq: np.ndarray) -> np.ndarray:
    """interp1(x, v, xq, 'makima', NaN): modified Akima, NaN outside [x0, xn]."""
    from scipy.interpolate import Akima1DInterpolator
    return Akima1DInterpolator(x, v, method="makima", extrapolate=False)(xq)


def fine_q_axis(n_qf: int = 768, q_min: float = 0.5, q_max: float = 6.0):
    """Centres of n_qf equal bins on [q_min, q_max] (XRD_builder.m:25-27)."""
    edges = q_min + np.arange(n_qf + 1) / n_qf * (q_max - q_min)
    return 0.5 * (edges[:-1] + edges[1:])


def theory_cross_sections(q: np.ndarray, mean_a1: float, mean_a2: float, grid_edges,
                          S1_grid: np.ndarray, S2_grid: np.ndarray,
                          hbarc: float = HBARC_XRD, mc2: float = MC2_XRD):
    """(S1, S2) per unit volume of a material with mean attenuation coefficients
    (mean_a1, mean_a2) at momentum transfers q (XRD_builder.m:54-61): the 0..255
    tables at (log2(q hbarc / mc2 / 2), log2(a1/a2)) by tensor-product makima (q
    first), times mean_a2 * S_max / 255.  Points off the table grid are NaN, as
    interpn returns there."""
    q_min, q_max, a_min, a_max, s1_max, s2_max = grid_edges
    nq, na = S1_grid.shape
    qg = q_min + np.arange(nq) * (q_max - q_min) / (nq - 1)
    ag = a_min + np.arange(na) * (a_max - a_min) / (na - 1)
    lq = np.log2(np.asarray(q, np.float64) * hbarc / mc2 / 2.0)
    la = float(np.log2(mean_a1 / mean_a2))
    tol = 1e-12 * max(1.0, abs(q_max - q_min), abs(a_max - a_min))
    off = (lq < q_min - tol) | (lq > q_max + tol)
    out = []
    for S, smax in ((S1_grid, s1_max), (S2_grid, s2_max)):
        if la < a_min - tol or la > a_max + tol:
            v = np.full(lq.shape, np.nan)
        else:
            v = makima2(qg, ag, np.asarray(S, np.float64), lq, np.array([la]))[:, 0]
            v[off] = np.nan
        out.append(mean_a2 * smax / 255.0 * v)
    return out[0], out[1]


def xrd_builder(a1: np.ndarray, a2: np.ndarray, segmentation: np.ndarray, grid_edges,
                S1_grid: np.ndarray, S2_grid: np.ndarray, patterns, n_qf: int = 768,
                q_min: float = 0.5, q_max: float = 6.0):
    """(ifine [n_qf, M], q_fine) of XRD_builder.m for M = len(patterns) materials.

    patterns[m] = (D, calibration, polarization) for material id m + 1, D the measured
    [2theta deg, counts] table.  Material m's form factor on the fine q axis is
    calibration * makima(q_exp, I_exp) inside the measured q range, and the theory
    S1 + S2 at the material's mean (a1, a2) (a1, a2 in 1/cm) outside it
    (XRD_builder.m:36-66)."""
    qf = fine_q_axis(n_qf, q_min, q_max)
    seg = np.asarray(segmentation)
    out = np.zeros((n_qf, len(patterns)))
    for m, (D, calib, pol) in enumerate(patterns):
        mask = seg == m + 1
        if not mask.any():
            raise ValueError(f"material {m + 1} has no voxels in the segmentation")
        qx, ix = pattern_q(D, polarization=pol)
        f = calib * makima1(qx, ix, qf)
        s1, s2 = theory_cross_sections(qf, float(np.mean(a1[mask])), float(np.mean(a2[mask])),
                                       grid_edges, S1_grid, S2_grid)
        nan = np.isnan(f)
        f[nan] = s1[nan] + s2[nan]
        out[:, m] = f
    return out, qf


# ---------------------------------------------------------------------------
# phantom_generator.m

# photoelectric and Compton amplitudes of phantom_generator.m:31-36 (K1 cm^2 keV^3/mol,
# K2 cm^2/mol, energy slope m, Z slope n, mc2 keV)
PHANTOM_PHYS = dict(K1=13.96, K2=0.30, m=3.00, n=4.20, mc2=510.975)


def material_coefficients(density: float, Z: Sequence[int], weight: Sequence[float],
                          M: Sequence[float], phys: dict = PHANTOM_PHYS):
    """(a1, a2) in 1/cm of a material from its mass fractions `weight` of elements Z
    with molar masses M: a1 = K1/mc2^m rho sum_i w_i Z_i^n / M_i and
    a2 = K2 rho sum_i w_i Z_i / M_i (phantom_generator.m:54-60, :140-145).  A
    molecule N_i x Z_i has w_i = N_i M_i / sum_j N_j M_j (phantom_generator.m:54-60
    writes that case directly)."""
    Z = np.asarray(Z, np.float64)
    w = np.asarray(weight, np.float64)
    M = np.asarray(M, np.float64)
    a1 = phys["K1"] / phys["mc2"] ** phys["m"] * density * np.sum(w * Z ** phys["n"] / M)
    a2 = phys["K2"] * density * np.sum(w * Z / M)
    return float(a1), float(a2)


def _molecule(density, Z, N, M):
    N = np.asarray(N, np.float64)
    M = np.asarray(M, np.float64)
    return material_coefficients(density, Z, N * M / np.sum(N * M), M)


def phantom_generator(molar_mass: np.ndarray, nx: int = 100, ny: int = 100,
                      size_mm: float = 200.0):
    """(a1, a2 in 1/cm, segmentation) of the paper's phantom (phantom_generator.m):
    three cellulose ellipses (id 1), one aluminium-alloy ellipse (id 2) and two
    ammonium-nitrate ellipses (id 3) on an nx x ny grid over size_mm^2, painted in
    that order (later objects overwrite).  molar_mass[Z-1] in g/mol
    (physical_constants.mat phys.M) serves the alloy."""
    x, y = np.meshgrid((np.arange(1, nx + 1) - 0.5) * size_mm / nx,
                       (np.arange(1, ny + 1) - 0.5) * size_mm / ny, indexing="ij")
    a1 = np.zeros((nx, ny))
    a2 = np.zeros((nx, ny))
    seg = np.zeros((nx, ny), np.int32)
    cellulose = _molecule(0.1, [1, 6, 8], [10, 6, 5], [1.00784, 12.0107, 15.999])
    al_Z = [26, 14, 29, 30, 22, 25, 12, 13]
    al_w = np.array([0.12, 1.9, 0.005, 0.01, 0.009, 0.58, 5.17, 92.206]) / 100.0
    alloy = material_coefficients(2.6, al_Z, al_w, np.asarray(molar_mass)[np.array(al_Z) - 1])
    nitrate = _molecule(1.725, [1, 7, 8], [4, 2, 3], [1.00784, 14.0067, 15.999])
    # (material id, coefficients, x0, y0 mm, tilt deg, semi-axes a, b mm)
    objects = [(1, cellulose, 40.0, 110.0, 0.0, 24.0, 80.0),
               (1, cellulose, 130.0, 150.0, -5.0, 40.0, 40.0),
               (1, cellulose, 130.0, 60.0, 5.0, 40.0, 30.0),
               (2, alloy, 40.0, 130.0, 10.0, 12.0, 16.0),
               (3, nitrate, 40.0, 80.0, 10.0, 12.0, 12.0),
               (3, nitrate, 130.0, 60.0, 10.0, 20.0, 12.0)]
    for mid, (c1, c2), x0, y0, tilt, a, b in objects:
        t = tilt / 180.0 * np.pi
        inside = ((((x - x0) * np.cos(t) - (y - y0) * np.sin(t)) / a) ** 2
                  + (((x - x0) * np.sin(t) + (y - y0) * np.cos(t)) / b) ** 2) < 1
        a1[inside] = c1
        a2[inside] = c2
        seg[inside] = mid
    return a1, a2, seg


# ---------------------------------------------------------------------------
# main.m:12-63

def paper_problem(ref_dir: str, n_views: int = 32, n_cols: int = 1024, n_energy: int = 64):
    """The problem main.m assembles before Forward_XRD, from the reference's own input
    files under ref_dir: the SpekCalc HD spectrum through spectrum_builder (NRG = 64),
    the cross-section tables (NQ = NA = 64), the 100 x 100 phantom, and the three
    measured patterns through XRD_builder (NQF = 768, NWT = 256), on the scanner
    literals of integrate_diffraction.cu:81-140 with SRC = 32 views and COL = 1024
    columns (Forward_XRD.cu:16-23).  a1, a2 are given to the projector in 1/voxel
    with main.m's conversion vox_over_cm = 0.1 / (NX / 100) (main.m:48-61).
    x_true is the fine form factor linearly resampled at the coarse q centres, and
    compton_scale = XS_grid_edges(6) / 255 converts the projector's C output to
    photons (main.m:68)."""
    from .descriptors import Discretization, Geometry, q_axis
    from .matfile import load_mat
    from .synthetic import Problem

    phys = load_physics(os.path.join(ref_dir, "Physics_inputs"))
    e_hd, n_hd = load_spectrum_hd(os.path.join(ref_dir, "Spectrum_80kVp", "spectral_data_HD.mat"))
    ndet, _ = spectrum_response(n_energy, e_hd, n_hd)
    S1, S2, edges = xsection_tables(phys, 64, 64)
    a1, a2, seg = phantom_generator(phys["M"])
    pats = []
    for fname, calib, pol in PAPER_PATTERNS:
        D = load_mat(os.path.join(ref_dir, "open_source_data", fname), ["D"])["D"]
        pats.append((D, calib, pol))
    ifine, _ = xrd_builder(a1, a2, seg, edges, S1, S2, pats, n_qf=768)
    nx, ny = seg.shape
    g = Geometry.reference(nx, ny, n_views=n_views, n_cols=n_cols, n_rows=1)
    d = Discretization.reference(n_energy=n_energy, n_q=256, n_qf=768, n_mat=3)
    q, _, qf, _ = q_axis(d)
    x_true = np.stack([np.interp(q, qf, ifine[:, m]) for m in range(3)], axis=1)
    vox_over_cm = 0.1 / (nx / 100.0)
    prob = Problem(name="paper100", geometry=g, disc=d,
                   spectrum=ndet.astype(np.float32), ifine=ifine.astype(np.float32),
                   segmentation=seg, a1=(a1 * vox_over_cm).astype(np.float32),
                   a2=(a2 * vox_over_cm).astype(np.float32),
                   s2_table=np.asarray(S2, np.float32), lut=tuple(float(v) for v in edges[:4]),
                   x_true=x_true, photons_per_view=0.0, mlem_iterations=1000,
                   voxel_mm=nx / 100.0, compton_scale=float(edges[5]) / 255.0)
    return prob


if __name__ == "__main__":
    # python -m xrd_tomography_amd.builders REF_DIR OUT_DIR: write the paper's problem
    # (paper_problem) in the on-disk format of problem_io
    import sys
    from .problem_io import save_problem
    if len(sys.argv) != 3:
        raise SystemExit("usage: python -m xrd_tomography_amd.builders REF_DIR OUT_DIR")
    prob = paper_problem(sys.argv[1])
    save_problem(prob, sys.argv[2])
    print(f"wrote {sys.argv[2]}: {prob.name}, compton_scale {prob.compton_scale:.9g}")
