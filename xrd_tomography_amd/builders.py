"""Input builders restated from the reference's MATLAB builders.

Host-side, run once per scanner/material set; their outputs are inputs of
xrd_forward (the spectrum response matrix, the Compton S2 table and its grid
edges).  They read the reference's physics tables, which ship as MATLAB v5 files
(attenuation_data.mat, Hubbell_data.mat, physical_constants.mat; scipy.io reads
them as plain arrays) and a SpekCalc text export of the source spectrum.

* ``fit_basis``: photoelectric + Compton basis fit of mu/rho per element
  (XSection_builder.m:15-60, least squares without intercept).
* ``xsection_tables``: coherent and incoherent cross-sections over the
  (log2 q, log2 a1/a2) grid, scaled to 0..255 (XSection_builder.m:62-128).
* ``spectrum_response``: the detector-response-smeared spectrum matrix
  Ndet[e_src, e_det] (spectrum_builder.m:8-95).
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
