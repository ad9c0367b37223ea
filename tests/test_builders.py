"""Input builders and the on-disk problem format (CPU).

The builders restate XSection_builder.m and spectrum_builder.m; they are pinned by
closed forms (exact recovery of a two-basis mu/rho, the Thomson limit of the
Klein-Nishina basis, modified-Akima reproduction of bilinear data, photon
conservation of the detector-response integral).  Checks on the reference's own
physics tables run where XRD_REFERENCE_DIR points at them.
"""
import json
import os

import numpy as np
import pytest

from xrd_tomography_amd import builders as B
from xrd_tomography_amd.problem_io import load_problem, save_problem
from xrd_tomography_amd.synthetic import make_problem

REF = B.reference_dir()
needs_ref = pytest.mark.skipif(REF is None, reason="XRD_REFERENCE_DIR does not hold the "
                                                   "reference's Physics_inputs/")


# ---------------------------------------------------------------------------
# on-disk format

def test_problem_round_trip(tmp_path, oracle):
    p = make_problem("small")
    rows = p.n_rows
    y = np.arange(rows, dtype=np.float32)
    save_problem(p, str(tmp_path), counts=y, background=0.5 * y)
    q, extras = load_problem(str(tmp_path))
    assert q.geometry.as_dict() == p.geometry.as_dict()
    assert q.disc.as_dict() == p.disc.as_dict()
    assert q.lut == tuple(p.lut)
    for n in ("spectrum", "ifine", "segmentation", "a1", "a2", "s2_table", "x_true"):
        a, b = getattr(p, n), getattr(q, n)
        assert a.shape == b.shape and a.dtype == b.dtype and np.array_equal(a, b), n
        assert np.array_equal(p.flat(n), q.flat(n))
    assert np.array_equal(extras["counts"], y) and np.array_equal(extras["background"], 0.5 * y)
    N0, C0, _ = oracle.forward(p, with_A=False, nthreads=2)
    N1, C1, _ = oracle.forward(q, with_A=False, nthreads=2)
    assert np.array_equal(N0, N1) and np.array_equal(C0, C1)
    # the files are the ABI memory order: segmentation is x fastest
    raw = np.fromfile(os.path.join(str(tmp_path), "segmentation.bin"), dtype="<i4")
    assert raw[1] == p.segmentation[1, 0] and raw[p.geometry.nx] == p.segmentation[0, 1]


def test_problem_rejects_inconsistent_files(tmp_path):
    p = make_problem("small")
    save_problem(p, str(tmp_path))
    fn = os.path.join(str(tmp_path), "a1.bin")
    data = open(fn, "rb").read()
    open(fn, "wb").write(data[:-4])
    with pytest.raises(ValueError, match="size"):
        load_problem(str(tmp_path))
    open(fn, "wb").write(data)
    doc = json.load(open(os.path.join(str(tmp_path), "problem.json")))
    doc["geometry"]["nx"] += 1
    json.dump(doc, open(os.path.join(str(tmp_path), "problem.json"), "w"))
    with pytest.raises(ValueError):
        load_problem(str(tmp_path))
    with pytest.raises(ValueError):
        save_problem(p, str(tmp_path), counts=np.zeros(3, np.float32))


# ---------------------------------------------------------------------------
# XSection_builder.m

def test_klein_nishina_basis_thomson_limit():
    """f2(e) = sigma_KN / (2 pi r_e^2) = 4/3 (1 - 2e + 26e^2/5 - 133e^3/10 + 1144e^4/35
    - ...) for small e (the Thomson limit and its Klein-Nishina corrections)."""
    for e, tol in ((1e-2, 1e-7), (3e-2, 3e-6)):
        series = 4.0 / 3.0 * (1.0 - 2.0 * e + 26.0 / 5.0 * e ** 2 - 133.0 / 10.0 * e ** 3
                              + 1144.0 / 35.0 * e ** 4)
        assert abs(B.basis_f2(e) - series) <= tol * series, e
    assert abs(B.basis_f1(0.5) - 8.0) < 1e-15


def test_fit_basis_recovers_two_basis_data():
    e_mev = np.array([0.03, 0.04, 0.05, 0.06, 0.08, 0.1, 0.15, 0.2])
    att = []
    truth = []
    for z in range(1, 6):
        a1, a2 = 1e-5 * z ** 3, 0.15 + 0.01 * z
        mu = a1 * B.basis_f1(e_mev / 0.511) + a2 * B.basis_f2(e_mev / 0.511)
        att.append(np.stack([e_mev, mu], 1))
        truth.append((a1, a2))
    f1, f2 = B.fit_basis(att, z_range=range(1, 6), mc2_mev=0.511)
    assert np.allclose(f1, [t[0] for t in truth], rtol=1e-10)
    assert np.allclose(f2, [t[1] for t in truth], rtol=1e-10)


def test_makima2_reproduces_bilinear_data():
    """Modified Akima reproduces linear data in each dimension, so the tensor
    product reproduces v = c0 + c1 x + c2 y + c3 x y on non-uniform grids."""
    x = np.cumsum(np.r_[0.0, np.random.default_rng(1).uniform(0.2, 1.0, 11)])
    y = np.cumsum(np.r_[0.0, np.random.default_rng(2).uniform(0.1, 0.7, 8)])
    f = lambda X, Y: 1.5 - 0.3 * X + 2.0 * Y + 0.25 * X * Y  # noqa: E731
    v = f(x[:, None], y[None, :])
    xq = np.linspace(x[0], x[-1], 17)
    yq = np.linspace(y[0], y[-1], 13)
    out = B.makima2(x, y, v, xq, yq)
    assert np.allclose(out, f(xq[:, None], yq[None, :]), rtol=0, atol=1e-12)


@needs_ref
def test_xsection_tables_from_reference_data():
    phys = B.load_physics(os.path.join(REF, "Physics_inputs"))
    a1, a2 = B.fit_basis(phys["attenuation"])
    # least squares: the residual is orthogonal to both basis columns (the normal
    # equations), and the two-basis model stays within 25% of the tabulated mu/rho
    for z in range(2, 28):
        ac = phys["attenuation"][z - 1]
        e = ac[:, 0] / 0.5110
        X = np.stack([B.basis_f1(e), B.basis_f2(e)], 1)
        r = ac[:, 1] - X @ np.array([a1[z - 2], a2[z - 2]])
        assert np.all(np.abs(X.T @ r) <= 1e-9 * np.abs(X.T @ ac[:, 1])), z
        assert np.max(np.abs(r) / ac[:, 1]) < 0.25, z
    S1, S2, edges = B.xsection_tables(phys)
    assert S1.shape == (64, 64) and S2.shape == (64, 64)
    assert np.isfinite(S1).all() and np.isfinite(S2).all()
    q_min, q_max, a_min, a_max, s1_max, s2_max = edges
    assert q_min < q_max and a_min < a_max and s1_max > 0 and s2_max > 0
    # grid corners are data nodes: values there are the scaled cross-sections
    qi, zi = np.arange(1, 38), np.arange(1, 27)
    s2 = (phys["r_e"] ** 2 * phys["N_A"] * phys["Sincoh"][np.ix_(qi, zi)] / phys["M"][zi]
          / a2[None, :]) * 255.0 / s2_max
    for (i, j), (k, m) in (((0, 0), (0, 0)), ((-1, -1), (-1, -1)), ((0, -1), (0, -1))):
        assert abs(S2[i, j] - s2[k, m]) < 1e-9
    assert S2.max() <= 255.0 + 1e-6 and S2.min() > -1.0


# ---------------------------------------------------------------------------
# spectrum_builder.m

def test_spectrum_response_conserves_photons():
    """A flat spectrum on a fine table: every photon of a source bin well inside
    the range is detected in some bin, so the row sum is N * (bin width) * (mAs
    * 100^2) up to the quadrature of the Gaussian on the table grid."""
    E = np.arange(8.0, 80.0 + 1e-9, 0.05)
    N = np.full(E.shape, 2.0)
    nd, edges = B.spectrum_response(36, E, N)
    scale = 0.100e-3 * 10.0 * 100.0 ** 2
    width = edges[1] - edges[0]
    rows = nd.sum(axis=1)
    mid = slice(6, 30)
    assert np.allclose(rows[mid], 2.0 * scale * width, rtol=2e-3)
    # detected energy concentrates on the diagonal
    assert np.all(np.abs(np.argmax(nd, axis=1) - np.arange(36))[mid] <= 1)


@needs_ref
def test_spectrum_response_from_reference_spectrum():
    E, N = B.read_spekcalc(os.path.join(REF, "Spectrum_80kVp", "SpeckCacl_80kVp.txt"))
    assert E[0] == 8.0 and E[-1] == 80.0 and len(E) == 73
    nd, edges = B.spectrum_response(61, E, N)
    assert nd.shape == (61, 61) and np.all(nd >= 0)
    assert abs(edges[0] - 8.0) < 1e-12 and abs(edges[-1] - 80.0) < 1e-12
    # the emitted total is conserved up to what leaks past the 8-80 keV range
    scale = 0.100e-3 * 10.0 * 100.0 ** 2
    emitted = np.sum(N[:-1] * scale * np.diff(E))
    assert 0.9 * emitted < nd.sum() <= emitted * 1.0001
