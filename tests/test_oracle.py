"""Pins of the FP64 oracle by checks that do not restate its formulas.

The reference ships no tests, fixtures or golden vectors (it is MATLAB + CUDA
that cannot run here), so each oracle function is pinned by an independent
method: a brute-force or textbook computation (slab clipping, numpy.interp,
numpy matmul), a closed form (telescoping q edges, Gaussian normalisation,
Beer-Lambert attenuation), or a symmetry (90-degree rotation of the scanner).
"""
import math

import numpy as np
import pytest


# ---------------------------------------------------------------------------
# GetIntersection (GetIntersection.cu:9-57)

def _slab_entry(x1, y1, x2, y2, nx, ny):
    """Liang-Barsky clipping: parameter and point where P1->P2 enters [0,nx]x[0,ny]."""
    dx, dy = x2 - x1, y2 - y1
    t0, t1 = 0.0, 1.0
    for p, q in ((-dx, x1), (dx, nx - x1), (-dy, y1), (dy, ny - y1)):
        if p == 0:
            if q < 0:
                return None
            continue
        t = q / p
        if p < 0:
            t0 = max(t0, t)
        else:
            t1 = min(t1, t)
    if t0 > t1:
        return None
    return x1 + t0 * dx, y1 + t0 * dy


def test_intersection_matches_slab_clipping(oracle):
    rng = np.random.default_rng(0)
    nx, ny = 53, 31
    for _ in range(2000):
        a = rng.uniform(0, 2 * np.pi)
        r = rng.uniform(60, 400)
        x1, y1 = nx / 2 + r * np.cos(a), ny / 2 + r * np.sin(a)
        x2, y2 = rng.uniform(0.5, nx - 0.5), rng.uniform(0.5, ny - 0.5)
        side, ex, ey = oracle.get_intersection(x1, y1, x2, y2, nx, ny)
        ref = _slab_entry(x1, y1, x2, y2, nx, ny)
        assert side >= 0 and ref is not None
        assert abs(ex - ref[0]) < 1e-9 * max(1, r) and abs(ey - ref[1]) < 1e-9 * max(1, r)
        # the reported side is the one the entry point lies on
        on = [abs(ey) < 1e-7, abs(ex - nx) < 1e-7, abs(ey - ny) < 1e-7, abs(ex) < 1e-7]
        assert on[side]


def test_intersection_special_cases(oracle):
    nx, ny = 10, 8
    # horizontal ray: sides y=0 and y=ny are parallel (|den| <= eps), x=nx misses, x=0 hits
    assert oracle.get_intersection(-10, 0.5, 3.5, 0.5, nx, ny) == (3, 0.0, 0.5)
    # through the corner (0,0): y=0 is tested first and wins
    s, ex, ey = oracle.get_intersection(-10, -10, 5.5, 5.5, nx, ny)
    assert s == 0 and ex == 0.0 and ey == 0.0
    # from beyond the far corner: side x=nx precedes y=ny in the fixed order
    s, ex, ey = oracle.get_intersection(20, 18, 5, 3, nx, ny)
    assert s == 1 and abs(ex - 10) < 1e-12 and abs(ey - 8) < 1e-12
    # start inside: no crossing, NaN
    s, ex, ey = oracle.get_intersection(5.5, 5.5, 6.5, 6.5, nx, ny)
    assert s == -1 and math.isnan(ex) and math.isnan(ey)
    # nearly parallel to y=0 within eps: skipped, enters through x=0
    s, _, _ = oracle.get_intersection(-100.0, 0.5, 3.5, 0.5 + 1e-9, nx, ny)
    assert s == 3
    # the tolerance is on the segment parameter (eps*|P2-P1| in length): for a
    # 1e6-long segment the far side x=nx, 6.5 beyond the end point, is accepted and,
    # being tested before x=0, wins (GetIntersection.cu:43,47,50)
    s, ex, _ = oracle.get_intersection(-1e6, 0.5, 3.5, 0.5, nx, ny)
    assert s == 1 and abs(ex - 10.0) < 1e-9


def test_sample_count(oracle):
    nx = ny = 20
    # entry (0, 5.5) -> voxel centre (L, 5.5): L/step exact multiples and fractions
    for xv, step in ((3.5, 0.25), (3.5, 0.5), (2.75, 0.25), (7.3, 0.3)):
        side, ex, ey, n, _ = oracle.trace_ray(-5, 5.5, xv, 5.5, nx, ny, step)
        L = xv
        expect = int(math.floor(L / step + 1e-12)) + 1
        assert side == 3 and n == expect


# ---------------------------------------------------------------------------
# attenuator (attenuator.cu:9-109) with exact bilinear filtering and zero border

def _interp_zero_border(profile, t):
    """Linear interpolation of texel values centred at i+0.5, zero outside (numpy.interp
    on a zero-padded profile)."""
    n = len(profile)
    xs = np.concatenate(([-0.5], np.arange(n) + 0.5, [n + 0.5]))
    vs = np.concatenate(([0.0], profile, [0.0]))
    return np.interp(t, xs, vs, left=0.0, right=0.0)


def test_line_integral_separable_field(oracle):
    """For f(x,y) = g(x) h(y), bilinear filtering factorises into two 1-D linear
    interpolations; the Riemann sum is then computed here from the slab entry."""
    rng = np.random.default_rng(1)
    nx, ny, step = 23, 17, 0.25
    gx, hy = rng.uniform(0, 1, nx), rng.uniform(0, 1, ny)
    fmap = np.outer(hy, gx)  # [y, x] -> x fastest
    for _ in range(200):
        a = rng.uniform(0, 2 * np.pi)
        x1, y1 = nx / 2 + 80 * np.cos(a), ny / 2 + 80 * np.sin(a)
        x2, y2 = rng.integers(0, nx) + 0.5, rng.integers(0, ny) + 0.5
        ex, ey = _slab_entry(x1, y1, x2, y2, nx, ny)
        L = math.hypot(x2 - ex, y2 - ey)
        n = int(math.floor(L / step)) + 1
        t = step * np.arange(n)
        px, py = ex + (x2 - ex) / L * t, ey + (y2 - ey) / L * t
        ref = step * np.sum(_interp_zero_border(gx, px) * _interp_zero_border(hy, py))
        got = oracle.line_integrals(fmap.reshape(1, -1), nx, ny, step, x1, y1, x2, y2)[0]
        assert abs(got - ref) <= 1e-10 * max(abs(ref), 1e-12)


def test_line_integral_linear_and_zero(oracle):
    rng = np.random.default_rng(2)
    nx, ny, step = 12, 14, 0.25
    f, g = rng.uniform(0, 1, (2, nx * ny))
    seg = (-30.0, 3.0, 7.5, 9.5)
    lf = oracle.line_integrals(f, nx, ny, step, *seg)[0]
    lg = oracle.line_integrals(g, nx, ny, step, *seg)[0]
    lfg = oracle.line_integrals(2.5 * f - 0.5 * g, nx, ny, step, *seg)[0]
    assert abs(lfg - (2.5 * lf - 0.5 * lg)) < 1e-12
    assert oracle.line_integrals(np.zeros(nx * ny), nx, ny, step, *seg)[0] == 0.0
    both = oracle.line_integrals(np.stack([f, g]), nx, ny, step, *seg)
    assert both[0] == lf and both[1] == lg


def test_bilinear_texel_centres(oracle):
    rng = np.random.default_rng(3)
    m = rng.uniform(0, 1, (5, 7))  # [y, x]
    for j in range(5):
        for i in range(7):
            assert oracle.bilinear(m.reshape(-1), 7, 5, i + 0.5, j + 0.5) == m[j, i]
    assert oracle.bilinear(m.reshape(-1), 7, 5, -0.5, 2.5) == 0.0  # border texel
    assert abs(oracle.bilinear(m.reshape(-1), 7, 5, 0.0, 2.5) - 0.5 * m[2, 0]) < 1e-15


# ---------------------------------------------------------------------------
# find_src_nrg_range and wvt_value

def test_energy_window_table(oracle):
    for E, hw in ((64, 5), (61, 5), (16, 5), (11, 5), (20, 2)):
        for nrg in range(E):
            lo, hi = oracle.find_src_nrg_range(nrg, E, hw)
            assert hi - lo == 2 * hw and 0 <= lo and hi <= E - 1
            if hw <= nrg <= E - 1 - hw:
                assert (lo, hi) == (nrg - hw, nrg + hw)  # centred away from the ends
            elif nrg < hw:
                assert lo == 0
            else:
                assert hi == E - 1


@pytest.mark.parametrize("nq,qmin,qmax,dq0", [(256, 0.5, 6.0, 0.01), (128, 0.5, 3.5, 0.01)])
def test_q_axis_closed_form(oracle, nq, qmin, qmax, dq0):
    qs, dqs = zip(*[oracle.wvt_value(n, nq, qmin, qmax, dq0) for n in range(nq)])
    qs, dqs = np.array(qs), np.array(dqs)
    assert abs(dqs.sum() - (qmax - qmin)) < 1e-12          # edges telescope to the range
    assert abs(qs[0] - 0.5 * dqs[0] - qmin) < 1e-15         # first edge at q_min
    assert abs(qs[-1] + 0.5 * dqs[-1] - qmax) < 1e-12       # last edge at q_max
    assert np.all(np.diff(dqs) > 0)                         # parabolic: widths increase
    c = (qmax - qmin - dq0 * nq) / nq ** 2
    assert abs(dqs[0] - (dq0 + c)) < 1e-15


def test_channel_match_partition(oracle):
    """The Compton-shifted source bin splits over consecutive detector bins with
    fractions summing to 1 (integrate_diffraction.cu:323-328)."""
    edges = 19.5 + np.arange(62)
    for er in (1.0, 1.0003, 1.004, 1.05):
        for s in range(20, 55):
            fr = [oracle.channel_match(edges[s], edges[s + 1], er, edges[t], edges[t + 1])
                  for t in range(61)]
            assert abs(sum(fr) - 1.0) < 1e-12 and min(fr) >= 0.0


# ---------------------------------------------------------------------------
# scattering geometry (integrate_diffraction.cu:81-255)

def _geom(**kw):
    from xrd_tomography_amd.descriptors import Geometry
    base = dict(nx=64, ny=64, n_views=8, n_cols=16, n_rows=1, center_x=32.0, center_y=32.0,
                r_src=150.0, r_det=100.0, view_angle0=0.0, view_angle_step=2 * np.pi / 8,
                col_pitch=2.0, row_pitch=1.5, det_z0=6.0, target_angle=np.pi / 6,
                focal_spot=0.25, wedge_angle=np.radians(0.5), voxel_size=1.0)
    base.update(kw)
    return Geometry(**base)


def test_geometry_hand_triangle(oracle):
    """View 0, voxel at the rotation centre: a = (0, R, -R tan(w/2)),
    b = (rdet, Rd, z0 + R tan(w/2)); theta, sin(theta/2) and the geometric weight
    follow from elementary trigonometry."""
    g = _geom()
    col = 11
    R, Rd, w = g.r_src, g.r_det, g.wedge_angle
    rdet = (col - g.n_cols / 2 + 0.5) * g.col_pitch
    a = np.array([0.0, R, -R * np.tan(w / 2)])
    b = np.array([rdet, Rd, g.det_z0 + R * np.tan(w / 2)])
    theta = np.arccos(np.dot(a, b) / np.linalg.norm(a) / np.linalg.norm(b))
    o = oracle.scatter_geometry(g, 0, col, 0, 31, 31)  # voxel centre (31.5, 31.5)
    # the voxel is half a voxel off the centre: compare with the exact centre offsets
    a = np.array([31.5 - 32.0, 31.5 - (32.0 - R), 0.0])
    proj = a[1]  # n_src in-plane direction is +y at view 0
    a[2] = -proj * np.tan(w / 2)
    b = np.array([32.0 + rdet - 31.5, 32.0 + Rd - 31.5, g.det_z0 - a[2]])
    cos_t = np.dot(a, b) / np.linalg.norm(a) / np.linalg.norm(b)
    assert abs(o["cos_theta"] - cos_t) < 1e-14
    assert abs(o["sin_half"] - np.sin(np.arccos(cos_t) / 2)) < 1e-9
    cospsi = a[1] / np.hypot(a[0], a[1])
    g_sv = cospsi / np.linalg.norm(a) * np.tan(w)
    g_vd = b[1] / np.linalg.norm(b) ** 3  # -(n_det . b) with n_det = (0, -1, 0)
    assert abs(o["geom"] / (g_sv * g_vd) - 1) < 1e-12
    assert abs(o["rayleigh"] - (1 + cos_t ** 2) / 2) < 1e-15
    assert abs(theta - np.arccos(cos_t)) < 0.02  # the half-voxel offset changes theta little


def test_geometry_detector_rows(oracle):
    """Row r of the detector sits at height det_z0 + r*row_pitch
    (integrate_diffraction.cu:100-109): cos(theta) of each row from hand vectors."""
    g = _geom(n_rows=4)
    R, Rd, w = g.r_src, g.r_det, g.wedge_angle
    col = 5
    rdet = (col - g.n_cols / 2 + 0.5) * g.col_pitch
    a = np.array([31.5 - 32.0, 31.5 - (32.0 - R), 0.0])
    a[2] = -a[1] * np.tan(w / 2)
    for row in range(4):
        o = oracle.scatter_geometry(g, 0, col, row, 31, 31)
        b = np.array([32.0 + rdet - 31.5, 32.0 + Rd - 31.5,
                      g.det_z0 + row * g.row_pitch - a[2]])
        cos_t = np.dot(a, b) / np.linalg.norm(a) / np.linalg.norm(b)
        assert abs(o["cos_theta"] - cos_t) < 1e-14, row
        assert abs(o["det_x"] - (32.0 + rdet)) < 1e-12 and abs(o["det_y"] - (32.0 + Rd)) < 1e-12


def test_forward_rows_layout(oracle):
    """Row r of a multi-row detector (output row v + V*(col + COL*(row + ROW*e))) is the
    one-row detector placed at height det_z0 + r*row_pitch."""
    import dataclasses
    from xrd_tomography_amd.descriptors import Geometry
    from xrd_tomography_amd.synthetic import make_problem
    p1 = make_problem("small")
    g3 = Geometry(**p1.geometry.as_dict())
    g3.n_rows = 3
    N3, C3, A3 = oracle.forward(dataclasses.replace(p1, geometry=g3), nthreads=2)
    V, COL, E = g3.n_views, g3.n_cols, p1.disc.n_energy
    for r in range(3):
        gr = Geometry(**p1.geometry.as_dict())
        gr.det_z0 = p1.geometry.det_z0 + r * p1.geometry.row_pitch
        N1, C1, A1 = oracle.forward(dataclasses.replace(p1, geometry=gr), nthreads=2)
        assert np.array_equal(N3.reshape(E, 3, COL, V)[:, r].reshape(-1), N1)
        assert np.array_equal(C3.reshape(E, 3, COL, V)[:, r].reshape(-1), C1)
        assert np.array_equal(A3.reshape(E, 3, COL, V, -1)[:, r].reshape(A1.shape), A1)


def test_geometry_rotation_symmetry(oracle):
    """A quarter turn of the scanner (V/4 views) maps voxel offset (dx, dy) to (-dy, dx);
    every scalar of the scattering geometry is unchanged."""
    g = _geom()
    rng = np.random.default_rng(4)
    for _ in range(50):
        ix, iy = rng.integers(0, 64, 2)
        col = int(rng.integers(0, g.n_cols))
        v = int(rng.integers(0, 2))
        o1 = oracle.scatter_geometry(g, v, col, 0, int(ix), int(iy))
        o2 = oracle.scatter_geometry(g, v + 2, col, 0, int(63 - iy), int(ix))
        for k in ("cos_theta", "sin_half", "anorm2", "bnorm2", "res2", "geom", "rayleigh"):
            assert abs(o1[k] - o2[k]) <= 1e-9 * max(abs(o1[k]), 1e-300), k


# ---------------------------------------------------------------------------
# forward accumulation: closed forms on one-voxel problems

def _one_voxel_problem(mu2=0.0, flags=0, ifine_value=1.0, diag_spectrum=True, focal=2.0):
    """A single non-air voxel at the centre of a 16x16 grid, optional uniform a2 map,
    energies 20..35 keV, resolution blurred by a large focal spot (sigma_q >> bin widths)."""
    from xrd_tomography_amd.synthetic import make_problem
    p = make_problem("small", flags=flags)
    g = p.geometry
    g.nx = g.ny = 16
    g.center_x = g.center_y = 8.0
    g.n_views, g.n_cols = 2, 3
    g.view_angle_step = np.pi / 2
    g.r_src, g.r_det, g.det_z0, g.col_pitch = 60.0, 40.0, 6.0, 1.0
    g.focal_spot = focal
    d = p.disc
    d.n_energy, d.n_q, d.n_qf, d.n_mat = 16, 400, 1200, 1
    d.e_min, d.e_width = 19.5, 1.0
    d.q_min, d.q_max, d.q_dq0 = 0.05, 4.5, 0.005
    seg = np.zeros((16, 16), np.int32)
    seg[8, 8] = 1
    p.segmentation = seg
    p.a1 = np.zeros((16, 16), np.float32)
    p.a2 = np.full((16, 16), mu2, np.float32)
    p.ifine = np.full((d.n_qf, 1), ifine_value, np.float32)
    E = d.n_energy
    p.spectrum = (np.eye(E) if diag_spectrum else np.ones((E, E))).astype(np.float32)
    return p


def test_forward_gaussian_normalisation(oracle):
    """mu = 0, form factor 1, spectrum diagonal: each count row equals geom*rayleigh
    (the q-resolution Gaussian integrates to 1 on both the fine and the coarse axis)."""
    p = _one_voxel_problem()
    N, C, A = oracle.forward(p, nthreads=2)
    g, d = p.geometry, p.disc
    for v in range(g.n_views):
        for col in range(g.n_cols):
            o = oracle.scatter_geometry(g, v, col, 0, 8, 8)
            expect = o["geom"] * o["rayleigh"]
            for e in range(d.n_energy):
                r = v + g.n_views * (col + g.n_cols * e)
                assert abs(N[r] / expect - 1) < 1e-9, (v, col, e, N[r] / expect)
                assert abs(A[r].sum() / expect - 1) < 1e-5, (v, col, e, A[r].sum() / expect)


def test_forward_beer_lambert(oracle):
    """Uniform a2 = mu: each row is attenuated by exp(-(L_src + L_det)/f2(E)), with the
    two leg integrals from the (separately pinned) line integrals."""
    mu = 0.02
    p0 = _one_voxel_problem(0.0)
    p1 = _one_voxel_problem(mu)
    N0, _, _ = oracle.forward(p0, with_A=False, nthreads=2)
    N1, _, _ = oracle.forward(p1, with_A=False, nthreads=2)
    g, d = p1.geometry, p1.disc
    maps = np.stack([p1.flat("a1"), p1.flat("a2")]).astype(np.float64)
    for v in range(g.n_views):
        for col in range(g.n_cols):
            o = oracle.scatter_geometry(g, v, col, 0, 8, 8)
            ls = oracle.line_integrals(maps, 16, 16, d.ray_step, o["src_x"], o["src_y"], 8.5, 8.5)
            ld = oracle.line_integrals(maps, 16, 16, d.ray_step, o["det_x"], o["det_y"], 8.5, 8.5)
            for e in range(d.n_energy):
                ein = (d.e_min + (e + 0.5) * d.e_width) / d.mc2
                f2 = d.kn0 + d.kn1 * ein + d.kn2 * ein * ein
                r = v + g.n_views * (col + g.n_cols * e)
                assert abs(N1[r] / N0[r] - np.exp(-(ls[1] + ld[1]) / f2)) < 1e-12


def test_forward_coarse_fine_consistency(oracle):
    """A smooth form factor F(q): A . F(q_coarse) ~ N (main.m:214 compares A*Irecon with
    the counts); the residual is the quadrature difference of the two q grids."""
    p = _one_voxel_problem()
    d = p.disc
    from xrd_tomography_amd.descriptors import q_axis
    q, dq, qf, dqf = q_axis(d)
    F = lambda t: 1.0 + 0.5 * np.sin(2.0 * np.asarray(t))  # noqa: E731
    p.ifine = F(qf).reshape(-1, 1).astype(np.float32)
    N, C, A = oracle.forward(p, nthreads=2)
    est = A @ F(q)
    assert np.max(np.abs(est / N - 1)) < 2e-3


def test_forward_view_layout_and_subset(oracle):
    """Rows of a view subset equal the same rows of the all-views run."""
    from xrd_tomography_amd.synthetic import make_problem
    p = make_problem("small", flags=3)
    N, C, A = oracle.forward(p, nthreads=2)
    Ns, Cs, As = oracle.forward(p, views=[2], nthreads=2)
    g, d = p.geometry, p.disc
    rows = 2 + g.n_views * np.arange(g.n_cols * g.n_rows * d.n_energy)
    assert np.array_equal(N[rows], Ns) and np.array_equal(C[rows], Cs)
    assert np.array_equal(A[rows], As)


def test_forward_column_subset(oracle):
    """Rows of a (view, column) subset equal the same rows of the all-columns run, in
    the order of the lists (rows r = vi + nvl*(ci + ncl*(row + ROW*e)))."""
    from xrd_tomography_amd.synthetic import make_problem
    p = make_problem("small", flags=3, n_rows=2)
    g, d = p.geometry, p.disc
    N, C, A = oracle.forward(p, nthreads=2)
    views, cols = [3, 1], [g.n_cols - 1, 0, 2]
    Ns, Cs, As = oracle.forward(p, views=views, cols=cols, nthreads=2)
    sel = lambda X: X.reshape(d.n_energy, g.n_rows, g.n_cols, g.n_views, -1)[:, :, cols][:, :, :, views]  # noqa: E731
    assert np.array_equal(sel(N).reshape(-1), Ns) and np.array_equal(sel(C).reshape(-1), Cs)
    assert np.array_equal(sel(A).reshape(As.shape), As)
    for cv in ([], [-1], [g.n_cols]):
        with pytest.raises(ValueError):
            oracle.forward(p, views=[0], cols=cv)


def test_forward_quirks_only_touch_compton(oracle):
    """Both reference quirks change only the Compton counts (integrate_diffraction.cu:181,
    :259-260 feed only :318-356); N and A are identical."""
    from xrd_tomography_amd.synthetic import make_problem
    pa = make_problem("small", flags=0)
    pb = make_problem("small", flags=3)
    Na, Ca, Aa = oracle.forward(pa, nthreads=2)
    Nb, Cb, Ab = oracle.forward(pb, nthreads=2)
    assert np.array_equal(Na, Nb) and np.array_equal(Aa, Ab)
    assert not np.allclose(Ca, Cb)


# ---------------------------------------------------------------------------
# explicit-matrix operators and MLEM (main.m:141-176)

def test_operators_match_numpy(oracle):
    rng = np.random.default_rng(7)
    A = rng.random((300, 40))
    x, w = rng.random(40), rng.random(300)
    assert np.allclose(oracle.project(A, x), A @ x, rtol=1e-14)
    assert np.allclose(oracle.backproject(A, w), A.T @ w, rtol=1e-13)
    assert np.allclose(oracle.sensitivity(A), A.sum(0), rtol=1e-13)


def test_mlem_fixed_point_and_conservation(oracle):
    rng = np.random.default_rng(8)
    A = rng.random((400, 30)) ** 3
    xs = rng.uniform(1, 5, 30)
    y = A @ xs
    c = np.zeros(400)
    # a consistent noiseless y = A x* is a fixed point (up to the 1e-7 guard)
    x1 = oracle.mlem(A, y, c, xs, 1)
    assert np.max(np.abs(x1 / xs - 1)) < 1e-6
    # from x0 = 10: counts are conserved and the likelihood never decreases
    yn = rng.poisson(y * 3).astype(np.float64)
    x = np.full(30, 10.0)
    ll = oracle.loglik(A, yn, c, x)
    for _ in range(15):
        x = oracle.mlem(A, yn, c, x, 1)
        assert abs((A @ x).sum() / yn.sum() - 1) < 1e-6
        ll2 = oracle.loglik(A, yn, c, x)
        assert ll2 >= ll - 1e-9 * abs(ll)
        ll = ll2
        assert np.all(x >= 0)


def test_mlem_update_rule_numpy(oracle):
    """One update against a direct numpy evaluation of main.m:162-163,176."""
    rng = np.random.default_rng(9)
    A = rng.random((200, 24))
    A[:, 5] = 0.0  # zero sensitivity column: 0/(0+1e-7) -> 0
    y = rng.poisson(20, 200).astype(np.float64)
    c = rng.random(200)
    x = rng.uniform(0, 3, 24)
    ref = x * (A.T @ ((y - 0 * c) / (A @ x + c + 1e-20))) / (A.sum(0) + 1e-7)
    ref[~np.isfinite(ref)] = 0
    ref[ref < 0] = 0
    got = oracle.mlem(A, y, c, x, 1)
    assert np.allclose(got, ref, rtol=1e-13, atol=0)


# ---------------------------------------------------------------------------
# coefficient model: vector-Jacobian product pinned by FP64 central differences

def _nll_diff(yp, ym, y, c):
    """NLL(yp) - NLL(ym) for the Poisson NLL sum(f - y log f), f = yhat + c, formed row
    by row with log1p so the difference keeps its precision."""
    d = yp - ym
    return float(np.sum(d - y * np.log1p(d / (ym + c))))


def _coeff_case():
    """The differentiation workload's central-difference size: its 256x256 phantom
    generator and scanner at 32x32 voxels (16 mm, the same 512 mm field), 30 views,
    61 energies, 128 q-bins, 16 Gaussian-bump coefficients per voxel and the full
    61-energy attenuation map; 8 detector columns keep the FP64 forward near a second."""
    from xrd_tomography_amd.synthetic import make_coeff_problem, make_problem
    p = make_problem("fan256", nx=32, voxel_mm=16.0, n_views=30, n_cols=8)
    cp = make_coeff_problem(p, K=16, full_mu=True)
    assert cp.K == 16 and cp.B == 61 == p.disc.n_energy and p.disc.n_q == 128
    return cp


def test_mu_basis_equals_per_energy_map(oracle):
    """mu(x, E) = sum_b m_b(x) beta_b(E): the reference's two-basis form (B = 2) and the
    per-energy map it defines (B = NRG, beta = identity) give the same counts, since the
    line integral is linear in the map."""
    from xrd_tomography_amd.synthetic import make_coeff_problem, make_problem
    p = make_problem("small", nx=32, n_views=30)
    c2 = make_coeff_problem(p, K=6, full_mu=False)
    cf = make_coeff_problem(p, K=6, full_mu=True)
    full = c2.beta.astype(np.float64).T @ c2.maps.astype(np.float64)   # [NRG, npix]
    y2 = oracle.forward_coeffs(c2)
    yf = oracle.forward_coeffs(cf, maps=full)
    assert np.array_equal(cf.beta, np.eye(p.disc.n_energy, dtype=np.float32))
    assert np.max(np.abs(yf - y2) / np.maximum(np.abs(y2), 1e-300)) <= 1e-12


def test_vjp_coeffs_dot_product(oracle):
    """The model is linear in the coefficients: <J x, u> = <x, J' u>."""
    cp = _coeff_case()
    rng = np.random.default_rng(10)
    x = rng.normal(size=cp.coef.shape) * (cp.coef != 0)
    u = rng.normal(size=cp.base.n_rows)
    jx = oracle.forward_coeffs(cp, coef=x)
    jtu, _ = oracle.vjp_coeffs(cp, u, want_maps=False)
    a, b = np.dot(jx, u), np.sum(x * jtu)
    assert abs(a - b) <= 1e-12 * max(abs(a), abs(b))


def test_vjp_central_differences(oracle):
    """Poisson NLL gradient w.r.t. coefficients and per-energy attenuation maps against
    FP64 central differences (step 1e-6 relative), within 1e-6 of each element's
    J'|w| scale."""
    cp = _coeff_case()
    rng = np.random.default_rng(11)
    y0 = oracle.forward_coeffs(cp)
    c = 0.2 * y0.mean() * np.ones_like(y0)
    y = rng.poisson((y0 + c) / y0.mean() * 50) * y0.mean() / 50
    w = 1.0 - y / (y0 + c)
    gc, gm = oracle.vjp_coeffs(cp, w)
    # Each element's rounding scale is J'|w|, the sum of its terms' magnitudes: the
    # random-sign upstream gradient cancels some elements well below it, and the FP64
    # count differences at the 1e-6 step carry noise relative to that scale.
    sc, sm = oracle.vjp_coeffs(cp, np.abs(w))
    live = np.nonzero(cp.coef[:, 0])[0]
    for pix in live[np.argsort(-np.abs(gc[live, 0]))[:3]]:
        for k in (0, cp.K - 1):
            h = 1e-6 * abs(float(cp.coef[pix, k]))
            cpp, cpm = cp.coef.astype(np.float64).copy(), cp.coef.astype(np.float64).copy()
            cpp[pix, k] += h
            cpm[pix, k] -= h
            fd = _nll_diff(oracle.forward_coeffs(cp, coef=cpp),
                           oracle.forward_coeffs(cp, coef=cpm), y, c) / (2 * h)
            assert abs(fd - gc[pix, k]) <= 1e-6 * abs(sc[pix, k]), (pix, k, fd, gc[pix, k], sc[pix, k])
    npix = cp.maps.shape[1]
    # map pixels on the rays: take the largest-gradient pixels of two energies
    for b in (1, cp.B - 2):
        for pix in np.argsort(-np.abs(gm[b]))[:2]:
            h = 1e-6 * max(abs(float(cp.maps[b, pix])), 1e-3)
            mp, mm = cp.maps.astype(np.float64).copy(), cp.maps.astype(np.float64).copy()
            mp[b, pix] += h
            mm[b, pix] -= h
            fd = _nll_diff(oracle.forward_coeffs(cp, maps=mp),
                           oracle.forward_coeffs(cp, maps=mm), y, c) / (2 * h)
            assert abs(fd - gm[b, pix]) <= 1e-6 * abs(sm[b, pix]), (b, pix, fd, gm[b, pix], sm[b, pix])
    assert npix == cp.base.geometry.nx * cp.base.geometry.ny


# ---------------------------------------------------------------------------
# volume slices: 3-D attenuation legs (trilinear), pinned by the z-invariant limit,
# a numpy recomputation of the 3-D leg and the one-slice identity

def _z_invariant_volume(p, nz):
    import dataclasses
    rep = lambda a: np.repeat(a[:, :, None], nz, axis=2)  # noqa: E731
    return dataclasses.replace(p, segmentation=rep(p.segmentation), a1=rep(p.a1), a2=rep(p.a2))


def test_volume_constant_in_z_is_the_2d_model(oracle):
    """Identical slices: every leg stays inside the volume, the trilinear blend of equal
    slices is the bilinear value, and the slice's counts are the 2-D model's."""
    from xrd_tomography_amd.synthetic import make_problem
    p = make_problem("small")
    v = _z_invariant_volume(p, 12)
    N0, C0, A0 = oracle.forward(p, nthreads=2)
    N1, C1, A1 = oracle.forward_slice(v, 12, 5, nthreads=2)
    for a, b in ((N1, N0), (C1, C0), (A1, A0)):
        assert np.max(np.abs(a - b)) <= 1e-12 * np.abs(b).max()
    N2, C2, A2 = oracle.forward_slice(_z_invariant_volume(p, 1), 1, 0, nthreads=2)
    assert np.array_equal(N2, N0) and np.array_equal(C2, C0) and np.array_equal(A2, A0)


def test_line_integrals_3d_numpy(oracle):
    """3-D legs recomputed in numpy: the attenuator's 2-D samples from the ROI entry,
    z linear in the 2-D distance to the end point, trilinear with a zero border."""
    rng = np.random.default_rng(31)
    nx, ny, nz, step = 17, 13, 6, 0.25
    vol = rng.uniform(0.0, 1.0, (2, nz, ny, nx))        # [b, k, j, i]: x fastest in memory
    pad = np.zeros((2, nz + 2, ny + 2, nx + 2))
    pad[:, 1:-1, 1:-1, 1:-1] = vol

    def tri(px, py, pz):
        xt, yt, zt = px - 0.5, py - 0.5, pz - 0.5
        i, j, k = np.floor(xt).astype(int), np.floor(yt).astype(int), np.floor(zt).astype(int)
        ax, ay, az = xt - i, yt - j, zt - k
        out = 0.0
        for di, wx in ((0, 1 - ax), (1, ax)):
            for dj, wy in ((0, 1 - ay), (1, ay)):
                for dk, wz in ((0, 1 - az), (1, az)):
                    ii = np.clip(i + di, -1, nx) + 1
                    jj = np.clip(j + dj, -1, ny) + 1
                    kk = np.clip(k + dk, -1, nz) + 1
                    out = out + wx * wy * wz * pad[:, kk, jj, ii]
        return out
    for _ in range(40):
        x2, y2 = rng.uniform(0.5, nx - 0.5), rng.uniform(0.5, ny - 0.5)
        ang = rng.uniform(0, 2 * np.pi)
        x1, y1 = x2 + 60 * np.cos(ang), y2 + 60 * np.sin(ang)
        z1, z2 = rng.uniform(-1, nz + 1), rng.uniform(0.5, nz - 0.5)
        got = oracle.line_integrals_3d(vol, nx, ny, nz, step, x1, y1, z1, x2, y2, z2)
        side, ex, ey = oracle.get_intersection(x1, y1, x2, y2, nx, ny)
        L = np.hypot(x2 - ex, y2 - ey)
        H = np.hypot(x2 - x1, y2 - y1)
        q = np.arange(int(L / step) + 2)
        q = q[q * step <= L]
        t = q * step
        px, py = ex + (x2 - ex) / L * t, ey + (y2 - ey) / L * t
        pz = z2 + (z1 - z2) * ((L - t) / H)
        ref = step * tri(px, py, pz).sum(axis=1)
        assert np.allclose(got, ref, rtol=1e-12, atol=1e-14), (got, ref)


# ---------------------------------------------------------------------------
# CT transmission channel: chord by bisection, Beer-Lambert recomputed in numpy,
# the zero-attenuation closed form and FP64 central differences for the VJP

def _ct_case(full_mu=False):
    from xrd_tomography_amd.synthetic import make_coeff_problem, make_problem
    p = make_problem("small")
    cp = make_coeff_problem(p, K=2, full_mu=full_mu)
    return p, cp.maps.astype(np.float64), cp.beta.astype(np.float64)


def _inside(x, y, nx, ny):
    return (x >= 0) & (x <= nx) & (y >= 0) & (y <= ny)


def test_ct_chord_by_bisection(oracle):
    """Entry and exit of random segments through the grid, located by dense sampling
    of the segment and bisection of the inside/outside boundary."""
    rng = np.random.default_rng(21)
    nx, ny = 37, 23
    n_hit = 0
    for _ in range(300):
        sx, sy = rng.uniform(-60, 100, 2)
        dx, dy = rng.uniform(-60, 100, 2)
        hit, ex, ey, xx, xy = oracle.ct_chord(sx, sy, dx, dy, nx, ny)
        t = np.linspace(0.0, 1.0, 20001)
        ins = _inside(sx + t * (dx - sx), sy + t * (dy - sy), nx, ny)
        if not ins.any():
            assert not hit or math.hypot(xx - ex, xy - ey) < 1e-3
            continue
        assert hit
        n_hit += 1
        idx = np.nonzero(ins)[0]

        def edge(lo, hi, want_inside_at_hi):
            for _ in range(80):
                mid = 0.5 * (lo + hi)
                if bool(_inside(sx + mid * (dx - sx), sy + mid * (dy - sy), nx, ny)) == want_inside_at_hi:
                    hi = mid
                else:
                    lo = mid
            return hi
        t0 = 0.0 if idx[0] == 0 else edge(t[idx[0] - 1], t[idx[0]], True)
        t1 = 1.0 if idx[-1] == len(t) - 1 else edge(t[idx[-1] + 1], t[idx[-1]], True)
        assert abs(ex - (sx + t0 * (dx - sx))) < 1e-9 and abs(ey - (sy + t0 * (dy - sy))) < 1e-9
        assert abs(xx - (sx + t1 * (dx - sx))) < 1e-9 and abs(xy - (sy + t1 * (dy - sy))) < 1e-9
    assert n_hit > 50


def _np_bilinear(m, nx, ny, x, y):
    """Zero-border bilinear of m [nx*ny] (x fastest) at texel coordinates (x, y)."""
    xt, yt = x - 0.5, y - 0.5
    i, j = np.floor(xt).astype(int), np.floor(yt).astype(int)
    ax, ay = xt - i, yt - j
    M = np.zeros((ny + 4, nx + 4))
    M[2:ny + 2, 2:nx + 2] = m.reshape(ny, nx)

    def at(ii, jj):
        return M[np.clip(jj, -2, ny + 1) + 2, np.clip(ii, -2, nx + 1) + 2]
    return ((1 - ax) * (1 - ay) * at(i, j) + ax * (1 - ay) * at(i + 1, j)
            + (1 - ax) * ay * at(i, j + 1) + ax * ay * at(i + 1, j + 1))


def test_ct_forward_beer_lambert_numpy(oracle):
    """T recomputed in numpy: the chord sampled every ray_step from its entry (while
    q*step <= chord length), numpy bilinear line integrals, exp(-beta L) summed over
    each detector bin's source-energy window."""
    p, maps, beta = _ct_case()
    g, d = p.geometry, p.disc
    T = oracle.forward_ct(p, maps, beta)
    V, COL, E, DN = g.n_views, g.n_cols, d.n_energy, d.energy_halfwidth
    RF = p.flat("spectrum").astype(np.float64).reshape(E, E)  # RF[ed, e] = spectrum[e + E*ed]
    ref = np.zeros(V * COL * E)
    for v in range(V):
        for c in range(COL):
            sg = oracle.scatter_geometry(g, v, c, 0, 0, 0)
            hit, ex, ey, xx, xy = oracle.ct_chord(sg["src_x"], sg["src_y"], sg["det_x"],
                                                  sg["det_y"], g.nx, g.ny)
            assert hit
            Lc = math.hypot(xx - ex, xy - ey)
            q = np.arange(int(Lc / d.ray_step) + 2)
            q = q[q * d.ray_step <= Lc]
            px = ex + (xx - ex) / Lc * (q * d.ray_step)
            py = ey + (xy - ey) / Lc * (q * d.ray_step)
            L = np.array([d.ray_step * _np_bilinear(maps[b], g.nx, g.ny, px, py).sum()
                          for b in range(maps.shape[0])])
            P = np.exp(-(beta * L[:, None]).sum(0))
            for ed in range(E):
                lo = max(0, ed - DN)
                hi = lo + 2 * DN
                if hi >= E:
                    hi, lo = E - 1, E - 1 - 2 * DN
                ref[v + V * (c + COL * ed)] = np.dot(RF[ed, lo:hi + 1], P[lo:hi + 1])
    assert np.allclose(T, ref, rtol=1e-12, atol=0)
    # no attenuation: T is the window sum of the response matrix
    T0 = oracle.forward_ct(p, np.zeros_like(maps), beta)
    for ed in (0, E // 2, E - 1):
        lo = min(max(0, ed - DN), E - 1 - 2 * DN)
        assert np.allclose(T0[V * COL * ed:V * COL * (ed + 1)], RF[ed, lo:lo + 2 * DN + 1].sum(),
                           rtol=1e-14)


@pytest.mark.parametrize("full_mu", [False, True])
def test_ct_vjp_central_differences(oracle, full_mu):
    """<w, T(maps)> gradient against FP64 central differences, and along a random
    direction (the directional derivative must equal <grad, direction>)."""
    p, maps, beta = _ct_case(full_mu)
    rng = np.random.default_rng(22)
    w = rng.normal(size=p.geometry.n_views * p.geometry.n_cols * p.disc.n_energy)
    gm = oracle.vjp_ct(p, maps, beta, w)
    # beta = identity makes each map's effect small: a larger step keeps the
    # difference above rounding (the curvature error stays below 1e-9)
    hr, floor_ = (1e-4, 1e-2) if full_mu else (1e-6, 1e-3)
    for b in (0, maps.shape[0] - 1):
        for pix in np.argsort(-np.abs(gm[b]))[:3]:
            h = hr * max(abs(float(maps[b, pix])), floor_)
            mp, mm = maps.copy(), maps.copy()
            mp[b, pix] += h
            mm[b, pix] -= h
            fd = np.dot(w, oracle.forward_ct(p, mp, beta) - oracle.forward_ct(p, mm, beta)) / (2 * h)
            assert abs(fd - gm[b, pix]) <= 1e-6 * abs(gm[b, pix]), (b, pix, fd, gm[b, pix])
    dm = rng.normal(size=maps.shape) * 1e-3 * np.abs(maps).max()
    h = 1e-4
    fd = np.dot(w, oracle.forward_ct(p, maps + h * dm, beta) - oracle.forward_ct(p, maps - h * dm, beta)) / (2 * h)
    assert abs(fd - np.sum(gm * dm)) <= 1e-7 * abs(fd)


# ---------------------------------------------------------------------------
# committed FP64 fixtures (tests/golden/make_golden.py)

@pytest.mark.parametrize("name,flags", [("reference", 3), ("small", 0)])
def test_oracle_reproduces_golden(oracle, name, flags):
    import os
    from xrd_tomography_amd.synthetic import make_problem
    gold = np.load(os.path.join(os.path.dirname(__file__), "golden", f"{name}_flags{flags}.npz"))
    p = make_problem(name, flags=flags)
    assert np.array_equal(p.flat("spectrum"), gold["spectrum"])      # generator unchanged
    assert np.array_equal(p.flat("segmentation"), gold["segmentation"])
    N, C, A = oracle.forward(p, nthreads=3)
    assert np.allclose(N, gold["N"], rtol=1e-12, atol=0)
    assert np.allclose(C, gold["C"], rtol=1e-12, atol=0)
    assert np.allclose(A[::p.geometry.n_views], gold["A_view0"], rtol=1e-12, atol=0)
    assert np.allclose(A.sum(0), gold["A_colsum"], rtol=1e-12, atol=0)


# ---------------------------------------------------------------------------
# q-resolution Gaussian (integrate_diffraction.cu:284-288, :300-315) and the Compton
# term (:318-356), pinned by moments, first-order error propagation, Monte Carlo and
# a ramp S2 table -- none of which reuses the oracle's own formulas for wvt, sigma^2,
# the S/D/V vectors or the channel match.

def _view0_points(g, col):
    """View 0, detector column `col`, the voxel centred at (8.5, 8.5) of the one-voxel
    problem: source, pixel and voxel as 3-vectors built from the scanner description
    (source below the grid on -y, detector panel on +y at det_z0, the voxel at the
    wedge's mid-height), the focal-spot normal, the detector normal and the
    illuminated voxel height."""
    R, Rd, w = g.r_src, g.r_det, g.wedge_angle
    s = np.array([g.center_x, g.center_y - R, 0.0])
    rdet = (col - g.n_cols / 2 + 0.5) * g.col_pitch
    d = np.array([g.center_x + rdet, g.center_y + Rd, g.det_z0])
    along = 8.5 - s[1]  # in-plane distance from the source along the central ray (+y)
    v = np.array([8.5, 8.5, -along * np.tan(w / 2)])
    n_src = np.array([0.0, np.sin(g.target_angle), -np.cos(g.target_angle)])
    n_det = np.array([0.0, -1.0, 0.0])
    return s, d, v, n_src, n_det, along * np.tan(w)


def _q_exact(E, s, d, v, hbarc):
    """Momentum transfer 2E sin(theta/2)/hbar c with theta from atan2(|a x b|, a.b)."""
    a, b = v - s, d - v
    cr = np.cross(a, b)
    theta = np.arctan2(np.sqrt(np.sum(cr * cr, axis=-1)), np.sum(a * b, axis=-1))
    return 2.0 * E / hbarc * np.sin(theta / 2.0)


def _q_variance_first_order(g, disc, E, col):
    """Variance of q when the energy is uniform over its bin, the focal point uniform
    over a square of side focal_spot normal to the anode, the scattering point uniform
    over the voxel box (voxel_size x voxel_size x illuminated height) and the
    detection point uniform over a square of area col_pitch*row_pitch in the detector
    plane: the sum of (dq/du)^2 var(u), each derivative a central difference of the
    exact q."""
    s, d, v, n_src, _, hvox = _view0_points(g, col)
    q0 = _q_exact(E, s, d, v, disc.hbarc)
    var = (q0 / E) ** 2 * disc.e_width ** 2 / 12.0
    ex, ey, ez = np.eye(3)
    adet = g.col_pitch * g.row_pitch
    terms = [("s", ex, g.focal_spot ** 2), ("s", np.cross(n_src, ex), g.focal_spot ** 2),
             ("d", ex, adet), ("d", ez, adet),
             ("v", ex, g.voxel_size ** 2), ("v", ey, g.voxel_size ** 2), ("v", ez, hvox ** 2)]
    h = 1e-4
    for which, u, w2 in terms:
        def at(t):
            pts = {"s": s.copy(), "d": d.copy(), "v": v.copy()}
            pts[which] += t * u
            return _q_exact(E, pts["s"], pts["d"], pts["v"], disc.hbarc)
        dq = (at(h) - at(-h)) / (2.0 * h)
        var += dq * dq * w2 / 12.0
    return var


def _fine_moments(oracle, p):
    """N with F(q) = 1, q and q^2 on the fine axis, divided: the mean and variance of
    the q-resolution Gaussian of each (energy, column) row of view 0."""
    from xrd_tomography_amd.descriptors import q_axis
    g, d = p.geometry, p.disc
    qf = np.asarray(q_axis(d)[2])
    Ns = []
    for F in (np.ones_like(qf), qf, qf * qf):
        p.ifine = F.reshape(-1, 1)
        N, _, _ = oracle.forward(p, views=[0], with_A=False, nthreads=2)
        Ns.append(N.reshape(d.n_energy, g.n_cols))
    mean = Ns[1] / Ns[0]
    return mean, Ns[2] / Ns[0] - mean * mean


def pin_resolution_moments(oracle):
    """Form-factor moments on the one-voxel problem (mu = 0, diagonal spectrum, so
    row e holds the single source bin e): with F(q) = q the counts divided by the
    F = 1 counts are the Gaussian's centre, 2E sin(theta/2)/hbar c (wvt,
    integrate_diffraction.cu:284); with F(q) = q^2 they are wvt^2 + sigma^2, and
    sigma^2 (:287-288) equals the first-order propagation of the energy bin, focal
    spot, voxel and pixel extents (:220-237)."""
    p = _one_voxel_problem()
    g, d = p.geometry, p.disc
    mean, var = _fine_moments(oracle, p)
    for col in range(g.n_cols):
        s, dpix, v, _, _, _ = _view0_points(g, col)
        for e in range(d.n_energy):
            E = d.e_min + (e + 0.5) * d.e_width
            wvt = _q_exact(E, s, dpix, v, d.hbarc)
            assert abs(mean[e, col] / wvt - 1) < 1e-11, ("centre", col, e, mean[e, col], wvt)
            s2 = _q_variance_first_order(g, d, E, col)
            assert abs(var[e, col] / s2 - 1) < 1e-7, ("width", col, e, var[e, col], s2)


def pin_resolution_monte_carlo(oracle):
    """The same Gaussian against positions and energies drawn uniformly over the
    focal spot, voxel, pixel and energy bin, with q computed exactly per draw: the
    sample mean and variance of q agree with the Gaussian's centre and sigma^2 to
    the sampling error plus the second-order terms the first-order model leaves
    out (under 1%)."""
    p = _one_voxel_problem()
    g, d = p.geometry, p.disc
    mean, var = _fine_moments(oracle, p)
    rng = np.random.default_rng(1)
    n = 400_000
    ex, ey, ez = np.eye(3)
    for col, e in ((0, 3), (2, 12)):
        s, dpix, v, n_src, _, hvox = _view0_points(g, col)
        E0 = d.e_min + (e + 0.5) * d.e_width
        U = lambda w: rng.uniform(-0.5, 0.5, (n, 1)) * w  # noqa: E731
        pa = np.sqrt(g.col_pitch * g.row_pitch)
        E = E0 + U(d.e_width)[:, 0]
        S = s + U(g.focal_spot) * ex + U(g.focal_spot) * np.cross(n_src, ex)
        D = dpix + U(pa) * ex + U(pa) * ez
        V = v + U(g.voxel_size) * ex + U(g.voxel_size) * ey + U(hvox) * ez
        q = _q_exact(E, S, D, V, d.hbarc)
        assert abs(q.mean() / mean[e, col] - 1) < 1e-2, ("MC centre", col, e)
        assert abs(q.var() / var[e, col] - 1) < 2e-2, ("MC width", col, e, q.var(), var[e, col])


def _compton_case(flags):
    """One voxel in uniform (a1, a2) = (0.02*2^-12, 0.02) per voxel and a 16 x 16 S2
    table holding its own flat index, S2[nq + 16 na] = nq + 16 na (all < 256, so the
    uint8 quantisation of integrate_diffraction.cu:66-74 keeps them)."""
    p = _one_voxel_problem(flags=flags)
    mu2 = 0.02
    p.a1 = np.full((16, 16), mu2 * 2.0 ** -12, np.float32)
    p.a2 = np.full((16, 16), mu2, np.float32)
    p.disc.lut_nq = p.disc.lut_na = 16
    p.lut = (-9.0, -7.0, -15.0, -7.0)
    p.s2_table = (np.arange(16)[:, None] + 16 * np.arange(16)[None, :]).astype(np.float64)
    return p


def pin_compton_closed_form(oracle):
    """Compton counts of the one-voxel problem with a diagonal spectrum (only source
    bin e reaches detector bin e) in closed form (integrate_diffraction.cu:318-356):
      C_e = m_e Pin(E) Pout(E') g er^2 (er + 1/er - 1 + cos^2 theta)/2 a2 S2[nq + NQ na]
    with er = 1 + (E/mc^2)(1 - cos theta), E' = E/er, m_e = (E2 - er E1)/(E2 - E1) the
    part of the shifted bin [E1/er, E2/er] inside [E1, E2], Pin from the source leg
    and Pout at E' from the detector leg (the legs swap under XRD_QUIRK_SWAPPED_LEGS).
    Dividing C_e by everything but the table value must give an integer, which is the
    table's own index: it must decode to nq = floor((log2(w sqrt(er + w^2)/er) -
    q_lo)/dq) with w = (E/mc^2) sin(theta/2) (:331-339) and na = floor((log2|a1/a2| -
    a_lo)/da) (:186-189), dq and da the table range over 16.  With the stale-voxel
    quirk the only voxel sees (a1, a2) = 0: C = 0."""
    for flags in (0, 2):
        p = _compton_case(flags)
        g, d = p.geometry, p.disc
        _, C, _ = oracle.forward(p, with_A=False, nthreads=2)
        C = C.reshape(d.n_energy, g.n_cols, g.n_views)
        maps = np.stack([p.flat("a1"), p.flat("a2")]).astype(np.float64)
        a2 = float(p.a2[0, 0])
        qlo, qhi, alo, ahi = p.lut
        na = int(np.floor((np.log2(float(p.a1[0, 0]) / a2) - alo) / ((ahi - alo) / 16)))
        f2 = lambda t: d.kn0 + d.kn1 * t + d.kn2 * t * t  # noqa: E731
        seen = set()
        for v in range(g.n_views):
            for col in range(g.n_cols):
                o = oracle.scatter_geometry(g, v, col, 0, 8, 8)
                ls = oracle.line_integrals(maps, 16, 16, d.ray_step, o["src_x"], o["src_y"], 8.5, 8.5)
                ld = oracle.line_integrals(maps, 16, 16, d.ray_step, o["det_x"], o["det_y"], 8.5, 8.5)
                if flags & 2:
                    ls, ld = ld, ls
                sh, cth = o["sin_half"], o["cos_theta"]
                for e in range(d.n_energy):
                    E1 = d.e_min + e * d.e_width
                    E2 = E1 + d.e_width
                    ein = 0.5 * (E1 + E2) / d.mc2
                    er = 1.0 + ein * (1.0 - cth)
                    eout = ein / er
                    pin = np.exp(-(ls[0] / ein ** 3 + ls[1] / f2(ein)))
                    pout = np.exp(-(ld[0] / eout ** 3 + ld[1] / f2(eout)))
                    m = (E2 - er * E1) / (E2 - E1)
                    cf = er * er * (er + 1.0 / er - 1.0 + cth * cth) / 2.0
                    val = C[e, col, v] / (m * pin * pout * o["geom"] * cf * a2)
                    w = ein * sh
                    nq = int(np.floor((np.log2(w * np.sqrt(er + w * w) / er) - qlo) / ((qhi - qlo) / 16)))
                    assert abs(val - round(val)) < 1e-9 * val, ("not a table value", v, col, e, val)
                    assert int(round(val)) == nq + 16 * na, ("LUT index", v, col, e, val, nq, na)
                    seen.add(nq)
        assert len(seen) >= 4  # the rows exercise several q entries of the table
    _, C, _ = oracle.forward(_compton_case(1), with_A=False, nthreads=2)
    assert np.all(C == 0.0)


def test_resolution_moments(oracle):
    pin_resolution_moments(oracle)


def test_resolution_monte_carlo(oracle):
    pin_resolution_monte_carlo(oracle)


def test_compton_closed_form(oracle):
    pin_compton_closed_form(oracle)


# Plausible mistakes, planted one at a time into a copy of the oracle's source and
# compiled next to the test: each must be caught by the pin named with it.
PLANTED = [
    ("S2 index transposed", "s2u[nq + LQ * na]", "s2u[na + LA * nq]", "compton"),
    ("Compton Pout at E instead of E'", "const double g1 = eout * eout * eout, g2 = xo_f2(d, eout);",
     "const double g1 = ein * ein * ein, g2 = xo_f2(d, ein);", "compton"),
    ("LUT material index from a2/a1", "log2(fabs(vx->a1c / vx->a2c))",
     "log2(fabs(vx->a2c / vx->a1c))", "compton"),
    ("angular factor (E'/E)^2 instead of the reference's (E/E')^2",
     "er * er * (er + 1.0 / er - 1.0 + cth * cth) / 2.0",
     "(er + 1.0 / er - 1.0 + cth * cth) / 2.0 / (er * er)", "compton"),
    ("channel match over the unshifted bin", "const double Eo1 = es1 / eratio, Eo2 = es2 / eratio;",
     "const double Eo1 = es1, Eo2 = es2;", "compton"),
    ("wvt = E sin(theta/2)/hbar c", "*wvt = 2.0 * Es / d->hbarc * s;",
     "*wvt = Es / d->hbarc * s;", "moments"),
    ("energy-bin term without /12", "d->e_width * d->e_width / 12.0", "d->e_width * d->e_width",
     "moments"),
    ("geometric term with 2E/hbar c", "const double k2 = Es / d->hbarc;",
     "const double k2 = 2.0 * Es / d->hbarc;", "moments"),
    ("focal-spot term without /12", "g->focal_spot * g->focal_spot / 12.0",
     "g->focal_spot * g->focal_spot", "moments"),
    ("pixel term without the normal component removed",
     "(Dx * Dx + Dy * Dy + Dz * Dz - Dn * Dn) * adet / 12.0", "(Dx * Dx + Dy * Dy + Dz * Dz) * adet / 12.0",
     "moments"),
    ("voxel height from the half wedge", "const double hvox = proj * tw;",
     "const double hvox = proj * tw2;", "moments"),
]
PINS = {"compton": pin_compton_closed_form, "moments": pin_resolution_moments}


def _planted_oracle(tmp_path, old, new):
    import os
    import subprocess
    here = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
    src = open(os.path.join(here, "xrd_oracle.c")).read()
    assert src.count(old) == 1, old
    c = tmp_path / "planted.c"
    c.write_text(src.replace(old, new))
    so = tmp_path / "libplanted.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
                    "-fno-fast-math", "-std=c11", "-I", here, "-o", str(so), str(c), "-lm"],
                   check=True, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    return str(so)


@pytest.mark.parametrize("label,old,new,pin", PLANTED, ids=[m[0] for m in PLANTED])
def test_pins_catch_planted_mistake(oracle, tmp_path, label, old, new, pin):
    so = _planted_oracle(tmp_path, old, new)
    with oracle.using_library(so):
        with pytest.raises(AssertionError):
            PINS[pin](oracle)
    PINS[pin](oracle)  # the real oracle still passes after the swap back
