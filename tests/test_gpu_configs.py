"""GPU parity at the full workload sizes, against the FP64 oracle.

* fan256 (256x256 voxels at 2 mm, 8 materials, 360 views, 256 pixels x 61
  energies, 128 q-bins): the whole forward on the GPU; N, C and A rows of 2 views x
  8 columns against the oracle at the forward bars of test_gpu_forward (counts
  <= 1e-5 per row, A <= 1e-5 per row); texel lists of fan256 source and detector
  legs bit-exact; 5 MLEM updates on the rows of 36 of the 360 views <= 2e-5.
* vjp256 sub-phantom (32x32 voxels, 30 views, 61 energies, K = 16 coefficients,
  B = 61 per-energy maps): counts, coefficient and map gradients against the
  oracle; dot-product test <J x, u> = <x, J' u> <= 1e-6 on the 256x256 geometry at
  12 of the 360 views.
* vol192 (slice 16 of 192x192x32 voxels, 720 views, 256 pixels x 61 energies, 10
  materials): N, C and A rows of 2 views x 8 columns against the oracle, with FP32
  tables and with BF16 tables decoded by the kernels.
* batch48 (48 phantoms of 64^2..256^2 voxels, 90..360 views, 2..10 materials, BF24
  storage): every batched x equals its own single-problem plan's x, and the 4
  smallest match the FP64 oracle MLEM on the stored matrix (2e-5).
"""
import numpy as np
import pytest

from test_gpu_forward import _compare

pytestmark = pytest.mark.gpu

FAN256_VIEWS = [0, 97]
FAN256_COLS = [0, 31, 64, 127, 128, 190, 233, 255]


def _select(prob, X, views, cols, width=None):
    """Rows of a GPU output [E, ROW, COL, V(, width)] for the listed views and columns,
    in the oracle's order r = vi + nvl*(ci + ncl*(row + ROW*e))."""
    import torch
    g, d = prob.geometry, prob.disc
    shape = (d.n_energy, g.n_rows, g.n_cols, g.n_views) + ((width,) if width else ())
    X = X.reshape(shape)
    vi = torch.tensor(views, device=X.device)
    ci = torch.tensor(cols, device=X.device)
    Y = X.index_select(2, ci).index_select(3, vi)
    out = Y.double().cpu().numpy()
    return out.reshape(-1, width) if width else out.reshape(-1)


@pytest.fixture(scope="module")
def fan256(xrd):
    import torch
    from xrd_tomography_amd.synthetic import make_problem
    prob = make_problem("fan256")
    print("\n[fan256] building N, C, A on the GPU", flush=True)
    N, C, A = xrd.forward_problem(prob)
    torch.cuda.synchronize()
    yield prob, N, C, A
    del N, C, A
    torch.cuda.empty_cache()


def test_fan256_forward_rows_vs_oracle(fan256, oracle):
    prob, N, C, A = fan256
    d = prob.disc
    Ng = _select(prob, N, FAN256_VIEWS, FAN256_COLS)
    Cg = _select(prob, C, FAN256_VIEWS, FAN256_COLS)
    Ag = _select(prob, A, FAN256_VIEWS, FAN256_COLS, d.n_mat * d.n_q)
    No, Co, Ao = oracle.forward(prob, views=FAN256_VIEWS, cols=FAN256_COLS)
    _compare("fan256 2 views x 8 columns", Ng, Cg, Ag, No, Co, Ao)


def test_fan256_texel_lists_bit_exact(fan256, xrd, oracle):
    """Source legs (focal spot -> voxel) and detector legs (pixel -> voxel) of the
    fan256 scanner: entry side, entry point, sample count and every texel index."""
    import torch
    prob = fan256[0]
    g, d = prob.geometry, prob.disc
    rng = np.random.default_rng(8)
    ix, iy = np.nonzero(prob.segmentation > 0)
    segs = []
    for _ in range(1500):
        k = int(rng.integers(0, ix.size))
        v, col = int(rng.integers(0, g.n_views)), int(rng.integers(0, g.n_cols))
        o = oracle.scatter_geometry(g, v, col, 0, int(ix[k]), int(iy[k]))
        vx, vy = ix[k] + 0.5, iy[k] + 0.5
        segs.append([o["src_x"], o["src_y"], vx, vy])
        segs.append([o["det_x"], o["det_y"], vx, vy])
    segs = np.asarray(segs, np.float64)
    maxs = int(np.ceil(np.hypot(g.nx, g.ny) / d.ray_step)) + 2
    side, entry, count, tex, _ = xrd.trace_rays(torch.from_numpy(segs).cuda(), g.nx, g.ny,
                                                d.ray_step, max_samples=maxs)
    torch.cuda.synchronize()
    side, entry, count, tex = (t.cpu().numpy() for t in (side, entry, count, tex))
    mism = 0
    for i, (a, b, c, e) in enumerate(segs):
        s, ex, ey, n, t = oracle.trace_ray(a, b, c, e, g.nx, g.ny, d.ray_step, maxs)
        ok = (s == side[i] and n == count[i] and np.array_equal(t, tex[i]) and
              np.float64(ex).tobytes() == entry[i, 0].tobytes() and
              np.float64(ey).tobytes() == entry[i, 1].tobytes())
        mism += not ok
    print(f"\n[fan256 trace] {len(segs)} rays, mean {count.mean():.0f} samples, "
          f"{mism} mismatches")
    assert mism == 0 and count.min() > 0


def test_fan256_mlem_view_subset_vs_oracle(fan256, xrd, oracle):
    """5 fused MLEM updates on every row of 36 evenly spaced views of the fan256 matrix,
    scaled to photons and with Poisson counts as in the benchmark."""
    import torch
    prob, N, C, A = fan256
    g = prob.geometry
    nrow, ncol = prob.n_rows, prob.n_cols
    views = np.linspace(0, g.n_views - 1, 36).round().astype(np.int64)
    rows = (views[None, :] + g.n_views * np.arange(nrow // g.n_views)[:, None]).reshape(-1)
    rt = torch.from_numpy(rows).cuda()
    n_sum = float(N.double().sum())
    kappa = 0.3 * n_sum / float(C.double().sum())
    scale = prob.photons_per_view * g.n_views / (n_sum * 1.3)
    As = (A.view(nrow, ncol).index_select(0, rt) * scale).contiguous()
    cs = (C.reshape(-1).index_select(0, rt) * (kappa * scale)).contiguous()
    expected = (N.reshape(-1).index_select(0, rt) * scale + cs).double().cpu().numpy()
    ys = torch.from_numpy(np.random.default_rng(11).poisson(expected).astype(np.float32)).cuda()
    xg = xrd.mlem(As, ys, cs, 5).cpu().numpy()
    xo = oracle.mlem(As.double().cpu().numpy(), ys.double().cpu().numpy(),
                     cs.double().cpu().numpy(), np.full(ncol, 10.0), 5)
    live = xo > 1e-6 * xo.max()
    rel = float(np.max(np.abs(xg[live] - xo[live]) / xo[live]))
    print(f"\n[fan256 MLEM, 36 views, 5 updates] max rel err vs FP64 {rel:.2e}")
    assert rel <= 2e-5


# ---- the differentiation workload ------------------------------------------------

def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def test_vjp256_subphantom_vs_oracle(xrd, oracle):
    """32x32 voxels of the 256x256 generator (16 mm, same field), 30 views, 32 columns,
    K = 16, the full 61-energy map: counts, coefficient and map gradients."""
    import torch
    from xrd_tomography_amd.synthetic import make_coeff_problem, make_problem
    cp = make_coeff_problem(make_problem("fan256", nx=32, voxel_mm=16.0, n_views=30, n_cols=32),
                            K=16, full_mu=True)
    assert cp.K == 16 and cp.B == 61
    g, d = cp.base.geometry, cp.base.disc
    args = xrd.coeff_problem_tensors(cp)
    y = xrd.forward_coeffs(*args, geometry=g, disc=d).cpu().numpy().reshape(-1)
    yo = oracle.forward_coeffs(cp)
    live = yo > 1e-9 * yo.max()
    ey = float(np.max(np.abs(y[live] - yo[live]) / yo[live]))
    rng = np.random.default_rng(3)
    w = rng.normal(size=cp.base.n_rows).astype(np.float32)
    gc, gm = xrd.vjp_coeffs(*args, torch.from_numpy(w).cuda(), geometry=g, disc=d)
    torch.cuda.synchronize()
    gco, gmo = oracle.vjp_coeffs(cp, w.astype(np.float64))
    sco, smo = oracle.vjp_coeffs(cp, np.abs(w).astype(np.float64))
    gc, gm = gc.cpu().numpy(), gm.cpu().numpy()
    ec, em = _rel(gc, gco), _rel(gm, gmo)
    ec_el = float(np.max(np.abs(gc - gco) / np.maximum(np.abs(sco), 1e-300)))
    em_el = float(np.max(np.abs(gm - gmo) / np.maximum(np.abs(smo), 1e-300)))
    print(f"\n[vjp256 32x32x30 K=16 B=61] counts {ey:.2e}; coef {ec:.2e} (elementwise "
          f"{ec_el:.2e}); maps {em:.2e} (elementwise {em_el:.2e})")
    assert ey <= 1e-5
    assert ec <= 1e-5 and em <= 1e-5 and ec_el <= 1e-5 and em_el <= 1e-5


def test_vjp256_dot_product_12_views(xrd):
    """The 256x256 phantom and scanner at 12 of the 360 views (every 30th), K = 16,
    B = 61: <J x, u> = <x, J' u> to 1e-6."""
    import torch
    from xrd_tomography_amd.synthetic import make_coeff_problem, make_problem
    cp = make_coeff_problem(make_problem("fan256", n_views=12), K=16, full_mu=True)
    g, d = cp.base.geometry, cp.base.disc
    assert abs(g.view_angle_step - 30 * 2 * np.pi / 360) < 1e-15
    sp, seg, coef, basis, maps, beta = xrd.coeff_problem_tensors(cp)
    rng = np.random.default_rng(4)
    x = torch.from_numpy((rng.normal(size=cp.coef.shape) * (cp.coef != 0)).astype(np.float32)).cuda()
    u = torch.from_numpy(rng.normal(size=cp.base.n_rows).astype(np.float32)).cuda()
    jx = xrd.forward_coeffs(sp, seg, x, basis, maps, beta, geometry=g, disc=d).reshape(-1)
    jtu, gm = xrd.vjp_coeffs(sp, seg, coef, basis, maps, beta, u, geometry=g, disc=d)
    a = float((jx.double() * u.double()).sum())
    b = float((x.double() * jtu).sum())
    print(f"\n[vjp256 dot test, 12 views] <Jx,u> = {a:.9e}, <x,J'u> = {b:.9e}, "
          f"rel {abs(a - b) / abs(a):.2e}")
    assert abs(a - b) <= 1e-6 * max(abs(a), abs(b))
    assert bool(torch.isfinite(gm).all())


def test_vjp256_large_slab_matches_one_view_slabs(xrd, monkeypatch):
    """24 views of the 256x256 phantom in one slab: 79.6 M detector legs, whose trace
    is more than one dispatch's 2^32 work-items, so it is issued as several capped
    launches.  Counts and the coefficient VJP must equal, bit for bit, those of
    one-view slabs (XRD_VJP_SLAB_VIEWS=1), each far below the limit: the same rays
    through the same kernels."""
    import torch
    from xrd_tomography_amd.synthetic import make_coeff_problem, make_problem
    cp = make_coeff_problem(make_problem("fan256", n_views=24), K=16, full_mu=True)
    g, d = cp.base.geometry, cp.base.disc
    assert g.n_views * g.n_cols * int((cp.base.segmentation > 0).sum()) * 64 > 2**32
    args = xrd.coeff_problem_tensors(cp)
    u = torch.from_numpy(np.random.default_rng(6).normal(size=cp.base.n_rows)
                         .astype(np.float32)).cuda()

    def run():
        y = xrd.forward_coeffs(*args, geometry=g, disc=d)
        gc, _ = xrd.vjp_coeffs(*args, u, want_maps=False, geometry=g, disc=d)
        torch.cuda.synchronize()
        return y.cpu(), gc.cpu()

    y_big, gc_big = run()
    monkeypatch.setenv("XRD_VJP_SLAB_VIEWS", "1")
    y_one, gc_one = run()
    print(f"\n[vjp256 24 views] counts equal: {bool(torch.equal(y_big, y_one))}, "
          f"coefficient VJP max |diff| {float((gc_big - gc_one).abs().max()):.3e}")
    assert torch.equal(y_big, y_one)
    # the coefficient partials add slab by slab, so the order of the FP64 sums differs
    assert float((gc_big - gc_one).abs().max()) <= 1e-12 * float(gc_one.abs().max())


# ---- the 3-D volume -----------------------------------------------------------------

VOL_VIEWS = [0, 359]
VOL_COLS = [0, 40, 97, 128, 129, 170, 211, 255]


@pytest.mark.parametrize("tables", ["f32", "bf16"])
def test_vol192_slice_vs_oracle(xrd, oracle, tables):
    """Slice 16 of the 192x192x32 volume with all 720 views on the GPU; 2 views x 8
    columns of N, C and A against the oracle.  With BF16 tables the form factors and
    the attenuation maps are stored as bfloat16 with a per-material FP32 scale and
    decoded by the kernels (xrd_forward_bf16); the oracle runs on the decoded values."""
    import torch
    from xrd_tomography_amd.synthetic import decoded_bf16_tables, make_volume_problem
    nz, sl = 32, 16
    vp = make_volume_problem("fan256", n_slices=nz, seed=7, nx=192, voxel_mm=2.0, n_views=720,
                             n_mat=10, mlem_iterations=30)
    d = vp.disc
    print(f"\n[vol192 {tables}] building slice {sl} on the GPU", flush=True)
    if tables == "bf16":
        N, C, A = xrd.forward_problem_bf16(vp, nz, sl)
        ref = decoded_bf16_tables(vp)
    else:
        N, C, A = xrd.forward_volume_slice(vp, nz, sl)
        ref = vp
    torch.cuda.synchronize()
    Ng = _select(vp, N, VOL_VIEWS, VOL_COLS)
    Cg = _select(vp, C, VOL_VIEWS, VOL_COLS)
    Ag = _select(vp, A, VOL_VIEWS, VOL_COLS, d.n_mat * d.n_q)
    del N, C, A
    torch.cuda.empty_cache()
    No, Co, Ao = oracle.forward_slice(ref, nz, sl, views=VOL_VIEWS, cols=VOL_COLS)
    _compare(f"vol192 slice {sl}, {tables} tables, 2 views x 8 columns", Ng, Cg, Ag, No, Co, Ao)


# ---- batching ---------------------------------------------------------------------------

def test_batch48_matches_single_plans_and_oracle(xrd, oracle):
    """The 48 phantoms of the batching workload with GPU-built matrices in BF24 storage:
    10 batched updates (persistent launches, graph replay) give each phantom the x of
    its own single-problem plan, and the 4 smallest agree with the FP64 oracle MLEM on
    the stored (decoded) matrix."""
    import torch
    from xrd_tomography_amd.synthetic import make_batch_problems
    probs = make_batch_problems()
    stored = []
    for k, p in enumerate(probs):
        N, C, A = xrd.forward_problem(p)
        A = A.view(p.n_rows, p.n_cols)
        N, C = N.reshape(-1), C.reshape(-1)
        n_sum = float(N.double().sum())
        kappa = 0.3 * n_sum / max(float(C.double().sum()), 1e-300)
        scale = p.photons_per_view * p.geometry.n_views / (n_sum * 1.3)
        A.mul_(scale)
        c = (C * (kappa * scale)).contiguous()
        expected = (N * scale + c).double().cpu().numpy()
        y = torch.from_numpy(np.random.default_rng(k).poisson(expected).astype(np.float32)).cuda()
        stored.append((xrd.convert_matrix(A, xrd.BF24), y, c,
                       torch.full((p.n_cols,), 10.0, device="cuda")))
        del N, C, A
        if k % 12 == 11:
            print(f"[batch48] built {k + 1}/48", flush=True)
    n_it = 10
    batch = xrd.MlemBatch(stored)
    batch.iterate(n_it // 2, use_graph=False)
    batch.iterate(n_it - n_it // 2, use_graph=True)
    torch.cuda.synchronize()
    worst_single = 0.0
    for (Ab, y, c, xb) in stored:
        xs = xrd.mlem(Ab, y, c, n_it)
        scale = float(xs.abs().max())
        worst_single = max(worst_single, float((xb - xs).abs().max()) / max(scale, 1e-30))
    order = sorted(range(len(stored)), key=lambda k: stored[k][0].numel())[:4]
    worst = 0.0
    for k in order:
        Ab, y, c, xb = stored[k]
        xo = oracle.mlem(Ab.float().double().cpu().numpy(), y.double().cpu().numpy(),
                         c.double().cpu().numpy(), np.full(Ab.shape[1], 10.0), n_it)
        xb = xb.cpu().numpy()
        live = xo > 1e-6 * xo.max()
        worst = max(worst, float(np.max(np.abs(xb[live] - xo[live]) / xo[live])))
    batch.close()
    print(f"\n[batch48 bf24] batched vs single plans {worst_single:.2e} (of each x's max); "
          f"4 smallest vs FP64 {worst:.2e}")
    assert worst_single <= 1e-6
    assert worst <= 2e-5
