"""GPU parity of the explicit-matrix operators and MLEM (main.m:141-176).

Bars: A x <= 1e-6 relative per row; A' w and sum(A,1) <= 1e-12 relative
(FP64 column sums of FP32 products of exactly-representable inputs);
MLEM iterate after k updates <= 2e-5 relative (max over columns above 1e-6 of
the maximum) against the FP64 oracle run on the same matrix.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(seed=0, rows=20000, cols=512):
    rng = np.random.default_rng(seed)
    A = (rng.random((rows, cols)) ** 4 * 1e-3).astype(np.float32)
    A[rng.random((rows, cols)) < 0.3] = 0.0
    x_true = rng.uniform(0.5, 20.0, cols)
    c = (rng.uniform(0.0, 0.2, rows) * (A.astype(np.float64) @ x_true).mean()).astype(np.float32)
    y = rng.poisson(A.astype(np.float64) @ x_true + c).astype(np.float32)
    return A, y, c


@pytest.mark.parametrize("cols", [48, 384, 512, 640, 1024, 1280, 1536])
def test_project_backproject_sensitivity(xrd, oracle, cols):
    import torch
    A, y, c = _problem(1, rows=3001, cols=cols)
    rng = np.random.default_rng(2)
    x = rng.uniform(0, 2, cols).astype(np.float32)
    w = rng.uniform(-1, 1, A.shape[0]).astype(np.float32)
    Ad = torch.from_numpy(A).cuda()
    f = xrd.project(Ad, torch.from_numpy(x).cuda()).cpu().numpy()
    b = xrd.backproject(Ad, torch.from_numpy(w).cuda()).cpu().numpy()
    s = xrd.sensitivity(Ad).cpu().numpy()
    A64 = A.astype(np.float64)
    fo = oracle.project(A64, x.astype(np.float64))
    bo = oracle.backproject(A64, w.astype(np.float64))
    so = oracle.sensitivity(A64)
    ef = np.max(np.abs(f - fo) / np.maximum(np.abs(fo), 1e-30))
    eb = np.max(np.abs(b - bo)) / np.max(np.abs(bo))
    es = np.max(np.abs(s - so) / so)
    print(f"\n[ops cols={cols}] Ax rel {ef:.1e}, A'w {eb:.1e}, sum(A) {es:.1e}")
    assert ef <= 1e-6 and eb <= 1e-12 and es <= 1e-12


@pytest.mark.parametrize("fused", [True, False])
def test_mlem_vs_oracle(xrd, oracle, fused):
    import torch
    A, y, c = _problem(3)
    n_iter = 20
    x = xrd.mlem(torch.from_numpy(A).cuda(), torch.from_numpy(y).cuda(),
                 torch.from_numpy(c).cuda(), n_iter, fused=fused).cpu().numpy()
    xo = oracle.mlem(A.astype(np.float64), y, c, np.full(A.shape[1], 10.0), n_iter)
    live = xo > 1e-6 * xo.max()
    rel = np.max(np.abs(x[live] - xo[live]) / xo[live])
    print(f"\n[mlem fused={fused}] max rel err vs FP64 after {n_iter} its: {rel:.2e}")
    assert rel <= 2e-5
    assert np.all(x[~live] <= 1e-5 * xo.max())


def test_mlem_graph_matches_eager(xrd):
    """Graph replays equal eager iterations bit for bit, for iteration counts served by
    the one-iteration graph (4), the eight-iteration graph (16) and both (11)."""
    import torch
    A, y, c = _problem(4, rows=5000, cols=256)
    Ad, yd, cd = (torch.from_numpy(v).cuda() for v in (A, y, c))
    x1 = torch.full((A.shape[1],), 10.0, device="cuda")
    x2 = x1.clone()
    p1 = xrd.MlemPlan(Ad, yd, cd, x1)
    p2 = xrd.MlemPlan(Ad, yd, cd, x2)
    for n in (4, 11, 16, 4):  # replays with eager work and syncs in between
        p1.iterate(n, use_graph=False)
        p2.iterate(n, use_graph=True)
        torch.cuda.synchronize()
        assert torch.equal(x1, x2)
    p1.close()
    p2.close()


def test_mlem_count_conservation_and_monotone(xrd):
    """With c = 0 each update preserves sum(A x) = sum(y) (up to the 1e-7 guard), and the
    Poisson log-likelihood never decreases."""
    import torch
    A, y, _ = _problem(5, rows=8000, cols=256)
    c = np.zeros_like(y)
    Ad, yd, cd = (torch.from_numpy(v).cuda() for v in (A, y, c))
    x = torch.full((A.shape[1],), 10.0, device="cuda")
    p = xrd.MlemPlan(Ad, yd, cd, x)
    ll_prev = p.loglik()
    for _ in range(10):
        p.iterate(1)
        ll = p.loglik()
        assert ll >= ll_prev - 1e-9 * abs(ll_prev)
        ll_prev = ll
        tot = float(xrd.project(Ad, x).double().sum())
        assert abs(tot - float(y.sum())) <= 1e-5 * float(y.sum())
    p.close()


def test_mlem_deterministic(xrd):
    import torch
    A, y, c = _problem(6, rows=10000, cols=1024)
    args = [torch.from_numpy(v).cuda() for v in (A, y, c)]
    a = xrd.mlem(*args, 5)
    b = xrd.mlem(*args, 5)
    assert torch.equal(a, b)


@pytest.mark.parametrize("dtype", ["bfloat16", "float16"])
@pytest.mark.parametrize("cols,offset", [(512, 0), (1024, 0), (516, 0), (1028, 0), (512, 4),
                                         (2048, 0), (768, 0), (1280, 0), (1284, 4)])
def test_reduced_precision_matrix(xrd, dtype, cols, offset):
    """BF16/F16 storage of A: operators equal the FP32 operators on the rounded matrix.
    Column counts that are multiples of 8 on 16-byte aligned storage take the 16-byte
    load path; 516/1028 columns and a matrix starting 8 bytes into its buffer take the
    8-byte path."""
    import torch
    A, y, c = _problem(7, rows=4000, cols=cols)
    A = A * 1e3  # keep F16 in its normal range
    Ad = torch.from_numpy(A).cuda()
    dt = getattr(torch, dtype)
    Ar = xrd.convert_matrix(Ad, dt)
    assert torch.equal(Ar, Ad.to(dt))  # round-to-nearest-even like torch
    if offset:
        buf = torch.empty(Ar.numel() + offset, dtype=dt, device="cuda")
        buf[offset:].copy_(Ar.reshape(-1))
        Ar = buf[offset:].view(Ar.shape)
        assert Ar.data_ptr() % 16 == 2 * offset % 16
    x = torch.rand(cols, device="cuda")
    w = torch.rand(Ar.shape[0], device="cuda") - 0.5
    Af = Ar.float().contiguous()
    f1 = xrd.project(Ar, x)
    f2 = xrd.project(Af, x)
    assert torch.allclose(f1, f2, rtol=1e-6, atol=0)
    s1 = xrd.sensitivity(Ar)
    s2 = xrd.sensitivity(Af)
    assert torch.allclose(s1, s2, rtol=1e-12, atol=0)
    b1 = xrd.backproject(Ar, w).double()
    b2 = xrd.backproject(Af, w).double()
    assert (b1 - b2).abs().max().item() <= 1e-12 * b2.abs().max().item()
    # fused MLEM on the 2-byte matrix equals the same update on its FP32 copy
    x1 = xrd.mlem(Ar, torch.from_numpy(y).cuda(), torch.from_numpy(c).cuda() * 1e3, 3)
    x2 = xrd.mlem(Af, torch.from_numpy(y).cuda(), torch.from_numpy(c).cuda() * 1e3, 3)
    assert torch.allclose(x1, x2, rtol=1e-5, atol=1e-6 * x2.abs().max().item())


def test_bf24_conversion_rounds_to_16_significant_bits(xrd):
    """BF24 storage keeps 16 significant bits, rounded to nearest (ties to even): the
    decoded matrix equals an FP64 rounding of each value to 16 significant bits."""
    import torch
    rng = np.random.default_rng(30)
    A = (rng.standard_normal((300, 64)) * np.exp(rng.uniform(-30, 30, (300, 64)))).astype(np.float32)
    A[0, :8] = [0.0, -0.0, 1.0, 1.0 + 2.0 ** -16, 1.0 + 3 * 2.0 ** -16, 3.0e38, 1e-40, -2.5]
    Ab = xrd.convert_matrix(torch.from_numpy(A).cuda(), xrd.BF24)
    got = Ab.float().cpu().numpy().astype(np.float64)
    a = A.astype(np.float64)
    e = np.floor(np.log2(np.abs(np.where(a == 0, 1.0, a))))
    e = np.maximum(e, -126.0)  # subnormals keep the spacing of the smallest binade
    q = 2.0 ** (e - 15)
    want = np.round(a / q) * q  # numpy rounds half to even
    assert np.array_equal(got, want)
    assert got[0, 3] == 1.0 and got[0, 4] == 1.0 + 4 * 2.0 ** -16  # ties to even


@pytest.mark.parametrize("cols", [512, 640, 1024, 1280, 2048])
def test_bf24_operators_and_mlem(xrd, oracle, cols):
    """BF24 matrices: A x, A' w, sum(A) and MLEM equal the FP32 operators on the decoded
    matrix, and MLEM is within 2e-5 of the FP64 oracle on the decoded matrix."""
    import torch
    A, y, c = _problem(9, rows=6000, cols=cols)
    Ab = xrd.convert_matrix(torch.from_numpy(A).cuda(), xrd.BF24)
    Af = Ab.float().contiguous()
    x = torch.rand(cols, device="cuda")
    w = torch.rand(A.shape[0], device="cuda") - 0.5
    assert torch.allclose(xrd.project(Ab, x), xrd.project(Af, x), rtol=1e-6, atol=0)
    assert torch.allclose(xrd.sensitivity(Ab), xrd.sensitivity(Af), rtol=1e-12, atol=0)
    b1, b2 = xrd.backproject(Ab, w), xrd.backproject(Af, w)
    assert (b1 - b2).abs().max().item() <= 1e-12 * b2.abs().max().item()
    yd, cd = torch.from_numpy(y).cuda(), torch.from_numpy(c).cuda()
    xb = xrd.mlem(Ab, yd, cd, 10).cpu().numpy()
    xo = oracle.mlem(Af.double().cpu().numpy(), y, c, np.full(cols, 10.0), 10)
    live = xo > 1e-6 * xo.max()
    rel = np.max(np.abs(xb[live] - xo[live]) / xo[live])
    xo32 = oracle.mlem(A.astype(np.float64), y, c, np.full(cols, 10.0), 10)
    storage = np.max(np.abs(xb[live] - xo32[live]) / xo32[live])
    print(f"\n[bf24 cols={cols}] vs FP64 on stored {rel:.2e}; vs FP64 on the FP32 matrix {storage:.2e}")
    assert rel <= 2e-5 and storage <= 1e-4


def test_bf24_rejects_unsupported_layouts(xrd):
    import torch
    from xrd_tomography_amd._lib import XrdError
    A = torch.rand(64, 516, device="cuda")
    Ab = xrd.convert_matrix(A, xrd.BF24)  # conversion itself takes any shape
    with pytest.raises(XrdError):
        xrd.sensitivity(Ab)  # n_cols % 8 != 0
    A = torch.rand(64, 512, device="cuda")
    Ab = xrd.convert_matrix(A, xrd.BF24)
    off = xrd.Bf24Matrix(torch.empty(Ab.data.numel() + 8, dtype=torch.uint8,
                                     device="cuda")[8:], 64, 512)
    with pytest.raises(XrdError):
        xrd.sensitivity(off)  # 8-byte aligned only


@pytest.mark.parametrize("dtype", ["float16", "bfloat16"])
def test_scaled_reduced_precision_mlem(xrd, oracle, dtype):
    """Per-column power-of-two scaled F16/BF16 storage: MLEM on the stored matrix equals
    the FP64 MLEM on the de-scaled rounded matrix (the scale is exact)."""
    import torch
    A, y, c = _problem(8, rows=20000, cols=512)
    A = A * 1e-6  # far below the F16 normal range without the scale
    c = c * 1e-6
    Ad = torch.from_numpy(A).cuda()
    dt = getattr(torch, dtype)
    A16, sc = xrd.convert_matrix_scaled(Ad, dt)
    scale = sc.cpu().numpy()
    assert np.all(np.log2(scale) == np.round(np.log2(scale)))  # powers of two
    Aeff = (A16.double().cpu().numpy() / scale[None, :])
    x = xrd.mlem(A16, torch.from_numpy(y).cuda(), torch.from_numpy(c).cuda(), 10,
                 col_scale=sc).cpu().numpy()
    xo = oracle.mlem(Aeff, y, c, np.full(A.shape[1], 10.0), 10)
    live = xo > 1e-6 * xo.max()
    rel = np.max(np.abs(x[live] - xo[live]) / xo[live])
    xo32 = oracle.mlem(A.astype(np.float64), y, c, np.full(A.shape[1], 10.0), 10)
    storage = np.max(np.abs(x[live] - xo32[live]) / xo32[live])
    print(f"\n[scaled {dtype}] vs FP64 on stored matrix {rel:.2e}; vs FP64 on the FP32 matrix "
          f"{storage:.2e}")
    assert rel <= 2e-5


@pytest.mark.parametrize("dtype,wide", [("float32", 1024), ("float16", 1024),
                                        ("float16", 1020), ("bfloat16", 1024), ("bf24", 1024),
                                        ("float32", 640), ("bf24", 1280), ("float16", 1276),
                                        ("bf24", 384)])
def test_batched_mlem_matches_single(xrd, oracle, dtype, wide):
    """Problems of uneven rows and columns in one persistent launch: each x equals the
    single-problem plan's x (1e-6) and the FP64 oracle's (2e-5 on the stored matrix).
    A 1020-column problem puts a 2-byte batch on its 8-byte load path."""
    import torch
    shapes = [(3001, 48), (20000, 512), (777, wide), (12000, 256), (5, 16)]
    probs, singles, refs = [], [], []
    for i, (r, cc) in enumerate(shapes):
        A, y, c = _problem(20 + i, rows=r, cols=cc)
        Ad = torch.from_numpy(A).cuda()
        sc = None
        if dtype == "bf24":
            Ad = xrd.convert_matrix(Ad, xrd.BF24)
        elif dtype != "float32":
            Ad, sc = xrd.convert_matrix_scaled(Ad, getattr(torch, dtype))
        yd, cd = torch.from_numpy(y).cuda(), torch.from_numpy(c).cuda()
        x = torch.full((cc,), 10.0, device="cuda")
        probs.append((Ad, yd, cd, x, sc) if sc is not None else (Ad, yd, cd, x))
        singles.append(xrd.mlem(Ad, yd, cd, 8, col_scale=sc))
        Aeff = Ad.float().double().cpu().numpy() / (sc.double().cpu().numpy()[None, :]
                                                    if sc is not None else 1.0)
        refs.append(oracle.mlem(Aeff, y, c, np.full(cc, 10.0), 8))
    batch = xrd.MlemBatch(probs)
    batch.set_launch(1, 1)  # 256 workgroups: several problems share a workgroup's chunk
    batch.iterate(4, use_graph=False)
    batch.iterate(4, use_graph=True)
    torch.cuda.synchronize()
    # a batch on the 8-byte path sums its other problems' rows in another FP32 order
    # than their single plans on the 16-byte path
    tol = 1e-6 if dtype == "float32" or wide % 8 == 0 else 1e-5
    worst = 0.0
    for (p, xs, xo) in zip(probs, singles, refs):
        xb = p[3].cpu().numpy()
        assert np.allclose(xb, xs.cpu().numpy(), rtol=tol, atol=tol * xs.abs().max().item())
        live = xo > 1e-6 * xo.max()
        if not live.any():
            # the 5-row problem has no counts: every column is driven to exactly zero
            assert np.all(xb == 0.0) and np.all(xo == 0.0)
            continue
        worst = max(worst, float(np.max(np.abs(xb[live] - xo[live]) / xo[live])))
    print(f"\n[batched {dtype}] worst rel err vs FP64 {worst:.2e}")
    assert worst <= 2e-5
    batch.close()


@pytest.mark.parametrize("levels", ["-1", "6"], ids=["static chunks", "guided chunks"])
@pytest.mark.parametrize("dtype", ["float32", "bf24"])
def test_batched_graph_matches_eager_every_replay(xrd, monkeypatch, dtype, levels):
    """Two identical batches of five row-span groups, one iterated eagerly and one
    by replaying its captured graph with a synchronization between replays, agree
    bit for bit after every iteration: each chunk writes its own partial slots and
    the update sums them in a fixed order, whichever workgroup took the chunk."""
    monkeypatch.setenv("XRD_BATCH_CHUNK_LEVELS", levels)
    import torch
    shapes = [(3001, 48), (20000, 512), (4000, 1024), (12000, 256), (2500, 384), (900, 768)]

    def make():
        probs = []
        for i, (r, cc) in enumerate(shapes):
            A, y, c = _problem(40 + i, rows=r, cols=cc)
            Ad = torch.from_numpy(A).cuda()
            if dtype == "bf24":
                Ad = xrd.convert_matrix(Ad, xrd.BF24)
            probs.append((Ad, torch.from_numpy(y).cuda(), torch.from_numpy(c).cuda(),
                          torch.full((cc,), 10.0, device="cuda")))
        return probs

    pe, pg = make(), make()
    be, bg = xrd.MlemBatch(pe), xrd.MlemBatch(pg)
    assert len(be.groups()) >= 4
    for _ in range(4):
        be.iterate(1, use_graph=False)
        bg.iterate(1, use_graph=True)
        torch.cuda.synchronize()
        for a, b in zip(pe, pg):
            assert torch.equal(a[3], b[3])
    be.close()
    bg.close()


def test_batched_static_and_guided_chunks_agree(xrd, monkeypatch):
    """The same problems with one chunk per workgroup and with guided chunks taken from
    a counter (0, 6 and 8 rounds, the last with far more chunks than 16-row blocks,
    so many chunks are empty): x after 6 iterations agrees to FP32 rounding, and two
    guided runs agree bit for bit."""
    import torch
    shapes = [(3001, 48), (20000, 512), (777, 1024), (12000, 256), (5, 16)]
    data = []
    for i, (r, cc) in enumerate(shapes):
        A, y, c = _problem(60 + i, rows=r, cols=cc)
        data.append((torch.from_numpy(A).cuda(), torch.from_numpy(y).cuda(),
                     torch.from_numpy(c).cuda()))

    def run(levels, blocks_per_cu):
        monkeypatch.setenv("XRD_BATCH_CHUNK_LEVELS", levels)
        probs = [(A, y, c, torch.full((A.shape[1],), 10.0, device="cuda")) for A, y, c in data]
        b = xrd.MlemBatch(probs)
        b.set_launch(blocks_per_cu, 1)
        b.iterate(6, use_graph=True)
        torch.cuda.synchronize()
        b.close()
        return [p[3].cpu().numpy() for p in probs]

    ref = run("-1", 2)
    for levels, bpc in (("0", 2), ("6", 2), ("8", 4)):
        got = run(levels, bpc)
        for xr, xg in zip(ref, got):
            assert np.allclose(xg, xr, rtol=1e-6, atol=1e-6 * max(float(np.abs(xr).max()), 1e-30))
    again = run("8", 4)
    for a, b in zip(got, again):
        assert np.array_equal(a, b)
