"""xrd_mlem_iterate_static, the MLEM entry point that allocates nothing, against the
plan entry point (xrd_mlem_plan_create + xrd_mlem_plan_iterate, default launch shape):
the same kernels in the same launch shape, so x must be bit-identical.  Called
through the C ABI twice in a row, the second time on a smaller problem, since the
fixed device scratch keeps the previous call's contents."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(seed, rows, cols):
    rng = np.random.default_rng(seed)
    A = (rng.random((rows, cols)) ** 4 * 1e-3).astype(np.float32)
    A[rng.random((rows, cols)) < 0.3] = 0.0
    x_true = rng.uniform(0.5, 20.0, cols)
    c = (rng.uniform(0.0, 0.2, rows) * (A.astype(np.float64) @ x_true).mean()).astype(np.float32)
    y = rng.poisson(A.astype(np.float64) @ x_true + c).astype(np.float32)
    return A, y, c


def _static(xrd, A, dtype_code, rows, cols, scale, y, c, x, n_iter):
    import torch
    L = xrd.lib()
    ctx = xrd.default_context(0)
    vp = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = L.xrd_mlem_iterate_static(ctx.handle, ctypes.c_void_p(A.data_ptr()), dtype_code, rows,
                                   cols, vp(scale), vp(y), vp(c), vp(x), n_iter, s)
    return rc, L.xrd_last_error_string().decode()


@pytest.mark.parametrize("dtype", ["float32", "float16"])
def test_static_matches_plan_two_calls(xrd, dtype):
    import torch
    for seed, (rows, cols) in enumerate([(20000, 1024), (3001, 256)]):  # second call smaller
        A, y, c = _problem(60 + seed, rows, cols)
        Ad = torch.from_numpy(A).cuda()
        sc = None
        code = xrd.DTYPE_F32
        if dtype == "float16":
            Ad, sc = xrd.convert_matrix_scaled(Ad, torch.float16)
            code = xrd.DTYPE_F16
        yd, cd = torch.from_numpy(y).cuda(), torch.from_numpy(c).cuda()
        x_plan = torch.full((cols,), 10.0, device="cuda")
        plan = xrd.MlemPlan(Ad, yd, cd, x_plan, col_scale=sc)
        plan.iterate(6, use_graph=False)
        x_st = torch.full((cols,), 10.0, device="cuda")
        rc, msg = _static(xrd, Ad, code, rows, cols, sc, yd, cd, x_st, 6)
        torch.cuda.synchronize()
        plan.close()
        assert rc == 0, msg
        print(f"\n[static {dtype} {rows}x{cols}] max |x_static - x_plan| = "
              f"{float((x_st - x_plan).abs().max()):.1e}")
        assert torch.equal(x_st, x_plan)


def test_static_allocates_no_device_memory(xrd):
    import torch
    A, y, c = _problem(70, 8000, 512)
    Ad, yd, cd = (torch.from_numpy(v).cuda() for v in (A, y, c))
    x = torch.full((512,), 10.0, device="cuda")
    rc, msg = _static(xrd, Ad, xrd.DTYPE_F32, 8000, 512, None, yd, cd, x, 1)  # warm-up
    torch.cuda.synchronize()
    assert rc == 0, msg
    free0, _ = torch.cuda.mem_get_info()
    for _ in range(3):
        rc, msg = _static(xrd, Ad, xrd.DTYPE_F32, 8000, 512, None, yd, cd, x, 2)
        assert rc == 0, msg
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0


def test_static_rejects_problems_beyond_the_scratch(xrd):
    """2,048 columns at 512 workgroups need 8.4 MB of scratch: refused before any launch,
    x untouched."""
    import torch
    rows, cols = 100000, 2048
    Ad = torch.zeros((rows, cols), dtype=torch.float32, device="cuda")
    yd = torch.ones(rows, device="cuda")
    cd = torch.zeros(rows, device="cuda")
    x = torch.full((cols,), 10.0, device="cuda")
    rc, msg = _static(xrd, Ad, xrd.DTYPE_F32, rows, cols, None, yd, cd, x, 3)
    torch.cuda.synchronize()
    assert rc == -3 and "scratch" in msg, (rc, msg)
    assert bool((x == 10.0).all())
