"""GPU parity of the forward projector against the FP64 oracle.

Bars (stated here and asserted):
  * ray traversal (entry side, entry point, sample count, texel indices): bit-exact;
  * line integrals (Kahan FP32): <= 2e-6 relative to the FP64 oracle;
  * XRD counts N and Compton counts C: <= 1e-5 relative per row, for rows above
    1e-9 of the largest row (below that, <= 1e-14 of the largest row absolute);
  * model matrix A: <= 1e-5 relative Frobenius error per row, elementwise
    <= 1e-5 relative + max(1e-8 of the row maximum, 1e-12 of the matrix maximum).
Every test prints the measured error against FP64.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL_COUNTS = 1e-5
RTOL_A_ROW = 1e-5


def _rel_rows(gpu, ora, floor_frac=1e-9):
    gpu = np.asarray(gpu, np.float64)
    ora = np.asarray(ora, np.float64)
    big = np.abs(ora).max()
    mask = np.abs(ora) > floor_frac * big
    rel = np.abs(gpu[mask] - ora[mask]) / np.abs(ora[mask])
    small_abs = np.abs(gpu[~mask] - ora[~mask]).max(initial=0.0) / max(big, 1e-300)
    return (rel.max(initial=0.0), small_abs)


def _run(xrd, prob, views=None, oracle=None, with_A=True):
    import torch
    N, C, A = xrd.forward_problem(prob, with_A=with_A)
    torch.cuda.synchronize()
    g, d = prob.geometry, prob.disc
    Nh = N.cpu().numpy().reshape(d.n_energy, g.n_rows * g.n_cols, g.n_views)
    Ch = C.cpu().numpy().reshape(d.n_energy, g.n_rows * g.n_cols, g.n_views)
    Ah = (A.cpu().numpy().reshape(d.n_energy, g.n_rows * g.n_cols, g.n_views, d.n_mat * d.n_q)
          if with_A else None)
    if views is not None:
        Nh, Ch = Nh[:, :, views], Ch[:, :, views]
        Ah = Ah[:, :, views, :] if with_A else None
    return Nh.reshape(-1), Ch.reshape(-1), (Ah.reshape(-1, d.n_mat * d.n_q) if with_A else None)


def _compare(name, Ng, Cg, Ag, No, Co, Ao):
    rn, an = _rel_rows(Ng, No)
    rc, ac = _rel_rows(Cg, Co)
    print(f"\n[{name}] N: max rel err vs FP64 {rn:.2e} (tiny rows abs {an:.1e}); "
          f"C: {rc:.2e} ({ac:.1e})")
    assert rn <= RTOL_COUNTS and an <= 1e-14, (rn, an)
    assert rc <= RTOL_COUNTS and ac <= 1e-14, (rc, ac)
    if Ag is not None:
        Ao = np.asarray(Ao, np.float64)
        Ag = np.asarray(Ag, np.float64)
        rown = np.linalg.norm(Ao, axis=1)
        live = rown > 1e-9 * rown.max()
        rowerr = np.linalg.norm(Ag - Ao, axis=1)[live] / rown[live]
        rowmax = np.abs(Ao).max(axis=1, keepdims=True)
        # rows whose every entry is ~1e-100 (fully attenuated energies) underflow in FP32:
        # the absolute floor is 1e-8 of the row maximum or 1e-12 of the matrix maximum
        elem = (np.abs(Ag - Ao) - RTOL_A_ROW * np.abs(Ao) -
                np.maximum(1e-8 * rowmax, 1e-12 * np.abs(Ao).max()))
        print(f"[{name}] A: max row rel err {rowerr.max(initial=0):.2e}, "
              f"elementwise excess {elem.max():.1e}")
        assert rowerr.max(initial=0.0) <= RTOL_A_ROW
        assert elem.max() <= 0.0


@pytest.mark.parametrize("flags", [0, 3])
def test_forward_small_all_views(xrd, oracle, flags):
    from xrd_tomography_amd.synthetic import make_problem
    prob = make_problem("small", flags=flags)
    Ng, Cg, Ag = _run(xrd, prob)
    No, Co, Ao = oracle.forward(prob, nthreads=4)
    _compare(f"small flags={flags}", Ng, Cg, Ag, No, Co, Ao)


@pytest.mark.parametrize("flags", [0, 3])
def test_forward_reference_geometry(xrd, oracle, flags):
    """The reference's own scanner literals (integrate_diffraction.cu:81-140)."""
    from xrd_tomography_amd.synthetic import make_problem
    prob = make_problem("reference", flags=flags)
    Ng, Cg, Ag = _run(xrd, prob)
    No, Co, Ao = oracle.forward(prob, nthreads=4)
    _compare(f"reference flags={flags}", Ng, Cg, Ag, No, Co, Ao)


@pytest.mark.parametrize("name,rows", [("small", 3), ("reference", 2)])
def test_forward_multi_row_detector(xrd, oracle, name, rows):
    """ROW > 1 (integrate_diffraction.cu:39-40,100-109): rows stacked in z."""
    from xrd_tomography_amd.synthetic import make_problem
    prob = make_problem(name, n_rows=rows, flags=3 if name == "reference" else 0)
    Ng, Cg, Ag = _run(xrd, prob)
    No, Co, Ao = oracle.forward(prob, nthreads=4)
    _compare(f"{name} rows={rows}", Ng, Cg, Ag, No, Co, Ao)


@pytest.mark.parametrize("flags", [0, 3])
def test_forward_paper100(xrd, oracle, flags):
    """The paper's own problem (main.m:12-63, built from the reference's input files:
    100x100 phantom, 3 measured materials, 768 fine and 256 coarse q-bins, 64
    energies), one view of a 16-column detector on the reference scanner."""
    import dataclasses
    from xrd_tomography_amd.descriptors import Discretization
    from xrd_tomography_amd.synthetic import make_problem
    prob = make_problem("paper100", n_cols=16)
    if flags:
        d = prob.disc
        prob = dataclasses.replace(prob, disc=Discretization.reference(
            n_energy=d.n_energy, n_q=d.n_q, n_qf=d.n_qf, n_mat=d.n_mat, flags=flags))
    views = [5]
    Ng, Cg, Ag = _run(xrd, prob, views=views)
    No, Co, Ao = oracle.forward(prob, views=views, nthreads=16)
    _compare(f"paper100 flags={flags}", Ng, Cg, Ag, No, Co, Ao)


def test_forward_bf16_tables(xrd, oracle):
    """xrd_forward_bf16: form factors and attenuation maps stored as bfloat16 with a
    per-material FP32 scale, decoded by the kernels.  The GPU matches the FP64 oracle
    on the decoded tables to the bars above.  A does not depend on the form factors
    and per-material constant maps decode exactly, so A is bit-identical to the
    FP32-table A; each form-factor entry moves by at most 2^-8 of itself and N is a
    positive combination of the entries, so N moves by at most 2^-8 from the
    FP32-table counts."""
    import torch
    from xrd_tomography_amd.synthetic import decoded_bf16_tables, make_problem
    p = make_problem("small")
    q = decoded_bf16_tables(p)
    assert np.array_equal(q.a1, p.a1) and np.array_equal(q.a2, p.a2)
    assert not np.array_equal(q.ifine, p.ifine)
    Nb, Cb, Ab = xrd.forward_problem_bf16(p)
    Ng, Cg, Ag = _host(p, Nb, Cb, Ab)
    No, Co, Ao = oracle.forward(q, nthreads=4)
    _compare("small, BF16 tables", Ng, Cg, Ag, No, Co, Ao)
    _, _, Ap = xrd.forward_problem(p)
    torch.cuda.synchronize()
    assert torch.equal(Ap, Ab)
    Np, _, _ = oracle.forward(p, with_A=False, nthreads=4)
    live = Np > 1e-9 * Np.max()
    rel = float(np.max(np.abs(np.asarray(Ng, np.float64)[live] - Np[live]) / Np[live]))
    print(f"[BF16 tables] counts vs FP64 on the FP32 tables: {rel:.2e} (bound 2^-8 = 3.9e-3)")
    assert rel <= 2.0 ** -8 + RTOL_COUNTS


def test_forward_bf16_tables_varying_maps_and_volume(xrd, oracle):
    """Maps that vary inside each material (so their BF16 rounding reaches A), in 2-D
    and as slice 2 of a 4-slice volume: GPU against the oracle on the decoded tables;
    the 2-D BF16 path with the FP32 tables' exact values is xrd_forward itself."""
    import torch
    from xrd_tomography_amd.synthetic import (decoded_bf16_tables, make_problem,
                                              make_volume_problem)
    rng = np.random.default_rng(12)
    p = make_problem("small")
    p.a1 = (p.a1 * rng.uniform(0.6, 1.4, p.a1.shape)).astype(np.float32)
    p.a2 = (p.a2 * rng.uniform(0.6, 1.4, p.a2.shape)).astype(np.float32)
    q = decoded_bf16_tables(p)
    assert not np.array_equal(q.a1, p.a1)
    Ng, Cg, Ag = _host(p, *xrd.forward_problem_bf16(p))
    No, Co, Ao = oracle.forward(q, nthreads=4)
    _compare("small, BF16 tables, varying maps", Ng, Cg, Ag, No, Co, Ao)
    # the FP32 path on the decoded tables computes with the same values: same bits
    N2, C2, A2 = _host(q, *xrd.forward_problem(q))
    assert np.array_equal(Ng, N2) and np.array_equal(Cg, C2) and np.array_equal(Ag, A2)
    vp = make_volume_problem("small", n_slices=4)
    vp.a2 = (vp.a2 * rng.uniform(0.6, 1.4, vp.a2.shape)).astype(np.float32)
    vq = decoded_bf16_tables(vp)
    Ng, Cg, Ag = _host(vp, *xrd.forward_problem_bf16(vp, 4, 2))
    No, Co, Ao = oracle.forward_slice(vq, 4, 2, nthreads=4)
    _compare("volume slice 2/4, BF16 tables", Ng, Cg, Ag, No, Co, Ao)
    torch.cuda.synchronize()


def _host(prob, N, C, A):
    import torch
    torch.cuda.synchronize()
    g, d = prob.geometry, prob.disc
    return (N.cpu().numpy().reshape(-1), C.cpu().numpy().reshape(-1),
            A.cpu().numpy().reshape(-1, d.n_mat * d.n_q))


@pytest.mark.parametrize("slice_", [0, 2])
def test_forward_volume_slice(xrd, oracle, slice_):
    """A slice of a 4-slice volume: both attenuation legs are 3-D segments through the
    volume, sampled trilinearly (xrd_forward_slice against xo_forward_slice)."""
    from xrd_tomography_amd.synthetic import make_volume_problem
    vp = make_volume_problem("small", n_slices=4)
    Ng, Cg, Ag = _host(vp, *xrd.forward_volume_slice(vp, 4, slice_))
    No, Co, Ao = oracle.forward_slice(vp, 4, slice_, nthreads=4)
    _compare(f"volume slice {slice_}/4", Ng, Cg, Ag, No, Co, Ao)


def test_forward_slice_of_one_is_forward(xrd):
    import torch
    from xrd_tomography_amd.synthetic import make_problem
    p = make_problem("small")
    N0, C0, A0 = xrd.forward_problem(p)
    N1, C1, A1 = xrd.forward_volume_slice(p, 1, 0)
    torch.cuda.synchronize()
    assert torch.equal(N0, N1) and torch.equal(C0, C1) and torch.equal(A0, A1)


def test_forward_fan64_view_subset(xrd, oracle):
    """64x64 voxels, 60 views, 64 pixels, 61 energies, 128 q-bins (3 views checked)."""
    from xrd_tomography_amd.synthetic import make_problem
    prob = make_problem("fan64")
    views = [0, 17, 44]
    Ng, Cg, Ag = _run(xrd, prob, views=views)
    No, Co, Ao = oracle.forward(prob, views=views, nthreads=8)
    _compare("fan64 views 0,17,44", Ng, Cg, Ag, No, Co, Ao)


def test_forward_without_A_matches(xrd):
    import torch
    from xrd_tomography_amd.synthetic import make_problem
    prob = make_problem("small")
    N1, C1, A1 = xrd.forward_problem(prob, with_A=True)
    N2, C2, A2 = xrd.forward_problem(prob, with_A=False)
    torch.cuda.synchronize()
    assert A2 is None
    assert torch.equal(N1, N2) and torch.equal(C1, C2)


def test_forward_deterministic(xrd):
    import torch
    from xrd_tomography_amd.synthetic import make_problem
    prob = make_problem("reference", flags=3)
    a = xrd.forward_problem(prob)
    b = xrd.forward_problem(prob)
    torch.cuda.synchronize()
    for u, v in zip(a, b):
        assert torch.equal(u, v)


def test_trace_rays_bit_exact(xrd, oracle):
    """Intersection lists: side, entry point, count, texel indices identical bits."""
    import torch
    rng = np.random.default_rng(5)
    nx, ny, step = 37, 29, 0.25
    n = 3000
    # endpoints outside the ROI, targets at voxel centres inside (the forward's rays)
    ang = rng.uniform(0, 2 * np.pi, n)
    rad = rng.uniform(40, 200, n)
    x1 = nx / 2 + rad * np.cos(ang)
    y1 = ny / 2 + rad * np.sin(ang)
    x2 = rng.integers(0, nx, n) + 0.5
    y2 = rng.integers(0, ny, n) + 0.5
    segs = np.stack([x1, y1, x2, y2], axis=1)
    # special cases: axis-aligned, diagonal through corners, source inside (no hit)
    extra = np.array([[-10.0, 0.5, 3.5, 0.5], [0.5, -10.0, 0.5, 3.5], [-10, -10, 5.5, 5.5],
                      [nx + 10.0, ny + 10.0, 1.5, 1.5], [5.5, 5.5, 6.5, 6.5],
                      [nx / 2, -50.0, nx / 2, 3.5]])
    segs = np.concatenate([segs, extra])
    maxs = int(np.ceil(np.hypot(nx, ny) / step)) + 2
    side, entry, count, tex, _ = xrd.trace_rays(torch.from_numpy(segs).cuda(), nx, ny, step,
                                                max_samples=maxs)
    torch.cuda.synchronize()
    side, entry, count, tex = side.cpu().numpy(), entry.cpu().numpy(), count.cpu().numpy(), tex.cpu().numpy()
    mism = 0
    for i, (a, b, c, d) in enumerate(segs):
        s, ex, ey, k, t = oracle.trace_ray(a, b, c, d, nx, ny, step, maxs)
        ok = (s == side[i] and k == count[i] and np.array_equal(t, tex[i]) and
              (np.isnan(ex) and np.isnan(entry[i, 0]) or
               np.float64(ex).tobytes() == entry[i, 0].tobytes() and
               np.float64(ey).tobytes() == entry[i, 1].tobytes()))
        mism += (not ok)
    print(f"\n[trace] {len(segs)} rays, {mism} mismatches")
    assert mism == 0


def test_line_integrals_vs_oracle(xrd, oracle):
    import torch
    rng = np.random.default_rng(6)
    nx, ny, step = 48, 40, 0.25
    maps = rng.uniform(0.0, 0.1, (2, ny, nx)).astype(np.float32)  # [b, y, x] -> x fastest
    n = 500
    ang = rng.uniform(0, 2 * np.pi, n)
    x1 = nx / 2 + 300 * np.cos(ang)
    y1 = ny / 2 + 300 * np.sin(ang)
    x2 = rng.integers(0, nx, n) + 0.5
    y2 = rng.integers(0, ny, n) + 0.5
    segs = np.stack([x1, y1, x2, y2], axis=1)
    _, _, _, _, ints = xrd.trace_rays(torch.from_numpy(segs).cuda(), nx, ny, step,
                                      maps=torch.from_numpy(maps.reshape(-1)).cuda())
    torch.cuda.synchronize()
    ints = ints.cpu().numpy().astype(np.float64)
    ref = np.array([oracle.line_integrals(maps.reshape(2, -1).astype(np.float64), nx, ny, step,
                                          *s) for s in segs]).T
    rel = np.abs(ints - ref) / np.maximum(np.abs(ref), 1e-30)
    print(f"\n[line integrals] max rel err vs FP64 {rel.max():.2e}")
    assert rel.max() <= 2e-6


def test_forward_launch_shapes_and_autotune_cache(xrd, oracle, tmp_path):
    """Every projector launch shape the autotuner may pick (voxel chunk, view batch)
    meets the forward bars against the oracle; the tuned shape is kept on the context
    and a cache file gives it back without re-timing."""
    import torch
    from xrd_tomography_amd.synthetic import make_problem
    prob = make_problem("reference", flags=3, n_views=8)
    No, Co, Ao = oracle.forward(prob, nthreads=4)
    try:
        for kc, vpb in ((8, 2), (12, 8), (16, 32), (32, 1), (5, 3)):
            xrd.set_forward_launch(kc, vpb)
            Ng, Cg, Ag = _host(prob, *xrd.forward_problem(prob))
            _compare(f"reference, chunk {kc}, view batch {vpb}", Ng, Cg, Ag, No, Co, Ao)
        with pytest.raises(xrd.XrdError):
            xrd.set_forward_launch(33, 0)
        cache = str(tmp_path / "tune.json")
        r1 = xrd.autotune_forward(prob, sample_views=8, cache=cache)
        assert r1["cached"] is False and r1["voxels_per_chunk"] in (8, 12, 16)
        assert r1["views_per_batch"] in (2, 8, 32)
        r2 = xrd.autotune_forward(prob, sample_views=8, cache=cache)
        assert r2["cached"] is True
        assert (r2["voxels_per_chunk"], r2["views_per_batch"]) == (r1["voxels_per_chunk"],
                                                                  r1["views_per_batch"])
        a = xrd.forward_problem(prob)
        b = xrd.forward_problem(prob)
        torch.cuda.synchronize()
        for u, v in zip(a, b):
            assert torch.equal(u, v)
        print(f"\n[autotune] {r1}")
    finally:
        xrd.set_forward_launch(0, 0)
