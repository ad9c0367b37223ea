"""Benchmark: MLEM forward+adjoint iterations/s on the 256x256-voxel, 360-view workload.

A step is one MLEM update x <- x .* A'(y ./ (A x + c)) ./ sum(A,1)' (main.m:162)
over the explicit FP32 model matrix built on the GPU by xrd_forward
(5,621,760 rows x 1,024 columns, 23.0 GB, resident in HBM before timing).

Prints one JSON line.  Roofline: the fused forward+adjoint kernel streams A
once per step; algorithmic bytes per step = 4|A| + 8 NROW + 12 NCOL.
The CPU baseline is the FP64 oracle MLEM update on a row subsample of the same
matrix (whole views), timed on this host's cores and scaled to all rows.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "MLEM fwd+adj iters/s @256²×360 views; % HBM roofline; rel err vs FP64"
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the streaming pass, measured with rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE (profiles/mlem_fused_fan256_hbm_traffic.md): FETCH_SIZE scaled by 2.0 and
# WRITE_SIZE by 1.0, the factors that map each counter onto a 2 GiB copy kernel.  Each
# figure carries the hash of the streaming kernels' sources it was measured on
# (stream_sources_sha16); once those sources change it no longer applies and the bench
# line reports traffic null until it is measured again.
MEASURED_TRAFFIC = {
    ("fan256", "f32", "fused"): (23_076_127_744, "59d043db38ec3d07",
                                 "profiles/mlem_fused_fan256_hbm_traffic.md, launch 2,1"),
    ("fan256", "bf24", "fused"): (17_321_587_098, "59d043db38ec3d07",
                                  "profiles/mlem_fused_fan256_hbm_traffic.md, launch 3,1"),
}
ESIZE = {"f32": 4, "bf16": 2, "f16": 2, "bf24": 3}  # bytes per stored element of A


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def stream_sources_sha16() -> str:
    """First 16 hex digits of the SHA-256 of the streaming kernels' sources."""
    from xrd_tomography_amd.api import stream_kernels_tag
    return stream_kernels_tag()


def measured_traffic(key):
    """(bytes, profile note) of a stamped traffic measurement that still applies."""
    entry = MEASURED_TRAFFIC.get(key)
    if entry is None or entry[1] != stream_sources_sha16():
        return None, None
    return entry[0], entry[2]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="fan256")
    ap.add_argument("--dtype", default="f32", choices=["f32", "bf16", "f16", "bf24"],
                    help="storage of A: FP32, BF16/F16 with per-column scale, or 3-byte BF24")
    ap.add_argument("--unfused", action="store_true", help="plain two-pass A x then A' r")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shape of the fused pass")
    ap.add_argument("--launch", default=None,
                    help="B,R: fix the fused pass to B workgroups per CU and R rows per wave "
                         "(no autotuning)")
    ap.add_argument("--tune-cache", default=os.path.join(ROOT, "xrd_tomography_amd", "data",
                                                         "launch_cache_mi355x.json"),
                    help="JSON cache of autotuned launch shapes (projector and MLEM pass), "
                         "read and extended; '' disables it")
    ap.add_argument("--cpu-views", type=int, default=36, help="views in the CPU-baseline sample")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="repeat the CPU-baseline update until this much time has passed")
    ap.add_argument("--cpu-a-cols", type=int, default=8,
                    help="detector columns of the one-view CPU A-build sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--vjp-views", type=int, default=360,
                    help="views of the vjp256 workload (the coefficient-model VJP)")
    ap.add_argument("--vjp-basis", default="full", choices=["full", "2"],
                    help="vjp256 attenuation: the full 61-energy map or the two-basis form")
    ap.add_argument("--volume-slice", type=int, default=16,
                    help="slice of the vol192 workload's 32-slice volume")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 2 GiB device-to-device copy 10 times (calibrates HBM "
                         "byte counters: 2 GiB read + 2 GiB written per launch) and a "
                         "read-only sum over 20 GiB (a plain streaming-read rate)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "this build runs on exactly one GPU; the MLEM path would "
                          "shard rows by view with one all-reduce of the 1,024-column "
                          "numerator per step (not built)"}))
        sys.exit(2)

    import numpy as np
    import torch
    from xrd_tomography_amd import api
    from xrd_tomography_amd.synthetic import make_problem

    dev = torch.device("cuda:0")
    if args.copy_probe:
        n = (2 << 30) // 4
        a = torch.ones(n, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(10):
            b.copy_(a)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 10
        # a read-only reduction of a 20 GiB array (the size of the fan256 FP32 matrix)
        big = torch.ones(5 << 30, dtype=torch.float32, device=dev)
        big.sum()
        torch.cuda.synchronize()
        e0.record()
        for _ in range(5):
            big.sum()
        e1.record()
        torch.cuda.synchronize()
        ms_read = e0.elapsed_time(e1) / 5
        print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4,
                          "ms_per_copy": round(ms, 4),
                          "copy_GBps_read_plus_write": round(2 * n * 4 / ms / 1e6, 1),
                          "sum_probe_bytes_read": big.numel() * 4,
                          "ms_per_sum": round(ms_read, 4),
                          "sum_GBps_read": round(big.numel() * 4 / ms_read / 1e6, 1)}))
        return
    if args.workload == "batch48":
        run_batch(args, dev)
        return
    if args.workload == "vjp256":
        run_vjp(args, dev)
        return
    if args.workload == "vol192":
        run_volume(args, dev)
        return
    prob = make_problem(args.workload)
    g, d = prob.geometry, prob.disc
    nrow, ncol = prob.n_rows, prob.n_cols
    log(f"[bench] workload {args.workload}: rows {nrow}, cols {ncol}, non-air voxels "
        f"{int((prob.segmentation > 0).sum())}")

    # ---- build the model matrix on the GPU (not in the timed region), in the
    # projector launch shape the autotuner picked for this problem shape
    fwd_tuned = None
    if not args.no_autotune:
        fwd_tuned = api.autotune_forward(prob, cache=args.tune_cache or None)
        log(f"[bench] projector launch: {fwd_tuned}")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    N, C, A = api.forward_problem(prob, device=dev)
    torch.cuda.synchronize()
    a_build_s = time.perf_counter() - t0
    log(f"[bench] xrd_forward (N, C, A) built in {a_build_s:.2f} s")
    # rows of view 0 at a few evenly spaced detector columns, kept for the comparison
    # with the FP64 oracle forward that is also the CPU A-build baseline
    a_cols = np.linspace(0, g.n_cols - 1, max(1, min(args.cpu_a_cols, g.n_cols))).round()
    a_cols = np.unique(a_cols.astype(np.int64))
    sel = lambda X, w=None: (X.reshape((d.n_energy, g.n_rows, g.n_cols, g.n_views) +  # noqa: E731
                                       ((w,) if w else ()))
                             .index_select(2, torch.from_numpy(a_cols).to(dev))[:, :, :, 0]
                             .double().cpu().numpy().reshape((-1, w) if w else (-1,)))
    a_rows = (sel(N), sel(C), sel(A, ncol))
    A = A.view(nrow, ncol)
    N = N.reshape(-1)
    C = C.reshape(-1)
    if prob.photons_per_view == 0:
        # absolute photons (the paper's problem): y ~ Poisson(N + C * XS_grid_edges(6)/255)
        # and background C * XS_grid_edges(6)/255, as main.m:68,138,145
        scale, kappa = 1.0, prob.compton_scale
    else:
        # photon scale: expected N + c totals photons_per_view per view; the Compton
        # calibration factor (the XS_grid_edges(6)/255 of main.m:68) is set so that the
        # background is 30% of the diffraction signal.
        n_sum = float(N.double().sum())
        kappa = 0.3 * n_sum / max(float(C.double().sum()), 1e-300)
        scale = prob.photons_per_view * g.n_views / (n_sum * 1.3)
    A.mul_(scale)
    c = (C * (kappa * scale)).contiguous()
    expected = (N * scale + c).double().cpu().numpy()
    y = torch.from_numpy(np.random.default_rng(11).poisson(expected).astype(np.float32)).to(dev)
    del N
    col_scale = None
    if args.dtype == "bf24":
        Am = api.convert_matrix(A, api.BF24)
    elif args.dtype != "f32":
        dt = torch.bfloat16 if args.dtype == "bf16" else torch.float16
        Am, col_scale = api.convert_matrix_scaled(A, dt)
    else:
        Am = A

    # ---- MLEM timing
    x = torch.full((ncol,), 10.0, dtype=torch.float32, device=dev)
    plan = api.MlemPlan(Am, y, c, x, fused=not args.unfused, col_scale=col_scale)
    tuned = None
    if args.launch:
        b, r = (int(v) for v in args.launch.split(","))
        plan.set_launch(b, r)
        tuned = {"blocks_per_cu": b, "rows_per_wave": r, "fixed": True}
    elif not args.unfused and not args.no_autotune:
        tuned = plan.autotune(reps=2, cache=args.tune_cache or None)
        log(f"[bench] autotuned launch: {tuned}")
    use_graph = not args.no_graph
    plan.iterate(args.warmup, use_graph=use_graph)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record()
    plan.iterate(args.steps, use_graph=use_graph)
    ev1.record()
    torch.cuda.synchronize()
    ms_total = ev0.elapsed_time(ev1)
    ms_step = ms_total / max(args.steps, 1)
    iters_s = 1000.0 / ms_step
    log(f"[bench] {args.steps} MLEM steps: {ms_step:.3f} ms/step, {iters_s:.1f} iters/s")

    # the dominant kernel alone: the streaming pass over A, timed live with HIP events
    ms_pass = plan.time_pass(reps=max(5, args.steps // 5))
    log(f"[bench] streaming pass alone: {ms_pass:.3f} ms")
    esize = ESIZE[args.dtype]
    bytes_step = nrow * ncol * esize + nrow * 8 + ncol * 12
    passes = 2 if args.unfused else 1
    achieved = bytes_step * passes / (ms_pass * 1e-3) / 1e9
    traffic, traffic_src = measured_traffic((args.workload, args.dtype,
                                             "unfused" if args.unfused else "fused"))

    # ---- error vs FP64 and CPU baseline on a view subsample (whole views, all rows of them)
    rel_err = None
    cpu = None
    cpu_a = None
    a_build_err = None
    if not args.no_cpu:
        from oracle import oracle as O
        nv = max(1, min(args.cpu_views, g.n_views))
        views = np.linspace(0, g.n_views - 1, nv).round().astype(np.int64)
        # row r = v + V*(p + P*e): select every row whose view is in `views`
        rows = (views[None, :] + g.n_views * np.arange(nrow // g.n_views)[:, None]).reshape(-1)
        rows_t = torch.from_numpy(rows).to(dev)
        As = A.index_select(0, rows_t).contiguous()
        ys = y.index_select(0, rows_t).contiguous()
        cs = c.index_select(0, rows_t).contiguous()
        k_err = 5
        # GPU MLEM on the stored matrix (FP32, or BF16/F16 with its column scale) vs
        # the FP64 oracle on the FP32 matrix
        if args.dtype == "bf24":
            As_stored = api.convert_matrix(As, api.BF24)  # elementwise: same values as Am's rows
        elif col_scale is not None:
            As_stored = Am.index_select(0, rows_t).contiguous()
        else:
            As_stored = As
        xg = api.mlem(As_stored, ys, cs, k_err, col_scale=col_scale)
        del As_stored
        A64 = As.double().cpu().numpy()
        y64 = ys.double().cpu().numpy()
        c64 = cs.double().cpu().numpy()
        del As
        x0 = np.full(ncol, 10.0)
        xo = O.mlem(A64, y64, c64, x0, k_err)
        live = xo > 1e-6 * xo.max()
        rel_err = float(np.max(np.abs(xg.cpu().numpy()[live] - xo[live]) / xo[live]))
        threads = O.max_threads()
        # one update at a time until --cpu-seconds have passed (at least 5); median
        s64 = O.sensitivity(A64)
        xc = x0.copy()
        times = []
        t_start = time.perf_counter()
        while len(times) < 5 or time.perf_counter() - t_start < args.cpu_seconds:
            t0 = time.perf_counter()
            O.mlem_update(A64, y64, c64, s64, xc)
            times.append(time.perf_counter() - t0)
        t_cpu = float(np.median(times))
        frac = len(rows) / nrow
        cpu_iters_s = frac / t_cpu
        cpu = {"value": round(cpu_iters_s, 4), "unit": "iters/s", "cores": threads,
               "kind": "port",
               "sample": f"FP64 oracle MLEM update on {nv} of {g.n_views} views "
                         f"({len(rows)} rows x {ncol} cols, {A64.nbytes / 1e9:.2f} GB FP64): "
                         f"median of {len(times)} updates ({sum(times):.1f} s in total) "
                         f"= {t_cpu * 1e3:.1f} ms, scaled x{1 / frac:.1f} to all rows"}
        log(f"[bench] cpu baseline {cpu_iters_s:.3f} iters/s on {threads} threads; "
            f"GPU-vs-FP64 MLEM rel err after {k_err} steps {rel_err:.2e}")
        del A64
        # A-build baseline: the FP64 oracle forward (N, C and A) on view 0 and a few
        # detector columns of the same problem, scaled to all views and columns (the
        # cost per (view, column) row is the same); its rows are compared with the GPU's
        t0 = time.perf_counter()
        No, Co, Ao = O.forward(prob, views=[0], cols=a_cols, with_A=True)
        dt_a = time.perf_counter() - t0
        k_a = g.n_views * g.n_cols / len(a_cols)
        cpu_a = {"value": round(dt_a * k_a, 1), "unit": "s per A build (N, C and A)",
                 "cores": threads, "kind": "port",
                 "sample": f"FP64 oracle forward on view 0 x {len(a_cols)} detector "
                           f"columns of the same problem ({dt_a:.2f} s), scaled x{k_a:.0f}"}
        log(f"[bench] cpu A build {dt_a * k_a:.0f} s (scaled) against {a_build_s:.2f} s")
        a_build_err = _abuild_errors(a_rows, (No, Co, Ao))
        log(f"[bench] A build vs FP64 on those rows: {a_build_err}")

    out = {
        "metric": METRIC,
        "value": round(iters_s, 3),
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": ("the reference's own inputs (phantom_generator, measured patterns, "
                 "SpekCalc spectrum) through the builders, Poisson counts"
                 if args.workload == "paper100" else "synthetic"),
        "config": {
            "workload": f"{args.workload}: {g.nx}x{g.ny} voxels, {g.n_views} views, "
                        f"{g.n_cols} pixels x {d.n_energy} energies, {d.n_q} q-bins x "
                        f"{d.n_mat} materials; explicit A {nrow} x {ncol} ({args.dtype}), "
                        f"{'unfused two-pass' if args.unfused else 'fused single-pass'} "
                        f"MLEM{', hipGraph replay' if use_graph else ''}",
            "rows": nrow, "cols": ncol,
            "launch": tuned,
            "forward_launch": fwd_tuned,
        },
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "traffic_source": traffic_src,
                     "kernel_ms": round(ms_pass, 4),
                     "algorithmic_bytes": bytes_step * passes},
        "cpu_baseline": cpu,
        "rel_err_vs_fp64": rel_err,
        "a_build_s": round(a_build_s, 3),
        "cpu_a_build": cpu_a,
        "a_build_rel_err_vs_fp64": a_build_err,
        "bytes_per_step": bytes_step * passes,
    }
    plan.close()
    print(json.dumps(out), flush=True)


def _abuild_errors(gpu, ora):
    """Worst relative error of the GPU's N, C (per row, rows above 1e-9 of the largest)
    and A (per-row Frobenius) against the FP64 oracle's rows."""
    import numpy as np
    out = {}
    for name, g_, o_ in zip(("N", "C"), gpu[:2], ora[:2]):
        live = np.abs(o_) > 1e-9 * np.abs(o_).max()
        out[name] = float(np.max(np.abs(g_[live] - o_[live]) / np.abs(o_[live]), initial=0.0))
    rn = np.linalg.norm(ora[2], axis=1)
    live = rn > 1e-9 * rn.max()
    out["A_row"] = float(np.max(np.linalg.norm(gpu[2] - ora[2], axis=1)[live] / rn[live],
                                initial=0.0))
    return {k: float(f"{v:.3e}") for k, v in out.items()}


def _scaled_counts(api, prob, dev):
    """A build for one problem, scaled like the headline workload: expected N + c totals
    photons_per_view per view and the Compton background is 30% of the diffraction."""
    import numpy as np
    import torch
    N, C, A = api.forward_problem(prob, device=dev)
    A = A.view(prob.n_rows, prob.n_cols)
    N, C = N.reshape(-1), C.reshape(-1)
    n_sum = float(N.double().sum())
    kappa = 0.3 * n_sum / max(float(C.double().sum()), 1e-300)
    scale = prob.photons_per_view * prob.geometry.n_views / (n_sum * 1.3)
    A.mul_(scale)
    c = (C * (kappa * scale)).contiguous()
    expected = (N * scale + c).double().cpu().numpy()
    y = torch.from_numpy(np.random.default_rng(11).poisson(expected).astype(np.float32)).to(dev)
    return A, y, c


def run_batch(args, dev):
    """48 phantoms of uneven size and view count: batched persistent MLEM (one launch
    per pass over all phantoms, hipGraph replay, autotuned launch shape) against the
    same phantoms one at a time through the plain unfused FP32 path."""
    import numpy as np
    import torch
    from xrd_tomography_amd import api
    from xrd_tomography_amd.synthetic import make_batch_problems
    from oracle import oracle as O

    probs = make_batch_problems()
    t0 = time.perf_counter()
    data = []
    for k, p in enumerate(probs):
        A, y, c = _scaled_counts(api, p, dev)
        data.append((A, y, c))
        if k % 8 == 7:
            log(f"[batch] built {k + 1}/{len(probs)} model matrices")
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    n_it = 30
    esize = ESIZE[args.dtype]
    stored = []
    for A, y, c in data:
        if args.dtype == "f32":
            stored.append((A, y, c, torch.full((A.shape[1],), 10.0, device=dev)))
        elif args.dtype == "bf24":
            stored.append((api.convert_matrix(A, api.BF24), y, c,
                           torch.full((A.shape[1],), 10.0, device=dev)))
        else:
            dt = torch.bfloat16 if args.dtype == "bf16" else torch.float16
            A16, sc = api.convert_matrix_scaled(A, dt)
            stored.append((A16, y, c, torch.full((A.shape[1],), 10.0, device=dev), sc))
    bytes_it = sum(p[0].shape[0] * p[0].shape[1] * esize + p[0].shape[0] * 8 + p[0].shape[1] * 12
                   for p in stored)
    batch = api.MlemBatch(stored)
    tuned = batch.autotune(reps=2)
    # algorithmic bytes per row-span group (one persistent launch each)
    spans = (256, 384, 512, 768, 1024, 1536, 2048)
    per_span = {}
    for p in stored:
        nr, nc = p[0].shape
        sp = next(v for v in spans if nc <= v)
        b = per_span.setdefault(sp, [0, 0])
        b[0] += 1
        b[1] += nr * nc * esize + nr * 8 + nc * 12
    for gr in tuned["groups"]:
        gr["problems"], gr["bytes"] = per_span.get(gr["span"], (0, 0))
    log(f"[batch] autotuned launch: {tuned}")
    for p in stored:
        p[3].fill_(10.0)
    batch.iterate(1, use_graph=True)  # capture
    for p in stored:
        p[3].fill_(10.0)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    batch.iterate(n_it, use_graph=True)
    e1.record()
    torch.cuda.synchronize()
    ms_it = e0.elapsed_time(e1) / n_it
    ms_pass = batch.time_pass(reps=10)
    log(f"[batch] {args.dtype} batched: {ms_it:.2f} ms per iteration over {len(stored)} phantoms "
        f"({bytes_it / ms_it / 1e6:.0f} GB/s); streaming pass alone {ms_pass:.2f} ms "
        f"({bytes_it / ms_pass / 1e6:.0f} GB/s)")

    # baselines: the same phantoms one at a time (a) through the fused single-matrix
    # plan on the same stored matrices, (b) through the plain unfused FP32 path
    def one_at_a_time(plan_args, fused):
        xs = [torch.full((pa[0].shape[1],), 10.0, device=dev) for pa in plan_args]
        plans = [api.MlemPlan(pa[0], pa[1], pa[2], x, fused=fused,
                              col_scale=pa[4] if len(pa) > 4 else None)
                 for pa, x in zip(plan_args, xs)]
        for pl in plans:
            pl.iterate(1)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(n_it):
            for pl in plans:
                pl.iterate(1)
        e1.record()
        torch.cuda.synchronize()
        for pl in plans:
            pl.close()
        return e0.elapsed_time(e1) / n_it

    ms_single_fused = one_at_a_time(stored, True)
    log(f"[batch] one at a time, fused {args.dtype}: {ms_single_fused:.2f} ms per iteration")
    ms_base = one_at_a_time([(A, y, c) for A, y, c in data], False)
    log(f"[batch] one at a time, unfused FP32: {ms_base:.2f} ms per iteration")

    # error vs FP64 on the smallest phantoms (oracle MLEM, 30 updates)
    order = sorted(range(len(data)), key=lambda k: data[k][0].numel())[:4]
    worst = 0.0
    host = []
    for k in order:
        A, y, c = data[k]
        h = (A.double().cpu().numpy(), y.double().cpu().numpy(), c.double().cpu().numpy())
        host.append(h)
        xo = O.mlem(*h, np.full(A.shape[1], 10.0), n_it)
        xb = stored[k][3].cpu().numpy()
        live = xo > 1e-6 * xo.max()
        worst = max(worst, float(np.max(np.abs(xb[live] - xo[live]) / xo[live])))
    batch.close()
    # CPU baseline: FP64 oracle MLEM updates of the same 4 phantoms, one pass over the 4
    # at a time until --cpu-seconds have passed (at least 5 passes); the median pass is
    # scaled by the matrix elements of all 48 over those of the 4 (the update streams
    # every element once)
    cpu = None
    if not args.no_cpu:
        sens = [O.sensitivity(h[0]) for h in host]
        xs_c = [np.full(h[0].shape[1], 10.0) for h in host]
        times = []
        t_start = time.perf_counter()
        while len(times) < 5 or time.perf_counter() - t_start < args.cpu_seconds:
            t0 = time.perf_counter()
            for h, s_, x_ in zip(host, sens, xs_c):
                O.mlem_update(h[0], h[1], h[2], s_, x_)
            times.append(time.perf_counter() - t0)
        t_cpu = float(np.median(times))
        el_sample = sum(h[0].size for h in host)
        el_all = sum(p[0].shape[0] * p[0].shape[1] for p in data)
        cpu = {"value": round(el_sample / el_all / t_cpu, 5),
               "unit": "iters/s (all 48 phantoms per iteration)", "cores": O.max_threads(),
               "kind": "port",
               "sample": f"FP64 oracle MLEM update of the 4 smallest phantoms "
                         f"({el_sample} matrix elements): median of {len(times)} passes "
                         f"({sum(times):.1f} s in total) = {t_cpu * 1e3:.1f} ms, scaled "
                         f"x{el_all / el_sample:.1f} to the {el_all} elements of all 48"}
        log(f"[batch] cpu baseline {cpu['value']:.4f} iters/s on {cpu['cores']} threads")
    del host
    out = {
        "metric": "batched MLEM fwd+adj iters/s (48 phantoms); % HBM roofline; rel err vs FP64",
        "value": round(1000.0 / ms_it, 3), "unit": "iters/s (all 48 phantoms per iteration)",
        "n_gpus": 1, "steps": n_it, "warmup": 1, "ms_per_step": round(ms_it, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": args.dtype,
        "data": "synthetic",
        "config": {"workload": "batch48: 48 phantoms, sizes {64..256}^2 voxels, 90-360 views, "
                               "64 pixels x 61 energies, 128 q-bins x 2-10 materials",
                   "rows": int(sum(p[0].shape[0] for p in stored)), "launch": tuned},
        "roofline": {"bound": "hbm", "achieved": round(bytes_it / ms_pass / 1e6, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(bytes_it / ms_pass / 1e6 / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel_ms": round(ms_pass, 4), "algorithmic_bytes": bytes_it,
                     "iteration_GBps": round(bytes_it / ms_it / 1e6, 1)},
        "one_at_a_time_fused": {"ms_per_step": round(ms_single_fused, 3),
                                "speedup": round(ms_single_fused / ms_it, 3)},
        "baseline_one_at_a_time_fp32_unfused": {"ms_per_step": round(ms_base, 3),
                                                "speedup": round(ms_base / ms_it, 3)},
        "rel_err_vs_fp64": worst, "rel_err_sample": "4 smallest phantoms, 30 updates",
        "a_build_s": round(build_s, 1), "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def run_volume(args, dev):
    """One slice of the 192x192x32-voxel volume at 2 mm (720 views, 256 pixels x 61
    energies, 10 materials): the slice's explicit model matrix (the attenuation legs are
    3-D segments through the whole volume), its counts against the FP64 oracle on one
    view, and the 30-iteration MLEM image with A stored in FP32, BF24, BF16 and F16
    against the FP64 oracle MLEM on the same FP32 matrix (36 of the 720 views).  The
    counts are also reported with the form-factor and attenuation tables in BF16 with
    a per-material FP32 scale (xrd_forward_bf16), whose A build is timed as well."""
    import numpy as np
    import torch
    from xrd_tomography_amd import api
    from xrd_tomography_amd.synthetic import decoded_bf16_tables, make_volume_problem
    from oracle import oracle as O

    nz, sl = 32, args.volume_slice
    vp = make_volume_problem("fan256", n_slices=nz, seed=7, nx=192, voxel_mm=2.0, n_views=720,
                             n_mat=10, mlem_iterations=30)
    g, d = vp.geometry, vp.disc
    nrow, ncol = vp.n_rows, vp.n_cols
    log(f"[vol] slice {sl} of {g.nx}x{g.ny}x{nz}: rows {nrow}, cols {ncol} "
        f"({nrow * ncol * 4 / 1e9:.1f} GB FP32), scattering voxels "
        f"{int((vp.segmentation[:, :, sl] > 0).sum())}")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    N, C, A = api.forward_volume_slice(vp, nz, sl)
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    log(f"[vol] A build {t_build:.2f} s")
    A = A.view(nrow, ncol)
    N, C = N.reshape(-1), C.reshape(-1)
    n_sum = float(N.double().sum())
    scale = vp.photons_per_view * g.n_views / (n_sum * 1.3)
    kappa = 0.3 * n_sum / max(float(C.double().sum()), 1e-300)
    A.mul_(scale)
    c = (C * (kappa * scale)).contiguous()
    expected = (N * scale + c).double().cpu().numpy()
    y = torch.from_numpy(np.random.default_rng(11).poisson(expected).astype(np.float32)).to(dev)

    # the same slice with BF16 tables (xrd_forward_bf16: form factors and maps as
    # bfloat16 with per-material FP32 scales, decoded by the kernels)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Nq, _, Aq = api.forward_problem_bf16(vp, nz, sl)
    torch.cuda.synchronize()
    t_build_bf16 = time.perf_counter() - t0
    del Aq
    log(f"[vol] A build with BF16 tables {t_build_bf16:.2f} s")
    Nq = Nq.view(d.n_energy, g.n_rows * g.n_cols, g.n_views)[:, :, 0].reshape(-1)
    Nq = Nq.double().cpu().numpy()

    # counts against the FP64 oracle on one view (all pixels, energies)
    cnt_err = cnt_err_bf16 = cnt_err_bf16_kernel = None
    if not args.no_cpu:
        t0 = time.perf_counter()
        No, Co, _ = O.forward_slice(vp, nz, sl, views=[0], with_A=False)
        t_oracle = time.perf_counter() - t0
        Ng = N.view(d.n_energy, g.n_rows * g.n_cols, g.n_views)[:, :, 0].reshape(-1)
        Ng = Ng.double().cpu().numpy()
        live = No > 1e-9 * No.max()
        cnt_err = float(np.max(np.abs(Ng[live] - No[live]) / No[live]))
        log(f"[vol] counts of view 0 against FP64: max rel err {cnt_err:.2e} "
            f"(oracle {t_oracle:.1f} s)")
        # BF16 tables against the FP64 counts of the FP32 tables (storage and kernel
        # error) and of the decoded tables (the kernels' own error)
        cnt_err_bf16 = float(np.max(np.abs(Nq[live] - No[live]) / No[live]))
        Nd, _, _ = O.forward_slice(decoded_bf16_tables(vp), nz, sl, views=[0], with_A=False)
        lived = Nd > 1e-9 * Nd.max()
        cnt_err_bf16_kernel = float(np.max(np.abs(Nq[lived] - Nd[lived]) / Nd[lived]))
        log(f"[vol] counts of view 0 with BF16 tables against FP64: max rel err "
            f"{cnt_err_bf16:.2e} (FP32 tables), {cnt_err_bf16_kernel:.2e} (decoded tables)")
    del Nq

    # MLEM image per storage format on 36 of the 720 views, against FP64 MLEM
    views = np.linspace(0, g.n_views - 1, 36).round().astype(np.int64)
    rows = (views[None, :] + g.n_views * np.arange(nrow // g.n_views)[:, None]).reshape(-1)
    rows_t = torch.from_numpy(rows).to(dev)
    As = A.index_select(0, rows_t).contiguous()
    ys, cs = y.index_select(0, rows_t).contiguous(), c.index_select(0, rows_t).contiguous()
    n_it = 30
    x_ref = None
    if not args.no_cpu:
        x_ref = O.mlem(As.double().cpu().numpy(), ys.double().cpu().numpy(),
                       cs.double().cpu().numpy(), np.full(ncol, 10.0), n_it)
    study = {}
    for name in ("f32", "bf24", "bf16", "f16"):
        sc = None
        if name == "f32":
            Am = As
        elif name == "bf24":
            Am = api.convert_matrix(As, api.BF24)
        else:
            Am, sc = api.convert_matrix_scaled(As, torch.bfloat16 if name == "bf16" else torch.float16)
        xg = api.mlem(Am, ys, cs, n_it, col_scale=sc).cpu().numpy()
        err = None
        if x_ref is not None:
            live = x_ref > 1e-6 * x_ref.max()
            err = float(np.max(np.abs(xg[live] - x_ref[live]) / x_ref[live]))
        study[name] = err
        log(f"[vol] MLEM {n_it} its, A in {name}: max rel err vs FP64 {err}")
        del Am
    # full-slice MLEM speed with FP32 A
    x = torch.full((ncol,), 10.0, device=dev)
    plan = api.MlemPlan(A, y, c, x)
    tuned = None if args.no_autotune else plan.autotune(reps=2, cache=args.tune_cache or None)
    log(f"[vol] MLEM launch: {tuned}")
    x.fill_(10.0)
    plan.iterate(2, use_graph=True)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    plan.iterate(n_it, use_graph=True)
    e1.record()
    torch.cuda.synchronize()
    ms_it = e0.elapsed_time(e1) / n_it
    bytes_it = nrow * ncol * 4 + nrow * 8 + ncol * 12
    plan.close()
    out = {
        "metric": "volume slice (192x192x32 at 2 mm, 720 views): A build s, MLEM iters/s, "
                  "rel err of counts and of the 30-iteration MLEM image vs FP64",
        "value": round(1000.0 / ms_it, 3), "unit": "MLEM iters/s (one slice, FP32 A)",
        "n_gpus": 1, "steps": n_it, "warmup": 2, "ms_per_step": round(ms_it, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"vol192: slice {sl} of 192x192x{nz} voxels, 720 views, 256 "
                               f"pixels x 61 energies, 128 q-bins x 10 materials; 3-D "
                               f"trilinear attenuation legs", "rows": nrow, "cols": ncol,
                   "launch": tuned},
        "a_build_s": round(t_build, 3),
        "a_build_bf16_tables_s": round(t_build_bf16, 3),
        "roofline": {"bound": "hbm", "achieved": round(bytes_it / ms_it / 1e6, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(bytes_it / ms_it / 1e6 / HBM_PEAK_GBS, 4), "traffic": None},
        "counts_rel_err_vs_fp64_view0": cnt_err,
        "counts_rel_err_bf16_tables_vs_fp64_view0": cnt_err_bf16,
        "counts_rel_err_bf16_tables_vs_fp64_decoded_view0": cnt_err_bf16_kernel,
        "mlem30_rel_err_vs_fp64": study,
        "mlem_sample": "36 of 720 views of the slice; FP64 oracle MLEM on the FP32 matrix",
    }
    print(json.dumps(out), flush=True)


def run_vjp(args, dev):
    """Coefficient model on the 256x256-voxel phantom (16 coefficients per voxel on a
    Gaussian-bump basis, attenuation as the 256^2 x 61 mu(E) map): one forward, the
    Poisson NLL, the VJP with respect to the coefficients alone and with respect to
    coefficients and maps.  Rate = detector rays (view, pixel, voxel) per second; each
    ray's line integrals are recomputed.  Full-size check: <J x, u> = <x, J' u> (the
    counts are linear in the coefficients)."""
    import numpy as np
    import torch
    from xrd_tomography_amd import api
    from xrd_tomography_amd.synthetic import make_coeff_problem, make_problem
    from oracle import oracle as O

    full = args.vjp_basis == "full"
    prob = make_problem("fan256", n_views=args.vjp_views)
    cp = make_coeff_problem(prob, K=16, full_mu=full)
    g, d = prob.geometry, prob.disc
    nvox = int((prob.segmentation > 0).sum())
    rays = g.n_views * g.n_cols * g.n_rows * nvox
    sp, seg, coef, basis, maps, beta = api.coeff_problem_tensors(cp)
    kw = dict(geometry=g, disc=d)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(fn):
        torch.cuda.synchronize()
        e0.record()
        out = fn()
        e1.record()
        torch.cuda.synchronize()
        return out, e0.elapsed_time(e1) / 1000.0

    log(f"[vjp] fan256 {g.n_views} views, {nvox} voxels, B={cp.B}: {rays} detector rays")
    y, t_fwd = timed(lambda: api.forward_coeffs(sp, seg, coef, basis, maps, beta, **kw))
    log(f"[vjp] forward {t_fwd:.2f} s")
    yobs = torch.poisson(y.double() + 0.05 * float(y.mean())).float().reshape(-1)
    (loss, w), t_nll = timed(lambda: api.poisson_nll(y.reshape(-1).contiguous(), yobs))
    (gc, _), t_coef = timed(lambda: api.vjp_coeffs(sp, seg, coef, basis, maps, beta, w,
                                                   want_maps=False, **kw))
    log(f"[vjp] VJP w.r.t. coefficients {t_coef:.2f} s")
    (gc2, gm), t_all = timed(lambda: api.vjp_coeffs(sp, seg, coef, basis, maps, beta, w, **kw))
    log(f"[vjp] VJP w.r.t. coefficients and maps {t_all:.2f} s")
    # dot-product identity at full size: counts are linear in the coefficients
    gen = torch.Generator(device="cpu").manual_seed(5)
    xr = torch.rand(coef.shape, generator=gen).to(dev) * (seg.reshape(-1, 1) > 0)
    jx = api.forward_coeffs(sp, seg, xr.float().contiguous(), basis, maps, beta, **kw)
    a = float((jx.double().reshape(-1) * w.double()).sum())
    b = float((xr.double() * gc.double()).sum())
    dot_rel = abs(a - b) / max(abs(a), abs(b))
    log(f"[vjp] dot test <Jx,w> {a:.9e} <x,J'w> {b:.9e} rel {dot_rel:.2e}")
    # CT transmission channel on the same maps: forward and map VJP
    t_ct, t_ctv = None, None
    if g.n_rows == 1:
        tct, t_ct = timed(lambda: api.forward_ct(sp, maps, beta, **kw))
        wct = torch.randn(tct.shape, generator=gen).to(dev)
        _, t_ctv = timed(lambda: api.vjp_ct(sp, maps, beta, wct, **kw))
        log(f"[vjp] CT forward {t_ct * 1e3:.2f} ms, CT map VJP {t_ctv * 1e3:.2f} ms "
            f"({g.n_views * g.n_cols} rays)")

    # CPU baseline: the FP64 oracle on the central pixel of one view, rays/s
    cpu = None
    if not args.no_cpu:
        ps = make_problem("fan256", n_views=1, n_cols=1)
        cps = make_coeff_problem(ps, K=16, full_mu=full)
        ws = np.random.default_rng(1).normal(size=ps.n_rows)
        t0 = time.perf_counter()
        O.vjp_coeffs(cps, ws)
        dt = time.perf_counter() - t0
        rs = ps.geometry.n_cols * int((ps.segmentation > 0).sum())
        cpu = {"value": round(rs / dt, 1), "unit": "detector rays/s (VJP w.r.t. coefficients "
               "and maps)", "cores": os.cpu_count(), "kind": "port",
               "sample": f"FP64 oracle VJP, fan256 phantom, 1 view x 1 central pixel "
                         f"({rs} rays, {dt:.1f} s)"}
        log(f"[vjp] CPU oracle {rs / dt:.0f} rays/s")
    out = {
        "metric": "coefficient-model VJP (w.r.t. 16 coefficients/voxel and the mu(E) map) "
                  "detector rays/s @256²×360 views",
        "value": round(rays / t_all, 1), "unit": "detector rays/s", "n_gpus": 1, "steps": 1,
        "warmup": 0, "ms_per_step": round(t_all * 1000.0, 1), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"vjp256: 256x256 voxels, {g.n_views} views, 256 pixels x 61 "
                               f"energies, 128 q-bins, K=16, B={cp.B}", "rays": rays},
        "forward_s": round(t_fwd, 3), "vjp_coef_s": round(t_coef, 3),
        "vjp_coef_maps_s": round(t_all, 3), "poisson_nll_s": round(t_nll, 4),
        "coef_grad_same_in_both": bool(torch.equal(gc, gc2)),
        "dot_test_rel": dot_rel, "ct_forward_s": t_ct, "ct_vjp_s": t_ctv, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
