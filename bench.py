"""H-matrix x multi-vector product benchmark on one MI355X.

Workload (the headline metric): N = 262,144 points uniform in [0,1)^3 (seeded),
kernel 1/r with 0 on the diagonal, cluster-tree leaf 64, eta 2, ACA+ tol 1e-6,
16 right-hand sides N(0,1); U/V and near field stored in FP32 with FP32
accumulation.  One step = one call of hmx_matvec over all 16 right-hand sides
with X and Y resident in HBM.  `value` is H-matvecs/s (16 / step time).

Also reported: the plain FP64 single-RHS GPU product (the drop-in mul!), FP64
and BF16 16-RHS, relative errors against an FP64 direct sum on 4096 sampled
rows, the HBM roofline of the dominant kernel (HIP events on the launch
stream), and the CPU restatement of mult.jl:6-32 timed on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hmatrixgpu.jl_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "H-matvecs/s + HBM GB/s, N=262k, 16 RHS, mixed vs FP64 at rel.err≤1e-5"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--n", type=int, default=262144)
    p.add_argument("--nrhs", type=int, default=16)
    p.add_argument("--leaf", type=int, default=64)
    p.add_argument("--eta", type=float, default=2.0)
    p.add_argument("--tol", type=float, default=1e-6)
    p.add_argument("--sample-rows", type=int, default=4096)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-sample-gb", type=float, default=4.0,
                   help="FP64 payload bytes of the CPU-baseline sample")
    p.add_argument("--profile-run", action="store_true",
                   help="workload for rocprofv3: a 4 GiB copy and a 4 GiB fill (counter "
                        "calibration), then 5 FP32 16-RHS products; prints no JSON")
    return p.parse_args()


PMC_FILE = os.path.join(ROOT, "profiles", "leaf_panels_fp32_16rhs_n262144.md")


def measured_traffic():
    """HBM bytes per launch of the leaf-panel kernel from the committed rocprofv3
    counter summary: FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by
    what the same counter read for a copy / fill of known size."""
    if not os.path.exists(PMC_FILE):
        return None, None
    vals = {}
    for line in open(PMC_FILE):
        if line.startswith("    ") and "=" in line:
            k, v = line.split("=", 1)
            try:
                vals[k.strip()] = float(v.split()[0])
            except ValueError:
                pass
    try:
        rscale = vals["copy_read_bytes"] / (vals["copy_fetch_kib"] * 1024.0)
        wscale = vals["fill_write_bytes"] / (vals["fill_write_kib"] * 1024.0)
        traffic = (vals["leaf_fetch_kib"] * 1024.0 * rscale + vals["leaf_write_kib"] * 1024.0 * wscale)
        return traffic, {"fetch_scale": rscale, "write_scale": wscale, "source": os.path.relpath(PMC_FILE, ROOT)}
    except KeyError:
        return None, None


def profile_run(args):
    import numpy as np
    import torch

    import hmatrixgpu as H
    dev = torch.device("cuda", 0)
    a = torch.empty(2 ** 29, dtype=torch.float64, device=dev)  # 4 GiB
    a.fill_(1.0)  # calibration: 4 GiB written
    b = torch.empty_like(a)
    b.copy_(a)  # calibration: 4 GiB read + 4 GiB written
    torch.cuda.synchronize()
    log("calibration copy/fill of 4 GiB done")
    del a, b
    n = args.n
    pts = np.random.default_rng(3).random((n, 3)).T.copy()
    X = H.ClusterTree(pts, max_points_per_leaf=args.leaf)
    hm = H.HMatrix(H.LaplaceKernel(pts), X, X, eta=args.eta, eps=args.tol, storage=("f32",))
    log("built: %s" % json.dumps(hm.info()))
    Xg = torch.randn((args.nrhs, n), dtype=torch.float32, device=dev).t()
    Y = torch.empty((args.nrhs, n), dtype=torch.float32, device=dev).t()
    for _ in range(5):
        H.mul_(Y, hm, Xg, storage="f32")
    torch.cuda.synchronize()
    log("5 FP32 %d-RHS products done" % args.nrhs)


def dense_reference_rows(torch, pts, rows, X64, dev):
    """FP64 direct sum  sum_j K(x_r, x_j) X[j, :]  for the sampled rows, on the GPU."""
    P = torch.tensor(pts.T, dtype=torch.float64, device=dev)  # N x 3
    R = P[torch.tensor(rows, device=dev)]
    out = torch.zeros((len(rows), X64.shape[1]), dtype=torch.float64, device=dev)
    chunk = 32768
    for j0 in range(0, P.shape[0], chunk):
        d = R[:, None, :] - P[None, j0:j0 + chunk, :]
        r = torch.sqrt((d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2])
        K = torch.where(r == 0, torch.zeros_like(r), 1.0 / r)
        out += K @ X64[j0:j0 + chunk]
    return out


def time_calls(torch, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def cpu_baseline(hm, info, sample_gb):
    """mult.jl:6-32 restated in C (the oracle), 1 thread, FP64, on a bounded
    sample: the first fraction of the dense and low-rank block lists (DFS
    order), scaled to one full product by FP64 payload bytes."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O  # the oracle: CPU baseline only

    lay = hm.layout(True)
    di = np.ascontiguousarray(lay["dense_block_indices"].T)
    ui = np.ascontiguousarray(lay["U_block_indices"].T)
    vi = np.ascontiguousarray(lay["V_block_indices"].T)
    D, U, V = lay["dense_blocks"], lay["U_matrices"], lay["V_matrices"]
    total = 8.0 * (D.size + U.size + V.size)
    frac = min(1.0, sample_gb * 1e9 / total)
    nd = max(1, int(round(frac * di.shape[0])))
    na = int(round(frac * ui.shape[0]))
    dsz = (di[:nd, 1] - di[:nd, 0] + 1) * (di[:nd, 3] - di[:nd, 2] + 1)
    ranks = ui[:na, 3] - ui[:na, 2] + 1
    usz = (ui[:na, 1] - ui[:na, 0] + 1) * ranks
    vcols = np.array([(vi[int(ui[b, 2]) - 1, 1] - vi[int(ui[b, 2]) - 1, 0] + 1) if ranks[b] > 0 else 0
                      for b in range(na)])
    sample_bytes = 8.0 * (dsz.sum() + usz.sum() + (vcols * ranks).sum())
    n = hm.n
    x = np.random.default_rng(9).standard_normal(n)
    xo = np.zeros(n)
    y = np.zeros(hm.m)
    sm = lay["source_index_map"]
    args = (O.pi64(sm), O.pi64(di), O.p64(D), 0, nd, O.pi64(ui), O.p64(U if U.size else np.zeros(1)),
            0, na, O.pi64(vi), O.p64(V if V.size else np.zeros(1)), O.p64(x), O.p64(xo), n,
            O.p64(y))
    for _ in range(2):
        O.lib.orc_mul_layout_partial(*args)
    ts = []
    for _ in range(10):
        y[:] = 0
        t0 = time.perf_counter()
        O.lib.orc_mul_layout_partial(*args)
        ts.append(time.perf_counter() - t0)
    t_sample = float(np.median(ts))
    t_full = t_sample * total / sample_bytes
    return {"value": 1.0 / t_full, "unit": "H-matvecs/s", "cores": 1, "kind": "port",
            "sample": "oracle C restatement of mult.jl:6-32 (FP64, 1 thread, 1 RHS) over the "
                      "first %d of %d dense and %d of %d low-rank blocks (%.2f GB of %.2f GB FP64 "
                      "payload), median of 10 after 2 warm-ups, scaled by payload bytes"
                      % (nd, di.shape[0], na, ui.shape[0], sample_bytes / 1e9, total / 1e9),
            "sample_seconds": t_sample, "host_threads_available": os.cpu_count()}


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print("bench.py: this build measures one GPU only (the H-matrix product is sharded "
              "by target rows in the design, but no multi-GPU path is built); run with --gpus 1",
              flush=True)
        sys.exit(2)
    if args.profile_run:
        profile_run(args)
        return
    import numpy as np
    import torch

    import hmatrixgpu as H

    dev = torch.device("cuda", 0)
    n, nrhs = args.n, args.nrhs
    pts = np.random.default_rng(3).random((n, 3)).T.copy()  # seed 3, (3, N)
    log("building cluster tree + H-matrix N=%d leaf=%d eta=%g tol=%g" % (n, args.leaf, args.eta,
                                                                          args.tol))
    t0 = time.perf_counter()
    X = H.ClusterTree(pts, max_points_per_leaf=args.leaf)
    hm = H.HMatrix(H.LaplaceKernel(pts), X, X, eta=args.eta, eps=args.tol,
                   storage=("f64", "f32", "bf16"))
    build_s = time.perf_counter() - t0
    info = hm.info()
    log("built in %.1fs: %s" % (build_s, json.dumps(info)))

    rng = np.random.default_rng(4)
    X64 = torch.tensor(rng.standard_normal((n, nrhs)), dtype=torch.float64, device=dev)
    Xc64 = X64.t().contiguous().t()  # column-major n x nrhs
    Xc32 = Xc64.float().t().contiguous().t()
    Y32 = torch.empty((nrhs, n), dtype=torch.float32, device=dev).t()
    Ybf = torch.empty((nrhs, n), dtype=torch.float32, device=dev).t()
    Y64 = torch.empty((nrhs, n), dtype=torch.float64, device=dev).t()
    x1 = Xc64[:, 0].contiguous()
    y1 = torch.empty(n, dtype=torch.float64, device=dev)

    # ---- headline: FP32 storage, 16 RHS ---------------------------------------------
    log("timing FP32 16-RHS: warmup %d, steps %d" % (args.warmup, args.steps))
    f32 = lambda: H.mul_(Y32, hm, Xc32, storage="f32")  # noqa: E731
    t32 = time_calls(torch, f32, args.steps, args.warmup)
    # per-stage kernel durations with HIP events on the launch stream
    hm.profile(True)
    for _ in range(max(3, args.steps // 2)):
        f32()
    torch.cuda.synchronize()
    ms, launches = hm.profile_read(reset=True)
    hm.profile(False)
    stage_avg = [float(ms[k] / launches[k]) if launches[k] else 0.0 for k in range(4)]

    log("timing FP64 1-RHS, FP64 16-RHS, BF16 16-RHS")
    t64_1 = time_calls(torch, lambda: H.mul_(y1, hm, x1, storage="f64"), args.steps, args.warmup)
    t64_16 = time_calls(torch, lambda: H.mul_(Y64, hm, Xc64, storage="f64"), args.steps,
                        args.warmup)
    tbf = time_calls(torch, lambda: H.mul_(Ybf, hm, Xc32, storage="bf16"), args.steps, args.warmup)

    # ---- accuracy vs an FP64 direct sum on sampled rows ---------------------------------
    rows = np.sort(np.random.default_rng(5).choice(n, size=min(args.sample_rows, n), replace=False))
    ref = dense_reference_rows(torch, pts, rows, X64, dev)
    ridx = torch.tensor(rows, device=dev)

    def relerr(Y):
        d = Y[ridx].double() - ref
        return float(torch.linalg.norm(d) / torch.linalg.norm(ref))

    H.mul_(y1, hm, x1, storage="f64")
    torch.cuda.synchronize()
    err = {"fp32_16rhs": relerr(Y32), "bf16_16rhs": relerr(Ybf), "fp64_16rhs": relerr(Y64),
           "fp64_1rhs": float(torch.linalg.norm(y1[ridx] - ref[:, 0]) / torch.linalg.norm(ref[:, 0]))}
    log("rel err vs FP64 direct sum: %s" % json.dumps(err))

    # ---- bytes ---------------------------------------------------------------------
    nnzD, nnzU, nnzV = info["nnz_D"], info["nnz_U"], info["nnz_V"]

    def balg(s, s_in, s_out, k):
        return s * (nnzD + nnzU + nnzV) + n * k * (s_in + s_out)

    b32 = balg(4, 4, 4, nrhs)
    leaf_bytes = 4 * (nnzD + nnzU) + n * nrhs * 4  # dominant kernel: leaf panels + Y
    achieved = leaf_bytes / (stage_avg[3] * 1e-3) / 1e9 if stage_avg[3] > 0 else 0.0
    value = nrhs / t32
    traffic, traffic_cal = measured_traffic() if n == 262144 and nrhs == 16 else (None, None)
    res = {
        "metric": METRIC, "value": value, "unit": "H-matvecs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": t32 * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "H-matrix x 16 vectors, N=%d uniform [0,1)^3, kernel 1/r (diag 0), "
                               "leaf %d, eta %g, ACA+ tol %g, FP32 storage + FP32 accumulation"
                               % (n, args.leaf, args.eta, args.tol),
                   "n": n, "nrhs": nrhs, "leaf": args.leaf, "eta": args.eta, "tol": args.tol,
                   "storage": "fp32", "parallelism": "single-gpu"},
        "hbm_gbs": b32 / t32 / 1e9,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_calibration": traffic_cal,
                     "kernel": "k_leaf: leaf panels (near field + U*T), FP32, 16 RHS",
                     "kernel_ms": stage_avg[3], "kernel_alg_bytes": leaf_bytes},
        "stages_ms": {"gather": stage_avg[0], "far_field_VX": stage_avg[1],
                      "partial_reduce": stage_avg[2], "leaf_panels": stage_avg[3]},
        "rel_err_vs_fp64_direct": err,
        "fp64_1rhs": {"ms": t64_1 * 1e3, "H-matvecs/s": 1.0 / t64_1,
                      "hbm_gbs": balg(8, 8, 8, 1) / t64_1 / 1e9},
        "fp64_16rhs": {"ms": t64_16 * 1e3, "H-matvecs/s": nrhs / t64_16,
                       "hbm_gbs": balg(8, 8, 8, nrhs) / t64_16 / 1e9},
        "bf16_16rhs": {"ms": tbf * 1e3, "H-matvecs/s": nrhs / tbf,
                       "hbm_gbs": balg(2, 4, 4, nrhs) / tbf / 1e9},
        "speedup_vs_fp64_1rhs": value * t64_1,
        "speedup_vs_fp64_16rhs": value / (nrhs / t64_16),
        "nnz": {"D": nnzD, "U": nnzU, "V": nnzV}, "build_s": build_s,
        "device_bytes": info["device_bytes"],
    }
    if not args.no_cpu_baseline:
        log("CPU baseline (oracle restatement of mult.jl:6-32, 1 thread)")
        res["cpu_baseline"] = cpu_baseline(hm, info, args.cpu_sample_gb)
    else:
        res["cpu_baseline"] = None
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
