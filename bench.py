#!/usr/bin/env python3
"""Benchmark of the sparse x dense execute path on one MI355X.

Workload (the headline configuration): FP64 SpMV y = A x on a random
power-law matrix, n = 4,194,304 rows and columns, row lengths
round(lognormal(2.5, 1.0)) clipped to [1, 20000], uniform distinct sorted
columns, N(0,1) values, x uniform[-1, 1] (seeded synthetic data).

A step is one SpMV over the whole matrix with x and y resident in HBM,
issued through perflibs_amd_spmv_exec_async_d on the kernel variant the
library's optimize() selected.  `value` is algorithmic GB/s of the timed steps:
    bytes = nnz*(8 + 4) + (n + 1)*4 + n*8 (x) + n*8 (y written, beta = 0)
(values + column indices + row pointers + x + y, each counted once).

Also reported: every SpMV variant's GB/s on the same matrix (the exhaustive
sweep) and the selected variant's regret, the HIP-event roofline of the
dominant kernel with the random-gather ceilings of the same matrix, and the
reference's CPU path (this project's OpenMP port of it) timed on the host.
`extra_lines` carries the other configurations with their algorithmic bytes
and HBM fractions: FP32 and BF16 value storage on the same matrix, the 5-point
Laplacian 316^2 on the reference's CPU path (and, cache-resident, on the GPU),
CSR/BSR x dense SpMM on the 27-point stencil 128^3, and the batch of 4096
banded products.  `--gpus N` with N > 1 exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "perf-libs-sparse_amd"))
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E peak
N_ROWS = 4_194_304
SEED_A, SEED_X = 0x5EED0002, 0x5EED1002
METRIC = "SpMV/SpMM GB/s (% of HBM peak) + learned-vs-exhaustive kernel regret"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(n: int, nnz: int) -> int:
    return nnz * (8 + 4) + (n + 1) * 4 + n * 8 + n * 8


def measured_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 PMC summaries (FETCH_SIZE and WRITE_SIZE, each scaled by the
    factor the same counter reads on a device copy of known size), or None."""
    for f in sorted((ROOT / "profiles").glob("*.md")):
        for line in f.read_text().splitlines():
            if line.startswith("traffic_bytes_per_launch:"):
                parts = line.split()
                if len(parts) == 4 and parts[1] == kernel and parts[2] == workload:
                    return int(parts[3]), f.name
    return None


def time_gpu(A, xd, yd, warmup: int, steps: int):
    import torch

    for _ in range(warmup):
        A.spmv_async(xd.data_ptr(), yd.data_ptr())
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for _ in range(steps):
        A.spmv_async(xd.data_ptr(), yd.data_ptr())
    e1.record()
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1e3 / steps
    return wall_ms, e0.elapsed_time(e1) / steps


def cpu_baseline(n, rp, ci, v, x, y_gpu) -> dict:
    """The reference's OpenMP CPU path (oracle/cpu.c, oracle/scs.c port):
    nnz-balanced first-touch CSR (K3), vanilla CSR (K1) and the generic
    SELL-C-sigma kernel with the audition's candidate rule (first C from the
    FP64 registry 16, 14, ..., 2 and sigma = C*{0, 200, 2000} whose padding
    efficiency reaches 0.96); the fastest is the CPU's pick, as the
    reference's optimize() would choose."""
    import numpy as np

    from oracle import oracle as O

    L = O.lib()
    threads = O.omp_threads()
    p = lambda a: a.ctypes.data
    nnz = int(rp[-1])
    reps, warm = 10, 3
    y = np.zeros(n)

    def timed(fn):
        for _ in range(warm):
            fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts))

    results = {}
    rows, inds = O.csr_decomp(n, rp, threads, 1)
    results["csr_first_touch"] = timed(lambda: L.orc_cpu_csr_first_touch(
        n, p(rp), p(ci), p(v), p(x), 1.0, 0.0, p(y), threads, p(rows), p(inds)))
    results["csr_vanilla"] = timed(lambda: L.orc_cpu_csr_vanilla(n, p(rp), p(ci), p(v), p(x), 1.0, 0.0, p(y)))
    pick = None
    for Cc in (16, 14, 12, 10, 8, 6, 4, 2):
        for sm in (0, 200, 2000):
            if O.scs_beta(n, n, rp, ci, Cc, Cc * sm) >= 0.96:
                pick = (Cc, Cc * sm)
                break
        if pick:
            break
    if pick:
        h = O.ScsHandle(n, n, rp, ci, v, *pick)
        results[f"sell_c{pick[0]}_s{pick[1]}"] = timed(lambda: h.spmv(x, y))
        h.close()
    best = min(results, key=results.get)
    t = results[best]
    # parity check: the GPU result against the CPU port's y of the best kernel
    y[:] = 0
    if best == "csr_vanilla":
        L.orc_cpu_csr_vanilla(n, p(rp), p(ci), p(v), p(x), 1.0, 0.0, p(y))
    else:
        L.orc_cpu_csr_first_touch(n, p(rp), p(ci), p(v), p(x), 1.0, 0.0, p(y), threads, p(rows), p(inds))
    diff = float(np.linalg.norm(y - y_gpu) / np.linalg.norm(y))
    gbs = algorithmic_bytes(n, nnz) / t / 1e9
    return {
        "value": round(gbs, 2), "unit": "GB/s", "cores": threads, "kind": "port",
        "sample": f"full matrix (nnz={nnz}), {warm} warm-up + median of {reps} SpMV per kernel, "
                  f"best of {sorted(results)}",
        "kernel": best, "ms": round(t * 1e3, 3),
        "all_ms": {k: round(s * 1e3, 3) for k, s in results.items()},
        "gpu_vs_cpu_rel_diff": diff,
    }


def batch_lines(nprob: int, steps: int, warmup: int, with_cpu: bool) -> dict:
    """The batched configuration: 4096 banded problems (rows log-uniform in
    [256, 65536], bandwidth 32-512, 5-40 non-zeros per row, k in {1, 4, 16}) as
    one persistent launch replayed from a hipGraph, for FP64, FP32 and BF16
    value storage.  GB/s = sum of every problem's algorithmic bytes / replay
    time; the FP32 and BF16 results are compared with the FP64 run's (the FP64
    batch itself is checked against the compensated oracle by the tests)."""
    import numpy as np
    import torch

    import plsparse as P
    from plsparse import workloads as W

    P.set_stream(torch.cuda.current_stream().cuda_stream)
    t0 = time.perf_counter()
    probs = W.banded_batch(nprob, 0x5EED0004)
    nnz = sum(int(p[1][-1]) for p in probs)
    log(f"{len(probs)} problems, {sum(p[0] for p in probs)} rows, nnz={nnz} in {time.perf_counter() - t0:.1f}s")
    rng = np.random.default_rng(0x5EED1004)
    X64 = [torch.from_numpy(rng.uniform(-1, 1, (m, k))).cuda() for (m, _, _, _, k) in probs]
    results, y64 = {}, None
    for prec in ("fp64", "fp32", "bf16"):
        dt = np.float64 if prec == "fp64" else np.float32
        vals = [p[3] if prec == "fp64" else p[3].astype(np.float32) for p in probs]
        mats = [P.SpMat.csr(m, m, rp, ci, vals[i], flags=P.CREATE_NOCOPY)
                for i, (m, rp, ci, _, _) in enumerate(probs)]
        B = P.Batch(mats, [p[4] for p in probs], P.VALUES_BF16 if prec == "bf16" else P.VALUES_NATIVE)
        X = X64 if prec == "fp64" else [x.float() for x in X64]
        Y = [torch.empty((m, k), dtype=X[0].dtype, device="cuda") for (m, _, _, _, k) in probs]
        B.set_operands([x.data_ptr() for x in X], [y.data_ptr() for y in Y])
        t0 = time.perf_counter()
        B.optimize(autotune=True)
        tune_s = time.perf_counter() - t0
        for _ in range(warmup):
            B.exec(1.0, 0.0)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            B.exec(1.0, 0.0)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        sv = {"fp64": 8, "fp32": 4, "bf16": 2}[prec]
        sx = 8 if prec == "fp64" else 4
        nbytes = sum(int(rp[-1]) * (sv + 4) + (m + 1) * 4 + m * k * sx * 2 for (m, rp, _, _, k) in probs)
        info = B.info()
        sell = B.sell_info()
        entry = {"ms": round(ms, 4), "GB/s": round(nbytes / ms / 1e6, 1), "algorithmic_bytes": nbytes,
                 "autotune_s": round(tune_s, 2), "tile_nnz": info["tile_nnz"], "tiles": info["ntiles"],
                 "lanes_per_row": info["lanes"], "cols_per_lane": info["cpl"],
                 "k1_layout": f"sell_c{sell['C']}_s{sell['sigma']}" if sell["C"] else "csr_tiles"}
        ys = np.concatenate([y.double().cpu().numpy().ravel() for y in Y])
        if prec == "fp64":
            y64 = ys
        else:
            entry["rel_err_vs_fp64"] = float(np.linalg.norm(ys - y64) / np.linalg.norm(y64))
        results[prec] = entry
        log(f"{prec}: {entry}")
        B.destroy()
        del mats, Y
        torch.cuda.empty_cache()
    for r in results.values():
        r["hbm_fraction"] = round(r["GB/s"] / HBM_PEAK_GBS, 4)
    cpu = batch_cpu_baseline(probs, X64, y64, results["fp64"]["algorithmic_bytes"]) if with_cpu else None
    del X64, probs
    torch.cuda.empty_cache()
    return {"problems": nprob, "nnz": nnz, "rhs": [1, 4, 16], "seed": 0x5EED0004,
            "precisions": results, "cpu_baseline": cpu}


def batch_bench(args) -> int:
    r = batch_lines(args.batch, args.steps, args.warmup, not args.no_cpu)
    out = {
        "metric": METRIC, "value": r["precisions"]["fp64"]["GB/s"], "unit": "GB/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": r["precisions"]["fp64"]["ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "batched_banded_4096", "problems": r["problems"], "nnz": r["nnz"],
                   "rhs": r["rhs"], "seed": r["seed"]},
        "precisions": r["precisions"],
        "cpu_baseline": r["cpu_baseline"],
    }
    print(json.dumps(out), flush=True)
    return 0


def batch_cpu_baseline(probs, X64, y64, nbytes) -> dict:
    """The batch on the host as the reference would run it (it has no batch
    API): one call of its row-blocked CSR x dense kernel per problem
    (oracle/cpu.c port, OpenMP inside each call), FP64, 1 warm-up + median of
    3 passes over all problems; the result is checked against the GPU batch."""
    import numpy as np

    from oracle import oracle as O

    L = O.lib()
    threads = O.omp_threads()
    p = lambda a: a.ctypes.data
    X = [x.cpu().numpy() for x in X64]
    Y = [np.zeros((m, k)) for (m, _, _, _, k) in probs]
    ts = []
    for it in range(4):
        for y in Y:
            y[:] = 0.0
        t0 = time.perf_counter()
        for i, (m, rp, ci, v, k) in enumerate(probs):
            L.orc_cpu_spmm_rowwise(m, k, p(rp), p(ci), p(v), p(X[i]), k, 1.0, 0.0, p(Y[i]), k)
        if it >= 1:
            ts.append(time.perf_counter() - t0)
    t = float(np.median(ts))
    ys = np.concatenate([y.ravel() for y in Y])
    return {
        "value": round(nbytes / t / 1e9, 2), "unit": "GB/s", "cores": threads, "kind": "port",
        "sample": "the whole FP64 batch, 1 warm-up + median of 3 passes", "ms": round(t * 1e3, 3),
        "gpu_vs_cpu_rel_diff": float(np.linalg.norm(ys - y64) / np.linalg.norm(ys)),
    }


def spmm_cpu_baseline(rp, ci, v, k, nbytes) -> dict:
    """The reference's row-blocked CSR x dense kernel (oracle/cpu.c port of
    spmm_rowwise_csr_blocked_m, OpenMP) on the same stencil and k, FP64,
    with its result checked against the GPU product of the same B."""
    import numpy as np
    import torch

    import plsparse as P
    from oracle import oracle as O

    L = O.lib()
    threads = O.omp_threads()
    n = len(rp) - 1
    rng = np.random.default_rng(0x5EED1003)
    B = rng.uniform(-1, 1, (n, k))
    Cm = np.zeros((n, k))
    p = lambda a: a.ctypes.data
    ts = []
    for it in range(3 + 5):
        Cm[:] = 0.0
        t0 = time.perf_counter()
        L.orc_cpu_spmm_rowwise(n, k, p(rp), p(ci), p(v), p(B), k, 1.0, 0.0, p(Cm), k)
        if it >= 3:
            ts.append(time.perf_counter() - t0)
    t = float(np.median(ts))
    A = P.SpMat.csr(n, n, rp, ci, v, flags=P.CREATE_NOCOPY)
    Bd = torch.from_numpy(B).cuda()
    Cd = torch.zeros((n, k), dtype=torch.float64, device="cuda")
    A.spmm_async(k, Bd.data_ptr(), k, Cd.data_ptr(), k)
    torch.cuda.synchronize()
    diff = float(np.linalg.norm(Cd.cpu().numpy() - Cm) / np.linalg.norm(Cm))
    A.destroy()
    return {
        "value": round(nbytes / t / 1e9, 2), "unit": "GB/s", "cores": threads, "kind": "port",
        "sample": f"full csr_fp64_k{k} product, 3 warm-up + median of 5", "ms": round(t * 1e3, 3),
        "gpu_vs_cpu_rel_diff": diff,
    }


def spmm_lines(steps: int, warmup: int, with_cpu: bool) -> dict:
    """CSR x dense SpMM: the 27-point stencil on 128^3 (n = 2,097,152) times a
    row-major B of k in {8, 32, 128} columns, FP64 and FP32, and the BSR
    operator of the same shape (27-point block pattern of dense N(0,1) 4x4 or
    8x8 blocks) in FP64, with the SpMV of that BSR operator on its native block
    kernel beside it.  Device-resident B and C (x and y), beta = 0; GB/s counts
    values + indices + B + C once (BSR: block values + block indices)."""
    import numpy as np
    import torch

    import plsparse as P
    from plsparse import workloads as W

    P.set_stream(torch.cuda.current_stream().cuda_stream)
    rp, ci, v = W.stencil27(128)
    n = len(rp) - 1
    nnz = int(rp[-1])
    log(f"27-point stencil n={n} nnz={nnz}")
    lines = {}

    def run(tag, A, k, tdt, nbytes):
        Bd = torch.rand((n, k), dtype=tdt, device="cuda") * 2 - 1
        Cd = torch.empty((n, k), dtype=tdt, device="cuda")
        for _ in range(warmup):
            A.spmm_async(k, Bd.data_ptr(), k, Cd.data_ptr(), k)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            A.spmm_async(k, Bd.data_ptr(), k, Cd.data_ptr(), k)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        lines[tag] = {"ms": round(ms, 4), "GB/s": round(nbytes / ms / 1e6, 1),
                      "hbm_fraction": round(nbytes / ms / 1e6 / HBM_PEAK_GBS, 4),
                      "GFLOP/s": None, "algorithmic_bytes": nbytes}
        log(f"{tag}: {ms:.3f} ms {nbytes / ms / 1e6:.1f} GB/s")
        del Bd, Cd

    for prec, dt, tdt, sv in (("fp64", np.float64, torch.float64, 8), ("fp32", np.float32, torch.float32, 4)):
        A = P.SpMat.csr(n, n, rp, ci, v.astype(dt), flags=P.CREATE_NOCOPY)
        for k in (8, 32, 128):
            nbytes = nnz * (sv + 4) + (n + 1) * 4 + n * k * sv * 2
            run(f"csr_{prec}_k{k}", A, k, tdt, nbytes)
            lines[f"csr_{prec}_k{k}"]["GFLOP/s"] = round(2 * nnz * k / lines[f"csr_{prec}_k{k}"]["ms"] / 1e6, 1)
            lines[f"csr_{prec}_k{k}"]["row_blocks"] = "clustered" if A.spmm_rows_clustered() else "consecutive"
        A.destroy()
    for b, grid in ((4, (128, 64, 64)), (8, (64, 64, 64))):
        brp, bci, bv = W.block_stencil(*grid, b, 0x5EED0003)
        nnzb = int(brp[-1])
        # SpMV on the native BSR plan (a handle of its own: the SpMM below turns
        # its handle into CSR, as the reference's SpMM does)
        S = P.SpMat.bsr(P.ROW_MAJOR, n, n, b, brp, bci, bv, flags=P.CREATE_NOCOPY)
        xd = torch.rand(n, dtype=torch.float64, device="cuda") * 2 - 1
        yd = torch.empty(n, dtype=torch.float64, device="cuda")
        for _ in range(warmup):
            S.spmv_async(xd.data_ptr(), yd.data_ptr())
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            S.spmv_async(xd.data_ptr(), yd.data_ptr())
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        nbytes = nnzb * b * b * 8 + nnzb * 4 + (n // b + 1) * 4 + n * 8 * 2
        lines[f"bsr{b}_fp64_spmv"] = {"ms": round(ms, 4), "GB/s": round(nbytes / ms / 1e6, 1),
                                      "hbm_fraction": round(nbytes / ms / 1e6 / HBM_PEAK_GBS, 4),
                                      "GFLOP/s": round(2 * nnzb * b * b / ms / 1e6, 1),
                                      "algorithmic_bytes": nbytes, "variant": S.variant()}
        log(f"bsr{b}_fp64_spmv: {ms:.3f} ms {nbytes / ms / 1e6:.1f} GB/s")
        S.destroy()
        del xd, yd
        A = P.SpMat.bsr(P.ROW_MAJOR, n, n, b, brp, bci, bv, flags=P.CREATE_NOCOPY)
        for k in (8, 32):
            nbytes = nnzb * b * b * 8 + nnzb * 4 + (n // b + 1) * 4 + n * k * 8 * 2
            run(f"bsr{b}_fp64_k{k}", A, k, torch.float64, nbytes)
            lines[f"bsr{b}_fp64_k{k}"]["GFLOP/s"] = round(2 * nnzb * b * b * k / lines[f"bsr{b}_fp64_k{k}"]["ms"] / 1e6, 1)
        A.destroy()
        del brp, bci, bv
    cpu = spmm_cpu_baseline(rp, ci, v, 32, lines["csr_fp64_k32"]["algorithmic_bytes"]) if with_cpu else None
    return {"rows": n, "nnz": nnz, "lines": lines, "cpu_baseline": cpu}


def spmm_bench(args) -> int:
    r = spmm_lines(args.steps, args.warmup, not args.no_cpu)
    best = r["lines"]["csr_fp64_k32"]
    out = {
        "metric": METRIC, "value": best["GB/s"], "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": best["ms"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "spmm_stencil27_128", "rows": r["rows"], "nnz": r["nnz"],
                   "value_line": "csr_fp64_k32"},
        "lines": r["lines"],
        "cpu_baseline": r["cpu_baseline"],
    }
    print(json.dumps(out), flush=True)
    return 0


def reduced_precision_lines(rp, ci, v, x, y64, steps: int, warmup: int) -> dict:
    """FP32 and BF16 value storage on the headline matrix (FP32 x and y, FP32
    accumulation; optimize() selects the variant): time, algorithmic GB/s with
    s_v = 4 or 2 and s_x = s_y = 4, and the normwise error against the FP64
    result of the same matrix and x."""
    import numpy as np
    import torch

    import plsparse as P

    n = len(rp) - 1
    nnz = int(rp[-1])
    out = {}
    v32 = v.astype(np.float32)
    xd = torch.from_numpy(x.astype(np.float32)).cuda()
    yd = torch.zeros(n, dtype=torch.float32, device="cuda")
    for tag, prec, sv in (("fp32", P.VALUES_NATIVE, 4), ("bf16", P.VALUES_BF16, 2)):
        A = P.SpMat.csr(n, n, rp, ci, v32, flags=P.CREATE_NOCOPY)
        A.hint(P.HINT_SPMV_INVOCATIONS, P.INVOCATIONS_MANY)
        A.set_precision(prec)
        A.optimize()
        _, kms = time_gpu(A, xd, yd, warmup, steps)
        y = yd.double().cpu().numpy()
        nbytes = nnz * (sv + 4) + (n + 1) * 4 + n * 4 * 2
        out[tag] = {"variant": P.VARIANTS[A.variant()], "ms": round(kms, 4),
                    "GB/s": round(nbytes / kms / 1e6, 1),
                    "hbm_fraction": round(nbytes / kms / 1e6 / HBM_PEAK_GBS, 4),
                    "algorithmic_bytes": nbytes,
                    "rel_err_vs_fp64": float(np.linalg.norm(y - y64) / np.linalg.norm(y64))}
        log(f"spmv {tag}: {out[tag]}")
        A.destroy()
    return out


def laplace_line(steps: int) -> dict:
    """The 5-point Laplacian on 316 x 316 (n = 99,856, 498,016 non-zeros,
    y = A x + 0.5 y, FP64) on the reference's CPU path (this project's OpenMP
    port of its first-touch and vanilla CSR kernels, the faster reported),
    and on the GPU (optimize() selects; cache-resident: the whole problem is
    8.8 MB, far inside the L2 and MALL, so its GB/s is not an HBM figure)."""
    import numpy as np
    import torch

    import plsparse as P
    from oracle import oracle as O
    from plsparse import workloads as W

    rp, ci, v = W.laplace2d(316)
    n = 316 * 316
    nnz = int(rp[-1])
    x = W.uniform(n, 0x5EED0001)
    y0 = W.uniform(n, 0x5EED1001)
    nbytes = nnz * 12 + (n + 1) * 4 + n * 8 + n * 8 * 2
    L = O.lib()
    threads = O.omp_threads()
    p = lambda a: a.ctypes.data
    rows, inds = O.csr_decomp(n, rp, threads, 1)
    y = y0.copy()
    cpu = {}
    for name, fn in (("csr_first_touch", lambda: L.orc_cpu_csr_first_touch(
            n, p(rp), p(ci), p(v), p(x), 1.0, 0.5, p(y), threads, p(rows), p(inds))),
            ("csr_vanilla", lambda: L.orc_cpu_csr_vanilla(n, p(rp), p(ci), p(v), p(x), 1.0, 0.5, p(y)))):
        for _ in range(3):
            fn()
        ts = []
        for _ in range(20):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        cpu[name] = float(np.median(ts))
    best = min(cpu, key=cpu.get)
    ref = O.spmv(n, rp, ci, v, x, 1.0, 0.5, y0)
    A = P.SpMat.csr(n, n, rp, ci, v, flags=P.CREATE_NOCOPY)
    A.hint(P.HINT_SPMV_INVOCATIONS, P.INVOCATIONS_MANY)
    A.optimize()
    xd = torch.from_numpy(x).cuda()
    yd = torch.from_numpy(y0.copy()).cuda()
    A.spmv_async(xd.data_ptr(), yd.data_ptr(), 1.0, 0.5)
    torch.cuda.synchronize()
    err = float(np.linalg.norm(yd.cpu().numpy() - ref) / np.linalg.norm(ref))
    for _ in range(5):
        A.spmv_async(xd.data_ptr(), yd.data_ptr(), 1.0, 0.5)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        A.spmv_async(xd.data_ptr(), yd.data_ptr(), 1.0, 0.5)
    e1.record()
    torch.cuda.synchronize()
    gms = e0.elapsed_time(e1) / steps
    out = {"rows": n, "nnz": nnz, "alpha": 1.0, "beta": 0.5, "algorithmic_bytes": nbytes,
           "cpu_path": {"kernel": best, "ms": round(cpu[best] * 1e3, 4),
                        "GB/s": round(nbytes / cpu[best] / 1e9, 2), "cores": threads,
                        "all_ms": {k: round(t * 1e3, 4) for k, t in cpu.items()}},
           "gpu_cache_resident": {"variant": P.VARIANTS[A.variant()], "ms": round(gms, 4),
                                  "GB/s": round(nbytes / gms / 1e6, 1), "rel_err_vs_oracle": err}}
    A.destroy()
    log(f"laplace 316^2: {out}")
    return out


def gather_ceiling(A_panel, xd, nnz: int, kern_ms: float, steps: int) -> dict:
    """Ceilings of the headline SpMV set by its x gathers: (1) the panel
    kernel's own index/value stream and x gathers without the row sums
    (perflibs_amd_probe_panel_gathers, same plan and grid); (2) nnz random
    8-byte gathers from a 512 KiB table (L2-resident) and (3) from a 32 MiB
    table (the size of x, MALL-resident), with no index stream
    (perflibs_amd_probe_random_gathers).  Each is HIP-event timed on the
    library stream over `steps` launches."""
    import torch

    import plsparse as P

    L = P.lib()
    info = A_panel.panel_info()
    out_p = torch.empty(max(info["grid"], 1) * 256, dtype=torch.float64, device="cuda")
    out_r = torch.empty(8192 * 256, dtype=torch.float64, device="cuda")
    tab_l2 = torch.rand(1 << 16, dtype=torch.float64, device="cuda")

    def timed(fn):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    t_panel = timed(lambda: P.check("probe_panel_gathers", L.perflibs_amd_probe_panel_gathers(
        A_panel.h, xd.data_ptr(), out_p.data_ptr())))
    t_l2 = timed(lambda: P.check("probe_random_gathers", L.perflibs_amd_probe_random_gathers(
        tab_l2.data_ptr(), 1 << 16, nnz, out_r.data_ptr())))
    mall_len = 1 << 22
    tab_mall = xd if xd.numel() >= mall_len else torch.rand(mall_len, dtype=torch.float64, device="cuda")
    t_mall = timed(lambda: P.check("probe_random_gathers", L.perflibs_amd_probe_random_gathers(
        tab_mall.data_ptr(), mall_len, nnz, out_r.data_ptr())))
    out = {"gathers": nnz,
           "panel_streams_and_gathers_ms": round(t_panel, 4),
           "random_gathers_l2_ms": round(t_l2, 4), "random_gathers_l2_per_s": round(nnz / t_l2 / 1e6, 1),
           "random_gathers_mall_ms": round(t_mall, 4),
           "random_gathers_mall_per_s": round(nnz / t_mall / 1e6, 1),
           "kernel_ms": round(kern_ms, 4),
           "frac_of_l2_gather_ceiling": round(t_l2 / kern_ms, 4),
           "frac_of_panel_stream_gather_ceiling": round(t_panel / kern_ms, 4)}
    log(f"gather ceiling: {out}")
    return out


def gather_bench(args) -> int:
    """The gather workload: the headline matrix's column-panel plan, its SpMV
    and the gather probes of gather_ceiling() (G gathers per second), one JSON
    line whose value is the L2 random-gather rate in G gathers/s."""
    import torch

    import plsparse as P
    from plsparse import workloads as W

    n = args.rows
    rp, ci, v = W.powerlaw(n, SEED_A)
    x = W.uniform(n, SEED_X)
    nnz = int(rp[-1])
    P.set_stream(torch.cuda.current_stream().cuda_stream)
    xd = torch.from_numpy(x).cuda()
    yd = torch.zeros(n, dtype=torch.float64, device="cuda")
    A = P.SpMat.csr(n, n, rp, ci, v, flags=P.CREATE_NOCOPY)
    A.force_variant(P.VARIANTS.index("csr_panel"))
    A.optimize()
    _, kms = time_gpu(A, xd, yd, args.warmup, args.steps)
    g = gather_ceiling(A, xd, nnz, kms, args.steps)
    A.destroy()
    out = {"metric": "random 8-byte gathers per second (L2-resident table)",
           "value": round(nnz / g["random_gathers_l2_ms"] / 1e6, 1), "unit": "Ggather/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": g["random_gathers_l2_ms"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "gather_powerlaw_4m", "rows": n, "nnz": nnz, "seed": SEED_A},
           "gather_ceiling": g}
    print(json.dumps(out), flush=True)
    return 0


def mtx_bench(args) -> int:
    """FP64 SpMV on a Matrix Market file (read with the acceptance rules of the
    reference's benchmark reader, bench/pl_bench_sparse_x_dense.c:124-240): optimize() selects a
    variant, every applicable variant is timed for the regret, x uniform[-1, 1]."""
    import numpy as np
    import torch

    import plsparse as P
    from plsparse import workloads as W

    t0 = time.perf_counter()
    m, n, rp, ci, v = W.read_mtx(args.mtx)
    nnz = int(rp[-1])
    log(f"{args.mtx}: {m} x {n}, nnz={nnz}, read in {time.perf_counter() - t0:.1f}s")
    nbytes = nnz * (8 + 4) + (m + 1) * 4 + n * 8 + m * 8
    P.set_stream(torch.cuda.current_stream().cuda_stream)
    xd = torch.from_numpy(W.uniform(n, SEED_X)).cuda()
    yd = torch.zeros(m, dtype=torch.float64, device="cuda")
    A = P.SpMat.csr(m, n, rp, ci, v, flags=P.CREATE_NOCOPY)
    A.hint(P.HINT_SPMV_INVOCATIONS, P.INVOCATIONS_MANY)
    A.optimize()
    chosen = P.VARIANTS[A.variant()]
    wall_ms, kern_ms = time_gpu(A, xd, yd, args.warmup, args.steps)
    A.destroy()
    sweep = {}
    for vi, name in enumerate(P.VARIANTS):
        b = 4 if name == "bsr_b4" else 8 if name == "bsr_b8" else 1
        if m % b or n % b:
            continue
        B = P.SpMat.csr(m, n, rp, ci, v, flags=P.CREATE_NOCOPY)
        B.force_variant(vi)
        B.optimize()
        sweep[name] = time_gpu(B, xd, yd, 3, 20)[1]
        B.destroy()
    best = min(sweep, key=sweep.get)
    value = nbytes / (wall_ms * 1e-3) / 1e9
    achieved = nbytes / (kern_ms * 1e-3) / 1e9
    print(json.dumps({
        "metric": METRIC, "value": round(value, 2), "unit": "GB/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall_ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": f"file {Path(args.mtx).name}",
        "config": {"workload": f"spmv_fp64_mtx_{Path(args.mtx).stem}", "rows": m, "cols": n,
                   "nnz": nnz, "alpha": 1.0, "beta": 0.0, "variant": chosen},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": chosen, "kernel_ms": round(kern_ms, 4), "algorithmic_bytes": nbytes},
        "cpu_baseline": None,
        "variants_gbs": {k: round(nbytes / ms / 1e6, 1) for k, ms in sweep.items()},
        "selector": {"picked": chosen, "exhaustive_best": best,
                     "regret": round(sweep[chosen] / sweep[best] - 1.0, 4)},
    }), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--variant", default="auto", help="force a variant name, or 'auto' (optimize selects)")
    ap.add_argument("--no-sweep", action="store_true", help="skip the per-variant exhaustive sweep")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--rows", type=int, default=N_ROWS)
    ap.add_argument("--workload", choices=("spmv", "batch", "spmm", "gather"), default="spmv",
                    help="spmv: the headline SpMV line (default); batch: the batched configuration; "
                         "spmm: CSR/BSR x dense on the 27-point stencil; gather: the random-gather "
                         "ceilings of the headline matrix")
    ap.add_argument("--no-extras", action="store_true",
                    help="headline line only: skip the SpMM, batch, FP32/BF16, Laplacian and gather lines")
    ap.add_argument("--batch", type=int, default=4096, help="problems of the batched workload")
    ap.add_argument("--mtx", default=None, help="FP64 SpMV on this Matrix Market file instead")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (asked for {args.gpus}); multi-GPU is not part of this build",
              file=sys.stderr)
        return 2
    if args.workload == "batch":
        return batch_bench(args)
    if args.workload == "spmm":
        return spmm_bench(args)
    if args.workload == "gather":
        return gather_bench(args)
    if args.mtx:
        return mtx_bench(args)

    import numpy as np
    import torch

    import plsparse as P
    from plsparse import workloads as W

    n = args.rows
    t0 = time.perf_counter()
    rp, ci, v = W.powerlaw(n, SEED_A)
    x = W.uniform(n, SEED_X)
    nnz = int(rp[-1])
    log(f"matrix n={n} nnz={nnz} generated in {time.perf_counter() - t0:.1f}s")
    nbytes = algorithmic_bytes(n, nnz)

    stream = torch.cuda.current_stream()
    P.set_stream(stream.cuda_stream)
    xd = torch.from_numpy(x).cuda()
    yd = torch.zeros(n, dtype=torch.float64, device="cuda")

    A = P.SpMat.csr(n, n, rp, ci, v, flags=P.CREATE_NOCOPY)
    A.hint(P.HINT_SPMV_INVOCATIONS, P.INVOCATIONS_MANY)
    if args.variant != "auto":
        A.force_variant(P.VARIANTS.index(args.variant))
    t0 = time.perf_counter()
    A.optimize()
    chosen = P.VARIANTS[A.variant()]
    log(f"optimize() selected {chosen} in {time.perf_counter() - t0:.1f}s")

    wall_ms, kern_ms = time_gpu(A, xd, yd, args.warmup, args.steps)
    y_gpu = yd.cpu().numpy()
    value = nbytes / (wall_ms * 1e-3) / 1e9
    achieved = nbytes / (kern_ms * 1e-3) / 1e9
    log(f"{chosen}: {wall_ms:.3f} ms/step wall, {kern_ms:.3f} ms kernel, {achieved:.1f} GB/s")
    gather = None
    if not args.no_extras:
        if chosen == "csr_panel":
            gather = gather_ceiling(A, xd, nnz, kern_ms, 20)
        else:
            G = P.SpMat.csr(n, n, rp, ci, v, flags=P.CREATE_NOCOPY)
            G.force_variant(P.VARIANTS.index("csr_panel"))
            G.optimize()
            _, gms = time_gpu(G, xd, yd, 3, 20)
            gather = gather_ceiling(G, xd, nnz, gms, 20)
            G.destroy()
    A.destroy()

    sweep = {}
    if not args.no_sweep:
        for vi, name in enumerate(P.VARIANTS):
            if name.startswith("bsr"):
                continue  # random columns: 4x4/8x8 blocks would hold ~1 value each
            B = P.SpMat.csr(n, n, rp, ci, v, flags=P.CREATE_NOCOPY)
            B.force_variant(vi)
            B.optimize()
            _, kms = time_gpu(B, xd, yd, 3, 20)
            sweep[name] = kms
            log(f"  sweep {name}: {kms:.3f} ms  {nbytes / kms / 1e6:.1f} GB/s")
            B.destroy()
    best = min(sweep, key=sweep.get) if sweep else chosen
    regret = (sweep[chosen] / sweep[best] - 1.0) if chosen in sweep else None

    cpu = None if args.no_cpu else cpu_baseline(n, rp, ci, v, x, y_gpu)
    workload = "spmv_fp64_powerlaw_4m" if n == N_ROWS else f"spmv_fp64_powerlaw_{n}"
    traffic = measured_traffic(chosen, workload)

    # the other configurations, each with its algorithmic bytes and HBM fraction
    # (fewer repetitions than the headline, so the whole run stays in minutes)
    extras = None
    if not args.no_extras:
        extras = {"spmv_powerlaw_4m_reduced_precision": reduced_precision_lines(rp, ci, v, x, y_gpu, 20, 3)}
        del xd, yd
        torch.cuda.empty_cache()
        extras["laplace2d_316"] = laplace_line(50)
        t0 = time.perf_counter()
        extras["spmm_stencil27_128"] = spmm_lines(10, 2, not args.no_cpu)
        log(f"spmm lines in {time.perf_counter() - t0:.1f}s")
        t0 = time.perf_counter()
        extras["batched_banded_4096"] = batch_lines(args.batch, 10, 2, not args.no_cpu)
        log(f"batch lines in {time.perf_counter() - t0:.1f}s")

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": workload, "rows": n, "cols": n, "nnz": nnz,
            "row_length": "round(lognormal(2.5, 1.0)) clipped to [1, 20000]",
            "alpha": 1.0, "beta": 0.0, "seed": SEED_A, "variant": chosen,
            "hbm_fraction": round(value / HBM_PEAK_GBS, 4),
            "gflops": round(2.0 * nnz / (wall_ms * 1e-3) / 1e9, 1),
        },
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic[0] if traffic else None,
            "traffic_source": f"profiles/{traffic[1]}" if traffic else None,
            "kernel": chosen, "kernel_ms": round(kern_ms, 4), "algorithmic_bytes": nbytes,
            "gather_ceiling": gather,
        },
        "cpu_baseline": cpu,
        "variants_gbs": {k: round(nbytes / ms / 1e6, 1) for k, ms in sweep.items()},
        "selector": {"picked": chosen, "exhaustive_best": best,
                     "regret": None if regret is None else round(regret, 4)},
        "extra_lines": extras,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
