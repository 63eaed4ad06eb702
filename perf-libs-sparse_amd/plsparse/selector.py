"""Training driver of the learned SpMV format selector.

    python -m plsparse.selector collect  --out DATA.npz --count N --seed S
    python -m plsparse.selector train    --data DATA.npz --model MODEL.txt
    python -m plsparse.selector evaluate --data HELDOUT.npz --model MODEL.txt

`collect` draws matrices from five generators (2-D/3-D stencils, banded,
power-law, 4x4/8x8 block-structured, uniform random; rows log-uniform, mean
non-zeros per row log-uniform in [2, 200]) and records, per matrix, the 12
selector features the library computes on the GPU (perflibs_amd_features),
the time of every SpMV variant (perflibs_amd_time_variants: the exhaustive
autotuner) and the variant the hand heuristic picks.  `train` fits the
12 -> 64 -> 64 -> K MLP on the GPU with the library's own forward, backward and
Adam kernels (perflibs_amd_mlp_*), targets log(t_variant / t_best) masked to
the variants that ran, and writes the model file the library loads at
optimize time.  `evaluate` scores a model on held-out matrices: regret
(geometric mean of t(pick) / t(best) - 1), speedup over the fixed CSR-vector
kernel, and the same for the heuristic (the selector without the network).
"""
from __future__ import annotations

import argparse
import sys
import time
from pathlib import Path

import numpy as np

from . import (CREATE_NOCOPY, HINT_SPMV_INVOCATIONS, INVOCATIONS_MANY, NUM_FEATURES, NUM_VARIANTS,
               SELECT_HEURISTIC, SELECT_LEARNED, SpMat, VARIANTS, check, lib)
from . import workloads as W

GENERATORS = ("stencil", "banded", "powerlaw", "block", "random")
DEFAULT_MODEL = Path(__file__).resolve().parent.parent / "selector" / "selector_mlp.txt"
MAX_NNZ = 16_000_000
HIDDEN = 64

# stencil shapes: (dimension, radius, star) -> points per interior row
_STENCILS = [(2, 1, True, 5), (2, 2, True, 9), (2, 3, True, 13), (2, 4, True, 17),
             (2, 1, False, 9), (2, 2, False, 25), (2, 3, False, 49), (2, 5, False, 121),
             (3, 1, True, 7), (3, 2, True, 13), (3, 3, True, 19), (3, 1, False, 27),
             (3, 2, False, 125)]


def _loguni(rng, lo, hi):
    return float(np.exp(rng.uniform(np.log(lo), np.log(hi))))


def sample_matrix(kind: str, seed: int, rows=(10_000, 8_000_000), mean=(2.0, 200.0),
                  max_nnz: int = MAX_NNZ):
    """One random matrix of generator `kind` as (m, n, row_ptr, col, vals, desc)."""
    rng = np.random.default_rng(seed)
    m_t = _loguni(rng, *rows)
    mu_t = _loguni(rng, *mean)
    m_t = min(m_t, max_nnz / mu_t)
    gseed = seed * 7919 + 17
    if kind == "stencil":
        d, r, star, pts = min(_STENCILS, key=lambda s: abs(np.log(s[3] / mu_t)))
        m_t = min(m_t, max_nnz / pts)
        if d == 2:
            nx = max(8, int(round(np.sqrt(m_t))))
            ny = max(8, int(round(m_t / nx)))
            rp, ci, v = W.stencil2d(nx, ny, r, star)
        else:
            nx = max(6, int(round(m_t ** (1 / 3))))
            ny = nx
            nz = max(6, int(round(m_t / (nx * ny))))
            rp, ci, v = W.stencil3d(nx, ny, nz, r, star)
        m = len(rp) - 1
        return m, m, rp, ci, v, f"stencil{d}d r={r} {'star' if star else 'box'}"
    if kind == "banded":
        m = max(1000, int(m_t))
        lo = max(1, int(round(mu_t / 2)))
        hi = max(lo, int(round(1.5 * mu_t)))
        bw = int(rng.integers(2 * hi + 1, 20 * hi + 2))
        rp, ci, v = W.banded(m, bw, lo, hi, gseed)
        return m, m, rp, ci, v, f"banded bw={bw}"
    if kind == "powerlaw":
        m = max(1000, int(m_t))
        sig = float(rng.uniform(0.5, 1.5))
        mu = float(np.log(mu_t) - sig * sig / 2)
        rp, ci, v = W.powerlaw(m, gseed, mu=mu, sigma=sig, lo=1, hi=min(m, 20000))
        return m, m, rp, ci, v, f"powerlaw sigma={sig:.2f}"
    if kind == "block":
        b = int(rng.choice([4, 8]))
        kb = max(1, int(round(mu_t / b)))
        w = int(rng.integers(kb, 8 * kb + 1))
        mb = max(256, int(m_t) // b)
        brp, bci, bv = W.block_random(mb, b, kb, w, gseed)
        rp, ci, v = W.bsr_to_csr_rowmajor(brp, bci, bv, b)
        return mb * b, mb * b, rp, ci, v, f"block b={b} k={kb} w={w}"
    if kind == "random":
        m = max(1000, int(m_t))
        lo = max(1, int(round(mu_t / 2)))
        hi = max(lo, int(round(1.5 * mu_t)))
        rp, ci, v = W.random_rows(m, m, lo, hi, gseed)
        return m, m, rp, ci, v, "uniform random"
    raise ValueError(kind)


def measure(m, n, rp, ci, v, warmup=3, reps=10):
    """Features, per-variant times (ms, < 0: not a candidate) and the heuristic pick."""
    A = SpMat.csr(m, n, rp, ci, v, flags=CREATE_NOCOPY)
    A.hint(HINT_SPMV_INVOCATIONS, INVOCATIONS_MANY)
    feat = A.features()
    ms = A.time_variants(warmup, reps)
    A.destroy()
    H = SpMat.csr(m, n, rp, ci, v, flags=CREATE_NOCOPY)
    H.hint(HINT_SPMV_INVOCATIONS, INVOCATIONS_MANY)
    H.select_mode(SELECT_HEURISTIC)
    H.optimize()
    heur = H.variant()
    H.destroy()
    return feat, ms, heur


def matrix_seed(run_seed: int, i: int) -> int:
    """Generator seed of matrix i of a collect run (sample_matrix's `seed`)."""
    return run_seed * 100_003 + i


def collect(count: int, seed: int, out: Path, log=print, max_nnz: int = MAX_NNZ):
    feats, times, heur, kinds, sizes, seeds = [], [], [], [], [], []
    t0 = time.time()
    for i in range(count):
        kind = GENERATORS[i % len(GENERATORS)]
        ms_seed = matrix_seed(seed, i)
        m, n, rp, ci, v, desc = sample_matrix(kind, ms_seed, max_nnz=max_nnz)
        f, ms, h = measure(m, n, rp, ci, v)
        feats.append(f)
        times.append(ms)
        heur.append(h)
        kinds.append(kind)
        sizes.append((m, int(rp[-1])))
        seeds.append(ms_seed)
        best = int(np.argmax(np.where(ms > 0, -ms, -np.inf)))
        log(f"[collect] {i + 1}/{count} {desc} m={m} nnz={int(rp[-1])} best={VARIANTS[best]} "
            f"{ms[best]:.4f} ms heuristic={VARIANTS[h]} ({time.time() - t0:.0f}s)")
        if (i + 1) % 20 == 0 or i + 1 == count:  # keep what was measured so far
            np.savez(out, features=np.array(feats, np.float32), times=np.array(times, np.float64),
                     heuristic=np.array(heur, np.int32), kinds=np.array(kinds),
                     sizes=np.array(sizes, np.int64), seeds=np.array(seeds, np.int64),
                     max_nnz=np.full(len(seeds), max_nnz, np.int64))


def merge(parts, out: Path, legacy=(), limit: int | None = None):
    """Concatenate collect outputs (the first `limit` matrices when given).
    legacy: (file, run seed, max_nnz) for files written before seeds were
    recorded, whose matrix i has seed matrix_seed(run seed, i)."""
    legacy = {str(f): (s, mx) for f, s, mx in legacy}
    cols = {k: [] for k in ("features", "times", "heuristic", "kinds", "sizes", "seeds", "max_nnz")}
    for f in parts:
        d = np.load(f)
        n = len(d["times"])
        for k in ("features", "times", "heuristic", "kinds", "sizes"):
            cols[k].append(d[k])
        if "seeds" in d:
            cols["seeds"].append(d["seeds"])
            cols["max_nnz"].append(d["max_nnz"])
        else:
            rs, mx = legacy[str(f)]
            cols["seeds"].append(np.array([matrix_seed(rs, i) for i in range(n)], np.int64))
            cols["max_nnz"].append(np.full(n, mx, np.int64))
    np.savez(out, **{k: np.concatenate(v)[:limit] for k, v in cols.items()})


def targets(times: np.ndarray):
    """log(t / t_best) per variant and the mask of variants that ran."""
    mask = (times > 0).astype(np.float32)
    t = np.where(times > 0, times, np.inf)
    best = t.min(axis=1, keepdims=True)
    y = np.where(mask > 0, np.log(np.where(times > 0, times, 1.0) / best), 0.0)
    return y.astype(np.float32), mask


def train(features: np.ndarray, times: np.ndarray, steps=4000, lr=3e-3, seed=0, log=print,
          val=None):
    """Fit the MLP on the GPU with the library's kernels; returns (params, mean, std).
    val = (features, times) of a validation split, scored at every log point."""
    import torch

    L = lib()
    X = np.asarray(features, np.float32)
    mean = X.mean(axis=0)
    std = X.std(axis=0)
    std[std < 1e-6] = 1.0
    Xs = ((X - mean) / std).astype(np.float32)
    Y, M = targets(times)
    n = len(Xs)
    K = NUM_VARIANTS
    P_count = int(L.perflibs_amd_mlp_num_params())
    rng = np.random.default_rng(seed)
    # He-style initialisation of W1, W2, W3 (biases 0)
    F = NUM_FEATURES
    parts = [rng.standard_normal(F * HIDDEN) * np.sqrt(2.0 / F), np.zeros(HIDDEN),
             rng.standard_normal(HIDDEN * HIDDEN) * np.sqrt(2.0 / HIDDEN), np.zeros(HIDDEN),
             rng.standard_normal(HIDDEN * K) * np.sqrt(1.0 / HIDDEN), np.zeros(K)]
    P0 = np.concatenate(parts).astype(np.float32)
    assert len(P0) == P_count
    dev = "cuda"
    Pd = torch.from_numpy(P0).to(dev)
    Xd = torch.from_numpy(Xs).to(dev)
    Td = torch.from_numpy(Y).to(dev)
    Md = torch.from_numpy(M).to(dev)
    m1 = torch.zeros_like(Pd)
    v1 = torch.zeros_like(Pd)
    h1 = torch.empty((n, HIDDEN), device=dev)
    h2 = torch.empty((n, HIDDEN), device=dev)
    out = torch.empty((n, K), device=dev)
    G = torch.empty_like(Pd)
    loss = torch.empty(2, device=dev)
    for step in range(1, steps + 1):
        check("mlp_forward", L.perflibs_amd_mlp_forward(Pd.data_ptr(), Xd.data_ptr(), n, h1.data_ptr(),
                                                        h2.data_ptr(), out.data_ptr()))
        check("mlp_backward", L.perflibs_amd_mlp_backward(Pd.data_ptr(), Xd.data_ptr(), n, h1.data_ptr(),
                                                          h2.data_ptr(), out.data_ptr(), Td.data_ptr(),
                                                          Md.data_ptr(), G.data_ptr(), loss.data_ptr()))
        step_lr = lr * min(1.0, step / 200) * (0.1 ** (step / steps))
        check("mlp_adam", L.perflibs_amd_mlp_adam(Pd.data_ptr(), G.data_ptr(), m1.data_ptr(), v1.data_ptr(),
                                                  step, step_lr, 0.9, 0.999, 1e-8))
        if step % 500 == 0 or step == 1:
            torch.cuda.synchronize()
            msg = f"[train] step {step} loss {float(loss[0].cpu()):.5f}"
            if val is not None:
                pv = predict(Pd.cpu().numpy(), mean, std, val[0])
                sv = score(picks_of(pv, val[1]), val[1])
                msg += f" validation regret {sv['regret']:.4f} within10 {sv['within_10pct']:.2f}"
            log(msg)
    torch.cuda.synchronize()
    return Pd.cpu().numpy(), mean.astype(np.float32), std.astype(np.float32)


def predict(params, mean, std, features):
    """Predicted log slowdown per variant, computed with the library's forward kernel."""
    import torch

    L = lib()
    Xs = ((np.asarray(features, np.float32) - mean) / std).astype(np.float32)
    n = len(Xs)
    Pd = torch.from_numpy(np.ascontiguousarray(params, np.float32)).cuda()
    Xd = torch.from_numpy(Xs).cuda()
    h1 = torch.empty((n, HIDDEN), device="cuda")
    h2 = torch.empty((n, HIDDEN), device="cuda")
    out = torch.empty((n, NUM_VARIANTS), device="cuda")
    check("mlp_forward", L.perflibs_amd_mlp_forward(Pd.data_ptr(), Xd.data_ptr(), n, h1.data_ptr(),
                                                    h2.data_ptr(), out.data_ptr()))
    torch.cuda.synchronize()
    return out.cpu().numpy()


def score(picks: np.ndarray, times: np.ndarray) -> dict:
    """Regret (geometric mean of t(pick)/t(best) - 1) and speedup over CSR-vector."""
    t = np.where(times > 0, times, np.inf)
    best = t.min(axis=1)
    chosen = t[np.arange(len(t)), picks]
    ratio = chosen / best
    speed = t[:, 0] / chosen
    return {"regret": float(np.exp(np.mean(np.log(ratio))) - 1.0),
            "worst_ratio": float(ratio.max()),
            "within_10pct": float(np.mean(ratio <= 1.10)),
            "speedup_vs_csr_vector": float(np.exp(np.mean(np.log(speed)))),
            "n": int(len(t))}


def picks_of(pred: np.ndarray, times: np.ndarray) -> np.ndarray:
    return np.argmin(np.where(times > 0, pred, np.inf), axis=1)


def save_model(path: Path, params, mean, std, fmin=None, fmax=None):
    """Model file; fmin / fmax = the training range of each feature (the
    library's learned mode runs the exhaustive sweep outside it)."""
    lines = ["plsparse-selector-mlp 1",
             f"features {NUM_FEATURES} hidden {HIDDEN} outputs {NUM_VARIANTS}",
             "mean " + " ".join(f"{x:.9g}" for x in mean),
             "std " + " ".join(f"{x:.9g}" for x in std),
             "params " + " ".join(f"{x:.9g}" for x in params)]
    if fmin is not None and fmax is not None:
        lines += ["min " + " ".join(f"{x:.9g}" for x in fmin),
                  "max " + " ".join(f"{x:.9g}" for x in fmax)]
    Path(path).write_text("\n".join(lines) + "\n")


def load_model(path: Path, with_range=False):
    rows = {}
    for line in Path(path).read_text().splitlines():
        key, _, rest = line.partition(" ")
        rows[key] = rest
    arr = lambda k: np.array([float(t) for t in rows[k].split()], np.float32)
    if with_range:
        rng = (arr("min"), arr("max")) if "min" in rows and "max" in rows else (None, None)
        return arr("params"), arr("mean"), arr("std"), rng[0], rng[1]
    return arr("params"), arr("mean"), arr("std")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("collect")
    c.add_argument("--out", required=True)
    c.add_argument("--count", type=int, default=300)
    c.add_argument("--seed", type=int, default=1)
    c.add_argument("--max-nnz", type=int, default=MAX_NNZ,
                   help="non-zeros per matrix at most (rows are cut to fit)")
    mg = sub.add_parser("merge")
    mg.add_argument("--out", required=True)
    mg.add_argument("parts", nargs="+")
    mg.add_argument("--limit", type=int, default=None, help="keep the first LIMIT matrices")
    mg.add_argument("--legacy", action="append", default=[],
                    help="FILE:RUN_SEED:MAX_NNZ for a part written without per-matrix seeds")
    t = sub.add_parser("train")
    t.add_argument("--data", required=True)
    t.add_argument("--model", default=str(DEFAULT_MODEL))
    t.add_argument("--steps", type=int, default=4000)
    t.add_argument("--val", type=float, default=0.0,
                   help="fraction of the data (every k-th matrix) scored as a validation split, not trained on")
    e = sub.add_parser("evaluate")
    e.add_argument("--data", required=True)
    e.add_argument("--model", default=str(DEFAULT_MODEL))
    e.add_argument("--end-to-end", type=int, default=0,
                   help="also rebuild the first N held-out matrices (--seed of their collect run) and check "
                        "that optimize() in learned mode picks what the model predicts")
    e.add_argument("--seed", type=int, default=2)
    args = ap.parse_args(argv)
    log = lambda s: print(s, flush=True)
    if args.cmd == "collect":
        import torch  # noqa: F401  (one HIP runtime: torch's)

        collect(args.count, args.seed, Path(args.out), log, args.max_nnz)
    elif args.cmd == "merge":
        legacy = []
        for item in args.legacy:
            f, rs, mx = item.rsplit(":", 2)
            legacy.append((f, int(rs), int(mx)))
        merge(args.parts, Path(args.out), legacy, args.limit)
    elif args.cmd == "train":
        d = np.load(args.data)
        feats, times = d["features"], d["times"]
        val = None
        if args.val > 0:
            every = max(2, int(round(1 / args.val)))
            vsel = np.arange(len(feats)) % every == every - 1
            val = (feats[vsel], times[vsel])
            feats, times = feats[~vsel], times[~vsel]
        params, mean, std = train(feats, times, steps=args.steps, log=log, val=val)
        save_model(Path(args.model), params, mean, std, feats.min(axis=0), feats.max(axis=0))
        pred = predict(params, mean, std, feats)
        log(f"[train] training-set {score(picks_of(pred, times), times)}")
    else:
        d = np.load(args.data)
        params, mean, std = load_model(Path(args.model))
        pred = predict(params, mean, std, d["features"])
        learned = score(picks_of(pred, d["times"]), d["times"])
        heur = score(d["heuristic"], d["times"])
        fixed = score(np.zeros(len(d["times"]), np.int64), d["times"])
        log(f"[evaluate] learned {learned}")
        log(f"[evaluate] heuristic {heur}")
        log(f"[evaluate] fixed csr_vector {fixed}")
        if args.end_to_end:
            check("load_selector_model", lib().perflibs_amd_load_selector_model(str(args.model).encode()))
            picks = picks_of(pred, d["times"])
            agree = 0
            seeds = d["seeds"] if "seeds" in d else [matrix_seed(args.seed, i) for i in range(len(picks))]
            caps = d["max_nnz"] if "max_nnz" in d else [MAX_NNZ] * len(picks)
            for i in range(args.end_to_end):
                kind = str(d["kinds"][i])
                m, n, rp, ci, v, desc = sample_matrix(kind, int(seeds[i]), max_nnz=int(caps[i]))
                A = SpMat.csr(m, n, rp, ci, v, flags=CREATE_NOCOPY)
                A.hint(HINT_SPMV_INVOCATIONS, INVOCATIONS_MANY)
                A.select_mode(SELECT_LEARNED)
                A.optimize()
                got = A.variant()
                A.destroy()
                agree += int(got == picks[i])
                log(f"[evaluate] end-to-end {i}: {desc} optimize() picked {VARIANTS[got]}, "
                    f"model predicts {VARIANTS[picks[i]]}")
            log(f"[evaluate] end-to-end agreement {agree}/{args.end_to_end}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
