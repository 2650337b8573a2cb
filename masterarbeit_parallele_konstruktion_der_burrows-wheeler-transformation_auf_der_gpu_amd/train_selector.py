"""Collect training runs for the learned (W,P) selector and train it on the GPU.

For every synthetic text the prefix-free-parsing pipeline is run once per candidate
(W in {5,10,20,35} x P in {11,...,81}, and the default W=5,P=16); the selector features and the measured dictionary
length, parse length and build time become one training sample each.  The MLP is trained by
the library's own HIP kernels (mibwt_mlp_train) and written to selector_model.bin next to the
library (or --out).  Held-out texts report the chosen candidate's time against the best grid
point and against fixed W=5,P=16.

  python train_selector.py [--texts 300] [--holdout 24] [--large 8 --large-holdout 2]
                           [--out PATH] [--data runs.npz]
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _pkg():
    if "mibwt" in sys.modules:
        return sys.modules["mibwt"]
    spec = importlib.util.spec_from_file_location("mibwt", os.path.join(HERE, "__init__.py"),
                                                  submodule_search_locations=[HERE])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["mibwt"] = mod
    spec.loader.exec_module(mod)
    return mod


def random_text(rng: np.random.Generator, max_len: int, min_len: int = 4 << 20) -> tuple[np.ndarray, dict]:
    """A repetitive collection: random base (DNA, protein-like or bytes) x copies, mutated."""
    from mibwt import gen
    # large texts lean to DNA collections (the application); small ones mix the three kinds
    kinds = ["dna", "dna", "dna", "protein", "bytes"] if min_len < (64 << 20) else \
        ["dna"] * 7 + ["protein"] * 2 + ["bytes"]
    kind = rng.choice(kinds)
    copies = int(np.exp(rng.uniform(np.log(2), np.log(192))))
    total = int(np.exp(rng.uniform(np.log(min_len), np.log(max_len))))
    base_len = max(4096, total // copies)
    sub = float(np.exp(rng.uniform(np.log(1e-5), np.log(5e-2))))
    indel = sub / float(rng.uniform(5, 20))
    indel_max = int(rng.integers(1, 51))
    seed = int(rng.integers(0, 2**31))
    idx = gen.repetitive_dna(base_len, copies, sub, indel, indel_max, seed)  # ACGT bytes
    if kind == "protein":
        table = np.frombuffer(b"ACDEFGHIKLMNPQRSTVWY", dtype=np.uint8)
        r2 = np.random.default_rng(seed ^ 0x5A5A)
        # map 4-symbol pairs to 20 letters by re-drawing the base with a larger alphabet
        base = r2.integers(0, 20, size=base_len).astype(np.uint8)
        parts = [gen.mutate(r2, base, sub, indel, indel_max) for _ in range(copies)]
        idx = table[np.concatenate(parts) % 20]
    elif kind == "bytes":
        r2 = np.random.default_rng(seed ^ 0xA5A5)
        base = r2.integers(0, 256, size=base_len).astype(np.uint8)
        parts = []
        for _ in range(copies):
            x = base.copy()
            hit = np.flatnonzero(r2.random(len(x)) < sub)
            x[hit] = r2.integers(0, 256, size=len(hit))
            parts.append(x)
        idx = np.concatenate(parts)
    return idx, {"kind": str(kind), "copies": copies, "base_len": base_len, "sub": sub}


def collect(mibwt, a, grid, rng, t0):
    """Runs of every candidate on the generated texts, in the order: regular training texts,
    large training texts, regular held-out, large held-out."""
    Xs, Ys, MS, FX = [], [], [], []
    sizes = ([(4 << 20, a.max_mib << 20)] * a.texts + [(96 << 20, a.large_max_mib << 20)] * a.large
             + [(4 << 20, a.max_mib << 20)] * a.holdout
             + [(96 << 20, a.large_max_mib << 20)] * a.large_holdout)
    total = len(sizes)
    for i, (lo, hi) in enumerate(sizes):
        T, info = random_text(rng, hi, lo)
        X, Y, ms, fms = run_grid(mibwt, T, grid)
        if not np.isfinite(fms):
            raise RuntimeError("W=5,P=16 itself is infeasible for a generated text")
        Xs.append(X)
        Ys.append(Y)
        MS.append(ms)
        FX.append(fms)
        print(f"[train_selector] text {i + 1}/{total} n={len(T)} {info['kind']} best "
              f"{grid[int(ms.argmin())]} {ms.min():.2f} ms, W5P16 {fms:.2f} ms "
              f"({time.time() - t0:.0f} s)", flush=True)
    return np.stack(Xs), np.stack(Ys), np.stack(MS), np.array(FX)


def run_grid(mibwt, T: np.ndarray, grid):
    """features, targets and ms of every grid candidate, and the ms of fixed W=5,P=16.  A fresh
    storage per text: its grow-only buffers would otherwise keep the largest dictionary of every
    text seen so far (a poor (W,P) on a low-repetition text can need many times n)."""
    with mibwt.Storage(len(T)) as st:
        return _run_grid(mibwt, st, T, grid)


def _run_grid(mibwt, st, T: np.ndarray, grid):
    n = len(T)
    dT = mibwt.DeviceBuffer.from_numpy(T)
    dO = mibwt.DeviceBuffer(n)
    X = st.wp_features(dT.address, n)
    Y = np.zeros((len(grid), 3), dtype=np.float32)
    ms = np.zeros(len(grid))
    st.bwt_device(dT.address, dO.address, n, "pfp", 5, 16)  # warm the buffers
    for c, (W, P) in enumerate(grid):
        best = None
        for _ in range(2):  # the faster of two builds: less timing noise in the targets
            try:
                st.bwt_device(dT.address, dO.address, n, "pfp", W, P)
            except mibwt.BwtError as e:
                # expected only for a dictionary beyond the 32-bit index range, or one that does
                # not fit the device
                if e.code not in (mibwt.INDEX_RANGE, mibwt.GPU_NOT_ENOUGH_MEMORY):
                    raise
                best = None
                break
            s = st.stats()
            if best is None or s["total_ms"] < best["total_ms"]:
                best = s
        if best is None:  # infeasible candidate: no training sample, never the best
            ms[c] = np.inf
            Y[c] = np.nan
            continue
        ms[c] = best["total_ms"]
        Y[c] = [np.log(best["dict_length"] / n), np.log(best["phrases"] / n), np.log(best["total_ms"] / n)]
    dT.free()
    dO.free()
    return X, Y, ms, float(ms[grid.index((5, 16))])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--texts", type=int, default=300)
    ap.add_argument("--holdout", type=int, default=24)
    ap.add_argument("--max-mib", type=int, default=64)
    ap.add_argument("--large", type=int, default=0, help="extra texts of 96 MiB .. --large-max-mib")
    ap.add_argument("--large-holdout", type=int, default=0)
    ap.add_argument("--large-max-mib", type=int, default=384)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--members", type=int, default=5, help="networks in the ensemble")
    ap.add_argument("--out", default=os.path.join(HERE, "selector_model.bin"))
    ap.add_argument("--data", default=None, help="write / reuse the collected runs (npz)")
    ap.add_argument("--report", default=None, help="write the held-out evaluation (json)")
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--reuse", default=None,
                    help="train on the runs of an earlier --data file instead of collecting new ones")
    ap.add_argument("--eval-workloads", default="",
                    help="comma-separated named workloads (gen.workload) evaluated, never trained on")
    a = ap.parse_args()
    mibwt = _pkg()
    grid = mibwt.wp_grid()
    rng = np.random.default_rng(a.seed)
    st = mibwt.Storage(1 << 20)  # model training and inference only
    t0 = time.time()
    named_data = {}
    if a.reuse:
        d = np.load(a.reuse)
        X, Y, MS, FX = d["X"], d["Y"], d["MS"], d["FX"]
        for k in d.files:
            if k.endswith("_X"):
                name = k[:-2]
                named_data[name] = (d[name + "_X"], d[name + "_MS"], float(d[name + "_FX"][0]))
        assert len(X) == a.texts + a.large + a.holdout + a.large_holdout, "counts differ from the reused runs"
    else:
        X, Y, MS, FX = collect(mibwt, a, grid, rng, t0)
        if a.data:
            np.savez(a.data, X=X, Y=Y, MS=MS, FX=FX)
    ntr = a.texts + a.large
    Xtr, Ytr = X[:ntr].reshape(-1, 64), Y[:ntr].reshape(-1, 3)
    ok = np.isfinite(Ytr).all(axis=1)
    loss = st.mlp_train(Xtr[ok], Ytr[ok], epochs=a.epochs, batch=512, lr=1e-3, seed=a.seed & 0xFFFF,
                        out_path=a.out, members=a.members)
    print(f"[train_selector] trained on {int(ok.sum())} runs ({int((~ok).sum())} infeasible skipped), "
          f"final loss {loss:.4f}, model {a.out}", flush=True)
    mibwt.selector_load(a.out)
    # held-out evaluation: predicted choice against the measured grid
    report = evaluate(mibwt, st, X[ntr:], MS[ntr:], FX[ntr:], grid, a.out)
    if a.large_holdout:
        k = a.large_holdout
        report["large_holdout"] = evaluate(mibwt, st, X[-k:], MS[-k:], FX[-k:], grid, a.out)
    if a.eval_workloads:
        from mibwt import gen
        named = {}
        for name in a.eval_workloads.split(","):
            if name in named_data:
                Xw, msw, fmsw = named_data[name]
            else:
                T = gen.workload(name)
                Xw, _, msw, fmsw = run_grid(mibwt, T, grid)
                named_data[name] = (Xw, msw, fmsw)
            n_w = int(round(2 ** (float(Xw[0, 0]) * 32)))  # feature 0 is log2(n) / 32
            ev = evaluate(mibwt, st, Xw[None], msw[None], np.array([fmsw]), grid, a.out)
            c = int(np.argmin(np.where(np.isfinite(msw), msw, np.inf)))
            named[name] = {"n": n_w, "chosen_wp": ev["chosen_wp"][0], "best_wp": list(grid[c]),
                           "best_ms": float(msw[c]), "w5p16_ms": float(fmsw),
                           "chosen_over_best": ev["chosen_over_best_mean"],
                           "speedup_over_w5p16": ev["speedup_over_w5p16_median"]}
            print(f"[train_selector] workload {name}: {named[name]}", flush=True)
        report["named_workloads"] = named
        if a.data:  # the named workloads' grids beside the training runs
            extra = {f"{k}_{f}": v for k, (Xw, msw, fmsw) in named_data.items()
                     for f, v in (("X", Xw), ("MS", msw), ("FX", np.array([fmsw])))}
            np.savez(a.data, X=X, Y=Y, MS=MS, FX=FX, **extra)
    report["train_runs"] = int(ok.sum())
    report["final_loss"] = loss
    print(json.dumps(report), flush=True)
    if a.report:
        with open(a.report, "w") as f:
            json.dump(report, f, indent=1)
    st.close()


def evaluate(mibwt, st, X, MS, FX, grid, model_path):
    """Choice of the trained model on held-out texts (features X[t][candidates][64], measured
    ms, inf for a candidate whose dictionary outgrows the index range); the last candidate is
    the fixed default W=5,P=16.  The choice is the library's: mean predicted time of the
    ensemble members, smoothed over each grid candidate's P-neighbours, argmin.  An infeasible
    choice is charged the library's fallback: a failed attempt (counted as one more W=5,P=16
    build) plus the W=5,P=16 build."""
    m = mibwt.read_selector_model(model_path)

    def predicted(Xt):
        Xn = (Xt - m["xmean"]) / m["xstd"]
        t = np.mean([st.mlp_predict(p, Xn)[:, 2] for p in m["params"]], axis=0)
        return mibwt.smooth_along_p(t * m["ystd"][2] + m["ymean"][2])

    ratios_best, ratios_fixed, ratios_oracle, picks = [], [], [], []
    for t in range(len(X)):
        c = int(np.argmin(predicted(X[t])))
        picks.append(list(grid[c]))
        chosen = MS[t][c] if np.isfinite(MS[t][c]) else 2.0 * FX[t]
        ratios_best.append(chosen / MS[t].min())
        ratios_fixed.append(FX[t] / chosen)
        ratios_oracle.append(FX[t] / MS[t].min())
    out = {"holdout_texts": len(X),
           "chosen_over_best_median": float(np.median(ratios_best)),
           "chosen_over_best_mean": float(np.mean(ratios_best)),
           "within_10pct_of_best": float(np.mean(np.array(ratios_best) <= 1.10))}
    out["speedup_over_w5p16_median"] = float(np.median(ratios_fixed))
    out["speedup_over_w5p16_geomean"] = float(np.exp(np.mean(np.log(ratios_fixed))))
    out["best_grid_over_w5p16_geomean"] = float(np.exp(np.mean(np.log(ratios_oracle))))
    out["chosen_wp"] = picks
    return out


if __name__ == "__main__":
    main()
