"""Benchmark: BWT build MB/s on one MI355X (prefix-free parsing, bit-exact).

One step = one complete BWT build of the workload text, already resident in HBM, into a
device output buffer (mibwt_pfp_bwt_device).  Default workload: the 1 GiB repetitive DNA
collection (8 MiB uniform-ACGT base x 128 copies, 0.02% mutations split 10:1 substitutions to
indels of length 1..10).  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload dna1g] [--W 5 --P 16]
                  [--wp auto|fixed] [--lyndon-steps 1] [--batch-texts 256] [--grid-steps 0]
                  [--control markov256m] [--control-steps 3] [--no-verify] [--no-cpu-baseline]
                  [--method pfp|lyndon]

With --wp auto (default) every timed step runs the learned (W,P) selector on the resident text
and then the build with its choice; `value` is that rate.  Beside it: the fixed W/P build
("fixed_wp", "speedup_vs_fixed"), the W=35,P=81 build ("w35_p81"), the low-repetition control
text with fixed and learned (W,P) ("control"), the Lyndon-grammar BWT of the same text, compared
byte for byte ("lyndon"), the host-buffer entry with its PCIe copies ("pcie_inclusive"), a
batch of independent texts of uneven length ("batch"), the peak device memory of the storage
("device_memory_bytes"), and the multi-threaded C++ restatement of the reference pipeline on the
host cores ("cpu_baseline").  --method lyndon times the Lyndon-grammar build of the workload as
the step instead (its own JSON line; used for the Lyndon profiles).
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "masterarbeit_parallele_konstruktion_der_burrows-wheeler-transformation_auf_der_gpu_amd")
HBM_PEAK_GBS = 8000.0
METRIC = "BWT build MB/s on 1 MI355X, bit-exact; learned (W,P) vs fixed W=5,P=16"


def load_pkg():
    if "mibwt" in sys.modules:
        return sys.modules["mibwt"]
    spec = importlib.util.spec_from_file_location("mibwt", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["mibwt"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(T, W, P, sais_sample_bytes):
    """The multi-threaded C++ restatement of the reference PFP pipeline (oracle/cpu_pfp.cpp) on
    the whole workload with every host core, and the single-threaded SA-IS oracle on a prefix."""
    sys.path.insert(0, ROOT)
    import oracle  # checker / CPU baseline only
    threads = len(os.sched_getaffinity(0))
    t0 = time.perf_counter()
    oracle.cpu_pfp_bwt(T, W, P, threads)
    dt = time.perf_counter() - t0
    sample = T[:sais_sample_bytes]
    t1 = time.perf_counter()
    oracle.bwt(sample)
    ds = time.perf_counter() - t1
    return {"value": round(len(T) / 1e6 / dt, 3), "unit": "MB/s", "cores": threads, "kind": "port",
            "nproc": os.cpu_count(), "cpu_model": cpu_model(),
            "sample": f"whole workload ({len(T)} B), PFP W={W},P={P}, std::thread restatement of the "
                      f"reference pipeline on {threads} threads, {dt:.2f} s",
            "sais_1thread": {"value": round(len(sample) / 1e6 / ds, 3), "unit": "MB/s",
                             "sample": f"first {len(sample) >> 20} MiB, SA-IS oracle, 1 thread, {ds:.2f} s"}}


def lyndon_only(a, mibwt, st, T, dT, dO):
    """The Lyndon-grammar build of the workload as the timed step (one JSON line)."""
    n = len(T)
    idx = None
    for i in range(a.warmup):
        idx = st.bwt_device(dT.address, dO.address, n, "lyndon")
    mibwt.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        idx = st.bwt_device(dT.address, dO.address, n, "lyndon")
    mibwt.synchronize()
    ms = (time.perf_counter() - t0) * 1000.0 / a.steps
    s = st.stats()
    verified = None
    if not a.no_verify:
        dU = mibwt.DeviceBuffer(n)
        st.unbwt_device(dO.address, dU.address, n, idx)
        verified = "unbwt_roundtrip_ok_gpu" if bool((dU.download() == T).all()) else "MISMATCH"
        dU.free()
    out = {"metric": "Lyndon-grammar BWT build MB/s on 1 MI355X, bit-exact", "value": round(n / 1e6 / (ms / 1000.0), 2),
           "unit": "MB/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(ms, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
           "config": {"workload": a.workload, "n_bytes": n, "method": "lyndon", "parallelism": "single-gpu"},
           "induction": {1: "grammar (distinct labels)", 2: "position-level"}.get(int(s["induction"])),
           "grammar": {"labels": s["distinct_phrases"], "child_slots": s["dict_length"], "roots": s["triggers"],
                       "up_contexts": s["mixed_entries"], "rounds": s["doubling_rounds_parse"], "items": s["phrases"]},
           "retries": s["retries"], "device_memory_bytes": int(s["device_bytes"]),
           "phases": {p["name"]: round(p["ms"], 3) for p in s["phases"]}, "verified": verified}
    print(json.dumps(out), flush=True)
    st.close()
    if verified == "MISMATCH":
        sys.exit(1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="dna1g")
    ap.add_argument("--W", type=int, default=5)
    ap.add_argument("--P", type=int, default=16)
    ap.add_argument("--no-verify", action="store_true")
    ap.add_argument("--batch-texts", type=int, default=256,
                    help="batched leg: this many independent texts of 64 KiB-16 MiB (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--sais-sample-mib", type=int, default=32)
    ap.add_argument("--wp", choices=["auto", "fixed"], default="auto",
                    help="auto: learned (W,P) choice timed per step (value), fixed W/P reported beside it")
    ap.add_argument("--w35-steps", type=int, default=5, help="timed W=35,P=81 builds (0: skip)")
    ap.add_argument("--control", default="markov256m", help="low-repetition control text ('' to skip)")
    ap.add_argument("--control-steps", type=int, default=3)
    ap.add_argument("--lyndon-steps", type=int, default=1, help="timed Lyndon-grammar builds (0: skip)")
    ap.add_argument("--pcie-steps", type=int, default=2, help="timed host-buffer builds (0: skip)")
    ap.add_argument("--tuning-steps", type=int, default=5,
                    help="timed fixed builds with the default launch shapes (0: skip)")
    ap.add_argument("--grid-steps", type=int, default=0,
                    help="also time every (W,P) candidate with this many builds (0: skip)")
    ap.add_argument("--radix-pass-items", type=int, default=1 << 25,
                    help="time the whole radix digit pass on this many random pairs (0: skip)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also launch a plain 256 MiB device copy kernel (byte-counter reference)")
    ap.add_argument("--method", choices=["pfp", "lyndon"], default="pfp",
                    help="lyndon: time the Lyndon-grammar build as the step (checked by the GPU inverse "
                         "BWT); the PFP legs are skipped")
    a = ap.parse_args()

    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path. Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    mibwt = load_pkg()
    from mibwt import gen

    log(f"generating workload {a.workload}")
    T = gen.workload(a.workload)
    n = len(T)
    log(f"n = {n} bytes")
    # device buffers through the library's own HIP runtime (the text stays resident in HBM)
    dT = mibwt.DeviceBuffer.from_numpy(T)
    dO = mibwt.DeviceBuffer(n)
    mibwt.synchronize()
    st = mibwt.Storage(n)
    if a.method == "lyndon":
        lyndon_only(a, mibwt, st, T, dT, dO)
        return

    def run(step, steps, warmup, label, size=n):
        idx = None
        for i in range(warmup):
            idx = step()
            log(f"{label} warmup {i} done ({st.stats()['total_ms']:.2f} ms build)")
        mibwt.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            idx = step()
        mibwt.synchronize()
        ms = (time.perf_counter() - t0) * 1000.0 / steps
        log(f"{label}: {ms:.3f} ms/step, {size / 1e3 / ms:.1f} MB/s")
        return ms, idx

    # fixed (W, P)
    def step_fixed():
        return st.bwt_device(dT.address, dO.address, n, "pfp", a.W, a.P)

    ms_fixed, idx = run(step_fixed, a.steps, a.warmup, f"fixed W={a.W},P={a.P}")
    stats_fixed = st.stats()
    O_fixed = dO.download() if not a.no_verify else None

    # learned (W, P): W = P = 0 asks the library to run the selector's feature pass and MLP
    # inference on the resident text before the build, inside the timed step
    chosen = {}

    def step_tuned():
        idx_t = st.bwt_device(dT.address, dO.address, n, "pfp", 0, 0)
        s_t = st.stats()
        chosen.update(W=int(s_t["W"]), P=int(s_t["P"]))
        return idx_t

    if a.wp == "auto":
        ms_tuned, idx_t = run(step_tuned, a.steps, a.warmup, "learned (W,P)")
        assert idx_t == idx, "primary index differs between (W,P) choices"
        stats = st.stats()
        W_run, P_run = chosen["W"], chosen["P"]
        ms_per_step = ms_tuned
    else:
        stats = stats_fixed
        W_run, P_run = a.W, a.P
        ms_per_step = ms_fixed
    value = n / 1e6 / (ms_per_step / 1000.0)
    O_run = dO.download() if not a.no_verify else None
    # the learned choice as a fixed build: the difference is the selection's own cost
    selection = None
    if a.wp == "auto" and (W_run, P_run) != (a.W, a.P):
        ms_chosen, _ = run(lambda: st.bwt_device(dT.address, dO.address, n, "pfp", W_run, P_run), a.steps, 1,
                           f"fixed W={W_run},P={P_run} (the learned choice)")
        selection = {"chosen_fixed_ms_per_step": round(ms_chosen, 4),
                     "selection_ms": round(ms_per_step - ms_chosen, 4)}
    if O_run is not None and not (O_run == O_fixed).all():
        log("BWT differs between the fixed and the learned (W,P)")
        sys.exit(1)

    # the other fixed (W, P) of the configuration
    w35 = None
    if a.w35_steps > 0:
        ms_35, idx_35 = run(lambda: st.bwt_device(dT.address, dO.address, n, "pfp", 35, 81), a.w35_steps, 1,
                            "fixed W=35,P=81")
        assert idx_35 == idx, "primary index differs between (W,P) choices"
        s35 = st.stats()
        w35 = {"value": round(n / 1e6 / (ms_35 / 1000.0), 2), "ms_per_step": round(ms_35, 4),
               "phases": {p["name"]: round(p["ms"], 3) for p in s35["phases"]},
               "phrases": s35["phrases"], "distinct_phrases": s35["distinct_phrases"]}
        if O_fixed is not None and not (dO.download() == O_fixed).all():
            log("BWT differs between W=5,P=16 and W=35,P=81")
            sys.exit(1)

    # device memory the storage holds after the PFP builds (its buffers only grow)
    device_mem = int(st.stats()["device_bytes"])

    # launch shapes: the autotuner's table (tuning_table.txt, when present) against the defaults
    tuning = {"table": os.path.exists(mibwt.TUNING_PATH), "shape": mibwt.tuning_shape(n)}
    if a.tuning_steps > 0:
        mibwt.tuning_use(False)
        tuning["default_shape"] = mibwt.tuning_shape(n)
        ms_def, idx_def = run(step_fixed, a.tuning_steps, 1, f"fixed W={a.W},P={a.P}, default shapes")
        mibwt.tuning_use(True)
        assert idx_def == idx
        tuning.update(default_ms_per_step=round(ms_def, 4), tuned_ms_per_step=round(ms_fixed, 4),
                      speedup=round(ms_def / ms_fixed, 4))
        # where the shapes matter most: a short text (1 MiB of the workload), tuned against default
        m = 1 << 20
        small = {}
        for label, on in (("default", False), ("tuned", True)):
            mibwt.tuning_use(on)
            small[label + "_shape"] = mibwt.tuning_shape(m)
            ms_s, _ = run(lambda: st.bwt_device(dT.address, dO.address, m, "pfp", a.W, a.P), 20, 2,
                          f"1 MiB prefix, {label} shapes", size=m)
            small[label + "_ms_per_step"] = round(ms_s, 4)
        mibwt.tuning_use(True)
        small["speedup"] = round(small["default_ms_per_step"] / small["tuned_ms_per_step"], 4)
        tuning["short_text_1mib"] = small

    # Lyndon-grammar BWT of the same text
    lyndon = None
    if a.lyndon_steps > 0:
        def step_lyndon():
            return st.bwt_device(dT.address, dO.address, n, "lyndon")
        ms_l, idx_l = run(step_lyndon, a.lyndon_steps, 1, "lyndon")
        sl = st.stats()
        lyndon = {"value": round(n / 1e6 / (ms_l / 1000.0), 2), "ms_per_step": round(ms_l, 4),
                  "steps": a.lyndon_steps, "retries": sl["retries"],
                  "induction": {1: "grammar (distinct labels)", 2: "position-level"}.get(int(sl["induction"])),
                  "grammar": {"labels": sl["distinct_phrases"], "child_slots": sl["dict_length"],
                              "roots": sl["triggers"], "up_contexts": sl["mixed_entries"],
                              "rounds": sl["doubling_rounds_parse"], "items": sl["phrases"]},
                  "device_memory_bytes": int(sl["device_bytes"]),
                  "phases": {p["name"]: round(p["ms"], 3) for p in sl["phases"]}}
        if O_fixed is not None:
            same = idx_l == idx and bool((dO.download() == O_fixed).all())
            lyndon["identical_to_pfp"] = same
            if not same:
                log("Lyndon BWT differs from the PFP BWT")
                sys.exit(1)

    # every candidate timed alone: how far the learned choice is from the best grid point
    grid = None
    if a.grid_steps > 0:
        res = {}
        for (W, P) in mibwt.wp_grid():
            ms_c, idx_c = run(lambda W=W, P=P: st.bwt_device(dT.address, dO.address, n, "pfp", W, P),
                              a.grid_steps, 1, f"grid W={W},P={P}")
            assert idx_c == idx, "primary index differs between (W,P) choices"
            res[(W, P)] = ms_c
        best = min(res, key=res.get)
        grid = {"candidates": len(res), "best_wp": list(best), "best_ms": round(res[best], 4),
                "chosen_build_ms": round(res[(W_run, P_run)], 4),
                "chosen_over_best": round(res[(W_run, P_run)] / res[best], 4),
                "learned_step_over_best": round(ms_per_step / res[best], 4),
                "fixed_over_best": round(ms_fixed / res[best], 4),
                "ms": {f"{W},{P}": round(v, 3) for (W, P), v in res.items()}}

    # dominant kernel: the LSD radix sort's digit pass, every launch bracketed with HIP events
    # on the storage stream during one extra (untimed) build with the fixed (W,P), the
    # configuration the hardware-counter profile in profiles/ was taken on
    st.set_kernel_timing("radix_scatter")
    step_fixed()
    ks = st.stats()
    st.set_kernel_timing("trig_mask")
    step_fixed()
    kscan = st.stats()
    kscan_run = None
    if (W_run, P_run) != (a.W, a.P):  # the scan of the configuration the value is measured on
        st.bwt_device(dT.address, dO.address, n, "pfp", W_run, P_run)
        kscan_run = st.stats()
    st.set_kernel_timing(None)
    if a.calibrate_copy:
        cms, cbytes = st.time_kernel("copy", reps=3, nbytes=float(256 << 20))
        log(f"copy kernel: {cbytes / (cms / 1000) / 1e9 * 2:.1f} GB/s (read+write) over {cbytes:.0f} B")
    # the whole digit pass (digit counts, their scan, downsweep) on uniformly random pairs
    radix_whole = None
    if a.radix_pass_items > 0:
        rp_items = a.radix_pass_items
        rp_ms, rp_bytes = st.time_kernel("radix", 3, float(rp_items))
        radix_whole = {"items": rp_items, "ms_per_pass": round(rp_ms, 4),
                       "GBps": round(rp_bytes / (rp_ms / 1e3) / 1e9, 1),
                       "frac": round(rp_bytes / (rp_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                       "kernels": "k_radix_up + k_scan_chain + k_radix_pass"}
    avg_ms = ks["kt_ms_total"] / max(1, ks["kt_launches"])
    avg_bytes = ks["kt_bytes_total"] / max(1, ks["kt_launches"])
    achieved = avg_bytes / (avg_ms / 1000.0) / 1e9
    scan_gbs = kscan["kt_bytes_total"] / (kscan["kt_ms_total"] / 1000.0) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_radix_pass.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        # the counters were collected on one workload at W=5,P=16; other workloads report null
        if str(tj.get("workload", "")).split(",")[0] == a.workload and (a.W, a.P) == (5, 16):
            traffic = tj.get("hbm_bytes_per_launch")
    roofline = {"bound": "hbm", "kernel": "k_radix_pass", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic, "launches": ks["kt_launches"], "avg_launch_us": round(avg_ms * 1000, 2),
                "algorithmic_bytes_per_launch": round(avg_bytes), "wp": [a.W, a.P],
                "radix_whole_pass": radix_whole}

    verified = None
    if not a.no_verify:
        t = time.perf_counter()
        # inverse BWT on the device (libcubwt_unbwt's kernels), compared on the host
        st.bwt_device(dT.address, dO.address, n, "pfp", a.W, a.P)
        dU = mibwt.DeviceBuffer(n)
        st.unbwt_device(dO.address, dU.address, n, idx)
        back = dU.download()
        dU.free()
        ok = bool((back == T).all())
        verified = "unbwt_roundtrip_ok_gpu" if ok else "MISMATCH"
        log(f"verification (inverse BWT on the GPU): {verified} in {time.perf_counter() - t:.1f} s")
        if not ok:
            sys.exit(1)

    # the host-buffer entry (libcubwt-style: host text in, host BWT out, learned (W,P)): the
    # PCIe copies are inside this rate, which is therefore not the metric
    pcie = None
    if a.pcie_steps > 0:
        def step_host():
            L, i = st.bwt(T, "pfp", 0, 0)
            return i
        ms_h, idx_h = run(step_host, a.pcie_steps, 1, "host buffers (PCIe included)")
        assert idx_h == idx
        pcie = {"value": round(n / 1e6 / (ms_h / 1000.0), 2), "ms_per_step": round(ms_h, 4),
                "entry": "mibwt_pfp_bwt (W=P=0), host text in, host BWT out"}
    dT.free()
    dO.free()

    # low-repetition control: the same fixed and learned (W,P) on the order-2 Markov text
    control = None
    if a.control:
        log(f"generating control {a.control}")
        C = gen.workload(a.control)
        nc = len(C)
        dC = mibwt.DeviceBuffer.from_numpy(C)
        dCO = mibwt.DeviceBuffer(nc)
        stc = mibwt.Storage(nc)
        res = {}
        for label, (W, P) in (("fixed", (a.W, a.P)), ("fixed_w35_p81", (35, 81)), ("learned", (0, 0))):
            def step_c(W=W, P=P):
                return stc.bwt_device(dC.address, dCO.address, nc, "pfp", W, P)
            step_c()
            mibwt.synchronize()
            t0 = time.perf_counter()
            for _ in range(a.control_steps):
                step_c()
            mibwt.synchronize()
            ms_c = (time.perf_counter() - t0) * 1000.0 / a.control_steps
            sc = stc.stats()
            res[label] = {"W": int(sc["W"]), "P": int(sc["P"]), "ms_per_step": round(ms_c, 4),
                          "value": round(nc / 1e6 / (ms_c / 1000.0), 2)}
            log(f"control {label}: {ms_c:.3f} ms/step")
        control = {"workload": a.control, "n_bytes": nc, **res,
                   "speedup_learned_vs_fixed": round(res["fixed"]["ms_per_step"] / res["learned"]["ms_per_step"], 4)}
        stc.close()
        dC.free()
        dCO.free()

    # batched leg: independent texts of uneven length (64 KiB-16 MiB, log-uniform; a random base
    # x 2..64 copies, mutation 0.001%-1%; alphabet 4 or 256), (W,P) chosen per build by the
    # selector.  Device-resident (mibwt_batch_bwt_device: texts and BWTs in HBM) is the batch
    # rate; the host-buffer entry (PCIe copies included) is reported beside it.  A sample of the
    # outputs is checked by the inverse BWT.
    batch = None
    if a.batch_texts > 0:
        import numpy as np
        texts = gen.batch_collection(a.batch_texts)
        tot = sum(len(t) for t in texts)
        offs = np.concatenate([[0], np.cumsum([len(t) for t in texts])]).astype(np.int64)
        dB = mibwt.DeviceBuffer.from_numpy(np.concatenate(texts))
        dBO = mibwt.DeviceBuffer(tot)
        stb = mibwt.Storage(max(len(t) for t in texts))
        stb.batch_bwt_device(dB.address, offs, dBO.address, 0, 0)  # warm the buffers
        mibwt.synchronize()
        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            bidx = stb.batch_bwt_device(dB.address, offs, dBO.address, 0, 0)
        mibwt.synchronize()
        dms = (time.perf_counter() - t0) * 1000.0 / reps
        sb = stb.stats()
        checked = 0
        dU = mibwt.DeviceBuffer(int(max(len(t) for t in texts)))
        for i in range(0, len(texts), max(1, len(texts) // 16)):
            stb.unbwt_device(dBO.address + int(offs[i]), dU.address, len(texts[i]), int(bidx[i]))
            if not (dU.download(count=len(texts[i])) == texts[i]).all():
                log(f"batched text {i}: inverse BWT differs")
                sys.exit(1)
            checked += 1
        dU.free()
        # the same texts built one by one through the single-text device entry (learned (W,P)
        # per text): what one multi-text build saves
        stb.bwt_device(dB.address, dBO.address, len(texts[0]), "pfp", 0, 0)
        mibwt.synchronize()
        t0 = time.perf_counter()
        for i, t in enumerate(texts):
            stb.bwt_device(dB.address + int(offs[i]), dBO.address + int(offs[i]), len(t), "pfp", 0, 0)
        mibwt.synchronize()
        per_ms = (time.perf_counter() - t0) * 1000.0
        t0 = time.perf_counter()
        res = stb.batch_bwt(texts, "pfp", 0, 0)
        hms = (time.perf_counter() - t0) * 1000.0
        if any(int(r[1]) != int(bidx[i]) for i, r in enumerate(res)):
            log("host and device batch entries give different primary indices")
            sys.exit(1)
        batch = {"texts": len(texts), "bytes": tot, "ms": round(dms, 2), "value": round(tot / 1e6 / (dms / 1000.0), 1),
                 "unit": "MB/s", "entry": "mibwt_batch_bwt_device (texts and BWTs in HBM)",
                 "wp_choice": "learned per build", "last_build_wp": [int(sb["W"]), int(sb["P"])],
                 "device_memory_bytes": int(sb["device_bytes"]),
                 "host_buffers": {"ms": round(hms, 2), "value": round(tot / 1e6 / (hms / 1000.0), 1),
                                  "entry": "mibwt_batch_bwt (host texts in, host BWTs out, PCIe included)"},
                 "per_text": {"ms": round(per_ms, 2), "value": round(tot / 1e6 / (per_ms / 1000.0), 1),
                              "entry": "mibwt_pfp_bwt_device once per text (learned (W,P) per text)"},
                 "speedup_vs_per_text": round(per_ms / dms, 3),
                 "verified_texts": checked}
        log(f"batch: {len(texts)} texts, {tot / 1e6:.0f} MB: device {dms:.1f} ms = {batch['value']} MB/s, "
            f"host buffers {hms:.1f} ms = {batch['host_buffers']['value']} MB/s, one by one {per_ms:.1f} ms")
        stb.close()
        dB.free()
        dBO.free()

    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline(T, a.W, a.P, min(n, a.sais_sample_mib << 20))
        log(f"cpu baseline: {cpu['value']} MB/s on {cpu['cores']} threads; SA-IS 1 thread "
            f"{cpu['sais_1thread']['value']} MB/s")

    phases = {p["name"]: {"ms": round(p["ms"], 3),
                          "GBps": round(p["bytes"] / max(p["ms"], 1e-9) / 1e6, 1)} for p in stats["phases"]}
    phases["scan_kernel"] = {"GBps": round(scan_gbs, 1), "frac": round(scan_gbs / HBM_PEAK_GBS, 4),
                             "wp": [a.W, a.P]}
    if kscan_run is not None:
        g = kscan_run["kt_bytes_total"] / (kscan_run["kt_ms_total"] / 1000.0) / 1e9
        phases["scan_kernel_run"] = {"GBps": round(g, 1), "frac": round(g / HBM_PEAK_GBS, 4), "wp": [W_run, P_run]}
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "MB/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": a.workload, "n_bytes": n, "method": "pfp", "W": W_run, "P": P_run,
                   "wp_choice": "learned" if a.wp == "auto" else "fixed", "parallelism": "single-gpu"},
        "fixed_wp": {"W": a.W, "P": a.P, "value": round(n / 1e6 / (ms_fixed / 1000.0), 2),
                     "ms_per_step": round(ms_fixed, 4),
                     "phases": {p["name"]: round(p["ms"], 3) for p in stats_fixed["phases"]}},
        "speedup_vs_fixed": round(ms_fixed / ms_per_step, 4),
        "selection": selection,
        "w35_p81": w35,
        "tuning": tuning,
        "control": control,
        "grid": grid,
        "lyndon": lyndon,
        "pcie_inclusive": pcie,
        "batch": batch,
        "device_memory_bytes": {"pfp": device_mem, "bytes_per_text_byte": round(device_mem / n, 2)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "verified": verified,
        "phases": phases,
        "counts": {k: stats[k] for k in ("triggers", "phrases", "distinct_phrases", "dict_length",
                                          "mixed_entries", "doubling_rounds_dict", "doubling_rounds_parse")},
    }
    print(json.dumps(out), flush=True)
    st.close()


if __name__ == "__main__":
    main()
