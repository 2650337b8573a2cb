"""Measure the launch-shape table of this GPU and write it next to lib/ (tuning_table.txt).

    python tune.py [--max-bytes 1073741824] [--reps 2] [--out PATH]

Every build afterwards takes its trigger-scan, radix-sort and chained-scan tile shapes and its
initial key width from the row of its length bucket (mibwt_autotune in include/mibwt.h).
"""
from __future__ import annotations

import argparse
import importlib.util
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))


def load():
    if "mibwt" in sys.modules:
        return sys.modules["mibwt"]
    spec = importlib.util.spec_from_file_location("mibwt", os.path.join(HERE, "__init__.py"),
                                                  submodule_search_locations=[HERE])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["mibwt"] = mod
    spec.loader.exec_module(mod)
    return mod


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--max-bytes", type=int, default=1 << 30)
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--out", default=os.path.join(HERE, "tuning_table.txt"))
    a = ap.parse_args()
    mibwt = load()
    t0 = time.perf_counter()
    with mibwt.Storage(a.max_bytes) as st:
        rows = st.autotune(a.max_bytes, a.reps, a.out)
    print(f"tuned {rows} length buckets in {time.perf_counter() - t0:.1f} s -> {a.out}")
    print(open(a.out).read())


if __name__ == "__main__":
    main()
