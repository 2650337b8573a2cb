"""Golden vectors (tests/golden/bwt_small.json, written by make_golden.py by brute force)
against the oracle, and the reference-format layout of the mississippi known answer."""
import json
import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "bwt_small.json")))


@pytest.mark.parametrize("case", GOLDEN, ids=[str(i) for i in range(len(GOLDEN))])
def test_oracle_matches_golden(case):
    T = bytes(case["text"])
    O, idx = oracle.bwt(T)
    assert O.tolist() == case["reference_format"] and idx == case["primary_index"]
    assert oracle.ref_to_cubwt(O, idx).tolist() == case["libcubwt_L"]
    for W, P in [(1, 2), (3, 5), (5, 16)]:
        O2, idx2, _ = oracle.pfp_bwt(T, W, P)
        assert O2.tolist() == case["reference_format"] and idx2 == idx
    O3, idx3 = oracle.lyndon_bwt(T)
    assert O3.tolist() == case["reference_format"] and idx3 == idx


def test_reference_output_format_kat():
    # "n\n" then "%d " per symbol, no final newline (PFPk:4776-4790)
    O, _ = oracle.bwt(b"mississippi")
    text = f"{len(O)}\n" + "".join(f"{v} " for v in O)
    assert text == "11\n112 115 115 109 105 112 105 115 115 105 105 "


@pytest.mark.gpu
@pytest.mark.parametrize("case", GOLDEN, ids=[str(i) for i in range(len(GOLDEN))])
def test_gpu_matches_golden(storage, mibwt, case):
    T = np.array(case["text"], dtype=np.uint8)
    for method, kw in [("pfp", {"W": 5, "P": 16}), ("pfp", {"W": 2, "P": 3}), ("lyndon", {})]:
        L, idx = storage.bwt(T, method, **kw)
        assert idx == case["primary_index"] and L.tolist() == case["libcubwt_L"], method
