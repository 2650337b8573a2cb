"""GPU trigger scan (Compute_D1/D2, PFPk:443-571) against the oracle's trigger set.

The single-pass kernel keeps the exact window value (32 bits for W <= 24, 64 bits for
W <= 56) and tests divisibility by multiplication with the inverse of P's odd part; wider
windows use the two-pass modular kernel.  Texts cross several 8192-window tiles, use the
whole byte range (largest exact values), repeat the first window (second trigger rule) and,
for P = 1 and P = 2, exceed the first-guess capacity of the position buffer (rescan path).
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

WS = [1, 2, 3, 5, 10, 16, 20, 24, 25, 35, 56, 57, 64]
PS = [1, 2, 4, 8, 16, 32, 11, 13, 81, 96, 1024, 65535]


def texts(seed):
    rng = np.random.default_rng(seed)
    out = [rng.integers(0, 256, size=int(rng.integers(20000, 40000))).astype(np.uint8),
           rng.choice(np.frombuffer(b"ACGT", dtype=np.uint8), size=30011)]
    head = rng.integers(250, 256, size=64).astype(np.uint8)
    body = np.tile(head, 400)
    body[rng.integers(0, len(body), size=300)] = rng.integers(0, 256, size=300)
    out.append(body)  # the first window recurs often; large byte values
    return out


@pytest.mark.parametrize("W", WS)
def test_trigger_positions_match_oracle(storage, W):
    for i, T in enumerate(texts(W)):
        for P in PS:
            E = storage.pfp_triggers(T, W, P)
            ref = oracle.pfp_triggers(T, W, P)
            assert len(E) == len(ref), (W, P, i)
            assert (E == ref).all(), (W, P, i)


def test_rescan_when_triggers_exceed_first_capacity(mibwt):
    rng = np.random.default_rng(7)
    T = rng.integers(0, 256, size=300000).astype(np.uint8)
    with mibwt.Storage(len(T)) as st:
        for W, P in [(5, 1), (5, 2), (30, 1)]:
            E = st.pfp_triggers(T, W, P)
            ref = oracle.pfp_triggers(T, W, P)
            assert len(E) == len(ref) and (E == ref).all(), (W, P)
        # the full build takes the same path and stays exact
        L, idx = st.bwt(T[:100000], "pfp", 5, 1)
        O, oidx = oracle.bwt(T[:100000])
        assert idx == oidx and (mibwt.cubwt_to_reference(L, idx) == O).all()


def test_short_texts(storage):
    rng = np.random.default_rng(9)
    for n in [1, 2, 5, 6, 31, 32, 33, 63, 64, 65, 8191, 8192, 8193]:
        T = rng.integers(0, 4, size=n).astype(np.uint8) * 60
        for W, P in [(5, 16), (5, 1), (1, 2), (35, 81), (60, 7)]:
            E = storage.pfp_triggers(T, W, P)
            ref = oracle.pfp_triggers(T, W, P)
            assert len(E) == len(ref) and (E == ref).all(), (n, W, P)
