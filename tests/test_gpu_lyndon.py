"""GPU parity of the Lyndon-grammar BWT: the Lyndon array of the block-local passes plus
merge tree against the oracle's right-to-left loop, and the BWT against the SA-IS oracle."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def texts():
    rng = np.random.default_rng(21)
    out = {
        "one": b"x", "two": b"ba", "unary": b"a" * 3000, "period2": b"ab" * 2000,
        "period5": b"abcab" * 900, "abb": b"ab" * 700 + b"b" + b"ab" * 300,
        "all_bytes": bytes(range(256)) * 9, "descending": bytes(range(255, -1, -1)) * 5,
        "zeros": bytes(4097),
    }
    for k in range(12):
        n = int(rng.integers(1, 5000))
        out[f"rand{k}"] = bytes(rng.integers(0, 1 + k % 4, size=n).astype(np.uint8) * (1 + 60 * (k % 3)))
    base = rng.choice(np.frombuffer(b"ACGT", dtype=np.uint8), size=3000)
    out["repetitive"] = bytes(np.concatenate([base, base[3:], base[:1000], base]))
    out["long_period"] = bytes(rng.choice(np.frombuffer(b"ACGT", dtype=np.uint8), size=37)) * 300
    return out


TEXTS = texts()


@pytest.mark.parametrize("name", sorted(TEXTS))
def test_lyndon_array(storage, name):
    T = TEXTS[name]
    assert (storage.lyndon_array(T).astype(np.int64) == oracle.lyndon_array(T)).all()


@pytest.mark.parametrize("name", sorted(TEXTS))
def test_lyndon_bwt(storage, mibwt, name):
    T = np.frombuffer(TEXTS[name], dtype=np.uint8)
    L, idx = storage.bwt(T, "lyndon")
    O, oidx = oracle.bwt(T)
    assert idx == oidx
    assert (mibwt.cubwt_to_reference(L, idx) == O).all()


def test_lyndon_dna32m(mibwt):
    from mibwt import gen
    T = gen.workload("dna32m")
    O, oidx = oracle.bwt(T)
    with mibwt.Storage(len(T)) as st:
        lam = st.lyndon_array(T[: 1 << 22])
        assert (lam.astype(np.int64) == oracle.lyndon_array(T[: 1 << 22])).all()
        L, idx = st.bwt(T, "lyndon")
    assert idx == oidx and (mibwt.cubwt_to_reference(L, idx) == O).all()
