"""Batched BWT of many independent texts of uneven length (alphabets 4 and 256): every
text bit-exact against the oracle, for fixed (W,P), the learned choice and the Lyndon method."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def batch_texts(count=40, seed=31):
    from mibwt import gen
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n_total = int(np.exp(rng.uniform(np.log(1 << 10), np.log(1 << 20))))
        copies = int(rng.integers(2, 9))
        base = max(64, n_total // copies)
        rate = float(np.exp(rng.uniform(np.log(1e-5), np.log(1e-2))))
        T = gen.repetitive_dna(base, copies, rate, rate / 10, 10, seed=seed * 1000 + i)
        if i % 3 == 0:  # 256-symbol texts
            T = (T.astype(np.uint32) * 97 + np.arange(len(T)) % 7 * 31).astype(np.uint8)
        out.append(T)
    return out


@pytest.mark.parametrize("method,W,P", [("pfp", 5, 16), ("pfp", 0, 0), ("lyndon", 5, 16)])
def test_batch_bit_exact(mibwt, method, W, P):
    texts = batch_texts()
    with mibwt.Storage(max(len(t) for t in texts)) as st:
        res = st.batch_bwt(texts, method, W, P, workers=4)
    for T, (L, idx) in zip(texts, res):
        O, oidx = oracle.bwt(T)
        assert idx == oidx and (mibwt.cubwt_to_reference(L, idx) == O).all()


def test_batch_rejects_bad_text(mibwt):
    with mibwt.Storage(1000) as st:
        with pytest.raises(mibwt.BwtError):
            st.batch_bwt([np.zeros(10, np.uint8), np.zeros(5000, np.uint8)])
