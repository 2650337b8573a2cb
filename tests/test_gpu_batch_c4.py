"""The batched configuration at its real shape: 4096 independent texts, lengths log-uniform in
[64 KiB, 16 MiB], each a random base x 2..64 copies with mutation 0.001%-1%, alphabets 4 and 256,
in one batch call.  Every text is checked by inverse BWT and symbol histogram; every 64th text is
compared byte for byte with the SA-IS oracle."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

COUNT = 4096


@pytest.fixture(scope="module")
def c4_texts(mibwt):
    from mibwt import gen
    return gen.batch_collection(COUNT, threads=16)


@pytest.mark.timeout(1200)
def test_c4_batch_every_text(mibwt, c4_texts):
    texts = c4_texts
    assert len(texts) == COUNT
    sizes = np.array([len(t) for t in texts])
    assert sizes.min() >= 60 << 10 and sizes.max() <= 17 << 20
    with mibwt.Storage(int(sizes.max())) as st:
        res = st.batch_bwt(texts, "pfp", 0, 0)
        # every text: symbol histogram and the inverse BWT
        for i, (T, (L, idx)) in enumerate(zip(texts, res)):
            assert 1 <= idx <= len(T), i
            assert (np.bincount(L, minlength=256) == np.bincount(T, minlength=256)).all(), i
            assert (st.unbwt(L, idx) == T).all(), i

    # a fixed sample against the SA-IS oracle (independent of the library)
    def check(i):
        O, oidx = oracle.bwt(texts[i])
        L, idx = res[i]
        return idx == oidx and bool((mibwt.cubwt_to_reference(L, idx) == O).all())

    sample = list(range(0, COUNT, 64))
    with ThreadPoolExecutor(16) as ex:
        ok = list(ex.map(check, sample))
    assert all(ok), [i for i, g in zip(sample, ok) if not g]
