"""The batched PFP pipeline on device-resident texts (mibwt_batch_bwt_device): many texts of one
build share every kernel launch, each keeps its own terminator, first-window rule and primary
index.  Every text is compared with the SA-IS oracle, including texts shorter than W, equal
texts in one build (equal phrases of different texts must stay apart), texts made of one
repeated byte, all 256 byte values, and builds split across several chunks."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def mixed_texts(seed=3):
    from mibwt import gen
    rng = np.random.default_rng(seed)
    out = [np.frombuffer(b"x", np.uint8), np.frombuffer(b"ba", np.uint8), np.frombuffer(b"ACGTA", np.uint8),
           np.full(3000, 65, np.uint8), np.frombuffer(bytes(range(256)) * 9, np.uint8),
           np.frombuffer(b"ACGTA" * 700, np.uint8)]
    base = gen.repetitive_dna(5000, 6, 0.002, 0.0002, 10, seed=seed)
    out += [base, base.copy(), base[:-1].copy(), base.copy()]  # equal texts in one build
    for i in range(30):
        n = int(rng.integers(1, 40000))
        k = int(rng.integers(1, 6))
        T = (rng.integers(0, k, size=n) * int(rng.integers(1, 60))).astype(np.uint8)
        if i % 4 == 0:
            T = np.concatenate([T, T, T[: n // 3]])
        out.append(T)
    out.append(np.frombuffer(b"x" * 8192 + b"y", np.uint8))  # a text ending at a sub-tile boundary
    out.append(np.full(8192 - 64, 7, np.uint8))
    return out


def run_batch(mibwt, texts, W, P, chunk=None, monkeypatch=None):
    if chunk is not None:
        monkeypatch.setenv("MIBWT_BATCH_CHUNK", str(chunk))
    offs = np.concatenate([[0], np.cumsum([len(t) for t in texts])]).astype(np.int64)
    dT = mibwt.DeviceBuffer.from_numpy(np.concatenate(texts))
    dO = mibwt.DeviceBuffer(int(offs[-1]))
    with mibwt.Storage(max(len(t) for t in texts)) as st:
        idx = st.batch_bwt_device(dT.address, offs, dO.address, W, P)
        stats = st.stats()
    O = dO.download()
    dT.free()
    dO.free()
    return offs, idx, O, stats


@pytest.mark.parametrize("W,P", [(5, 16), (1, 2), (10, 31), (35, 81), (0, 0)])
def test_batch_device_every_text_matches_oracle(mibwt, W, P):
    texts = mixed_texts()
    offs, idx, O, _ = run_batch(mibwt, texts, W, P)
    for i, T in enumerate(texts):
        Oo, oidx = oracle.bwt(T)
        assert idx[i] == oidx, (i, len(T))
        assert (O[offs[i]:offs[i + 1]] == Oo).all(), (i, len(T))


@pytest.mark.parametrize("chunk", [0, 100000, 1 << 20])
def test_batch_device_chunks(mibwt, monkeypatch, chunk):
    # builds of one text, of a few texts, and of all of them give the same bytes
    texts = mixed_texts(seed=9)
    offs, idx, O, _ = run_batch(mibwt, texts, 5, 16, chunk, monkeypatch)
    for i, T in enumerate(texts):
        Oo, oidx = oracle.bwt(T)
        assert idx[i] == oidx and (O[offs[i]:offs[i + 1]] == Oo).all(), (i, len(T), chunk)


def test_batch_device_rejects_bad_offsets(mibwt):
    dT = mibwt.DeviceBuffer(100)
    with mibwt.Storage(100) as st:
        for offs in ([0, 10, 10], [0, 50, 20], [0, 200]):
            with pytest.raises(mibwt.BwtError) as e:
                st.batch_bwt_device(dT.address, offs, dT.address, 5, 16)
            assert e.value.code == mibwt.BAD_PARAMETER
    dT.free()
