"""The launch-shape autotuner: the table file and the shape lookup (CPU), every shape giving the
oracle's BWT, and a measured table on the device (GPU)."""
import os

import numpy as np
import pytest

import oracle

SHAPES = ["1,8,16,32", "2,16,64,48", "4,8,64,40", "8,16,16,64", "1,16,16,64", "8,8,64,56"]


def test_shape_lookup_follows_the_table(mibwt, tmp_path, monkeypatch):
    monkeypatch.delenv("MIBWT_TUNE_FORCE", raising=False)
    path = tmp_path / "t.txt"
    path.write_text("# comment\n16 1 8 16 32\n20 2 16 64 48\n30 8 16 64 64\n99 1 1 1 1\n")
    assert mibwt.lib().mibwt_tuning_load(str(path).encode()) == mibwt.NO_ERROR
    mibwt.tuning_use(True)
    assert mibwt.tuning_shape(1000) == {"scan_sub": 1, "sort_items": 8, "chain_items": 16, "key_bits": 32}
    assert mibwt.tuning_shape(1 << 20) == {"scan_sub": 2, "sort_items": 16, "chain_items": 64, "key_bits": 48}
    assert mibwt.tuning_shape((1 << 20) + 1) == {"scan_sub": 8, "sort_items": 16, "chain_items": 64, "key_bits": 64}
    mibwt.tuning_use(False)  # the defaults
    assert mibwt.tuning_shape(1000) == {"scan_sub": 8, "sort_items": 16, "chain_items": 64, "key_bits": 64}
    monkeypatch.setenv("MIBWT_TUNE_FORCE", "4,8,16,40")
    assert mibwt.tuning_shape(1 << 25) == {"scan_sub": 4, "sort_items": 8, "chain_items": 16, "key_bits": 40}
    bad = tmp_path / "bad.txt"
    bad.write_text("20 3 16 64 48\n")  # 3 sub-tiles is not a shape
    assert mibwt.lib().mibwt_tuning_load(str(bad).encode()) == mibwt.BAD_PARAMETER
    mibwt.tuning_use(True)
    default = os.path.join(os.path.dirname(mibwt.LIB_PATH), "..", "tuning_table.txt")
    if os.path.exists(default):
        assert mibwt.lib().mibwt_tuning_load(default.encode()) == mibwt.NO_ERROR


@pytest.mark.gpu
@pytest.mark.parametrize("shape", SHAPES)
def test_every_shape_gives_the_same_bwt(mibwt, monkeypatch, shape):
    from mibwt import gen
    monkeypatch.setenv("MIBWT_TUNE_FORCE", shape)
    rng = np.random.default_rng(len(shape))
    texts = [gen.repetitive_dna(20000, 9, 0.002, 0.0002, 10, seed=5),
             rng.integers(0, 256, size=150000).astype(np.uint8),
             np.frombuffer(b"ACGTA" * 3000, np.uint8)]
    with mibwt.Storage(1 << 20) as st:
        for T in texts:
            for W, P in [(5, 16), (10, 31)]:
                L, idx = st.bwt(T, "pfp", W, P)
                O, oidx = oracle.bwt(T)
                assert idx == oidx and (mibwt.cubwt_to_reference(L, idx) == O).all(), (shape, len(T), W, P)
        L, idx = st.bwt(texts[0], "lyndon")
        O, oidx = oracle.bwt(texts[0])
        assert idx == oidx and (mibwt.cubwt_to_reference(L, idx) == O).all()
        # the batched pipeline under the same shape
        offs = np.concatenate([[0], np.cumsum([len(t) for t in texts])]).astype(np.int64)
        dT = mibwt.DeviceBuffer.from_numpy(np.concatenate(texts))
        dO = mibwt.DeviceBuffer(int(offs[-1]))
        bidx = st.batch_bwt_device(dT.address, offs, dO.address, 5, 16)
        out = dO.download()
        for i, T in enumerate(texts):
            O, oidx = oracle.bwt(T)
            assert bidx[i] == oidx and (out[offs[i]:offs[i + 1]] == O).all()
        dT.free()
        dO.free()


@pytest.mark.gpu
def test_autotune_writes_a_usable_table(mibwt, tmp_path, monkeypatch):
    from mibwt import gen
    monkeypatch.delenv("MIBWT_TUNE_FORCE", raising=False)
    path = tmp_path / "table.txt"
    with mibwt.Storage(1 << 20) as st:
        rows = st.autotune(1 << 20, 1, str(path))
        assert rows == 5  # buckets 2^16 .. 2^20
        lines = [l.split()[:5] for l in path.read_text().splitlines() if not l.startswith("#")]
        assert [int(l[0]) for l in lines] == [16, 17, 18, 19, 20]
        for l in lines:
            assert int(l[1]) in (1, 2, 4, 8) and int(l[2]) in (8, 16) and int(l[3]) in (16, 64)
            assert 32 <= int(l[4]) <= 64
        sh = mibwt.tuning_shape(1 << 18)
        row = [l for l in lines if int(l[0]) == 18][0]
        assert [sh["scan_sub"], sh["sort_items"], sh["chain_items"], sh["key_bits"]] == [int(x) for x in row[1:]]
        T = gen.repetitive_dna(30000, 8, 0.003, 0.0003, 10, seed=77)
        L, idx = st.bwt(T, "pfp", 5, 16)
    O, oidx = oracle.bwt(T)
    assert idx == oidx and (mibwt.cubwt_to_reference(L, idx) == O).all()
    # back to the committed table (or the defaults) for the tests that follow
    default = os.path.join(os.path.dirname(mibwt.LIB_PATH), "..", "tuning_table.txt")
    if os.path.exists(default):
        mibwt.lib().mibwt_tuning_load(default.encode())
    else:
        mibwt.tuning_use(False)
