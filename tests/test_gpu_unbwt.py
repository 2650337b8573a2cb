"""GPU inverse BWT (libcubwt_unbwt) and sampled indexes (libcubwt_bwt_aux) against the oracle.

Inputs are oracle BWTs (SA-IS) of random, periodic and unary texts; the round trip of the GPU
BWT of the 32 MiB workload; a wrong primary index must be reported, not answered.  The sampled
indexes are checked against the suffix array: I[j] = 1 + rank of the suffix at j*r among all
suffixes with the terminator, the convention of the primary index (I[0]).
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def texts():
    rng = np.random.default_rng(23)
    out = [b"x", b"ab", b"mississippi", b"a" * 1000, b"ab" * 777, bytes(range(256)) * 3]
    for _ in range(30):
        n = int(rng.integers(2, 20000))
        k = int(rng.integers(1, 5))
        out.append((rng.integers(0, k, size=n) * 61).astype(np.uint8).tobytes())
    return out


def test_unbwt_matches_text(storage):
    for T in texts():
        O, idx = oracle.bwt(T)
        L = oracle.ref_to_cubwt(O, idx)
        U = storage.unbwt(L, idx)
        assert bytes(U) == bytes(T), len(T)


def test_unbwt_rejects_wrong_index(storage, mibwt):
    T = np.frombuffer(b"mississippimississippi", dtype=np.uint8)
    O, idx = oracle.bwt(T)
    L = oracle.ref_to_cubwt(O, idx)
    for bad in (1, 2, 3):  # not the primary index: the LF cycle from row 0 is shorter than n
        if bad == idx:
            continue
        try:
            U = storage.unbwt(L, bad)
        except mibwt.BwtError as e:
            assert e.code == mibwt.BAD_PARAMETER
        else:
            assert bytes(U) != bytes(T)


def test_unbwt_of_gpu_bwt_32mib(mibwt):
    from mibwt import gen
    T = gen.workload("dna32m")
    dT = mibwt.DeviceBuffer.from_numpy(T)
    dO = mibwt.DeviceBuffer(len(T))
    dU = mibwt.DeviceBuffer(len(T))
    with mibwt.Storage(len(T)) as st:
        idx = st.bwt_device(dT.address, dO.address, len(T), "pfp", 5, 16)
        st.unbwt_device(dO.address, dU.address, len(T), idx)
    assert (dU.download() == T).all()
    for b in (dT, dO, dU):
        b.free()


def test_bwt_aux_rejects_rates_libcubwt_rejects(storage, mibwt):
    # libcubwt.cu:2437: the sampling rate is a power of two >= 4
    T = np.frombuffer(b"ACGT" * 100, dtype=np.uint8)
    for r in (0, 1, 2, 3, 6, 12):
        with pytest.raises(mibwt.BwtError) as e:
            storage.bwt_aux(T, r)
        assert e.value.code == mibwt.BAD_PARAMETER


@pytest.mark.parametrize("r", [4, 8, 64])
def test_bwt_aux_indexes(storage, r):
    rng = np.random.default_rng(r)
    for n in (16, 17, 1000, 5003):
        T = rng.choice(np.frombuffer(b"ACGT", dtype=np.uint8), size=n)
        L, I = storage.bwt_aux(T, r)
        O, idx = oracle.bwt(T)
        assert (L == oracle.ref_to_cubwt(O, idx)).all() and I[0] == idx
        SA = oracle.suffix_array(T)
        rank = np.empty(n, dtype=np.int64)
        rank[SA] = np.arange(n)
        assert (I == rank[np.arange(0, n, r)] + 1).all()
