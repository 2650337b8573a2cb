"""The allocation-free entry point (mibwt_pfp_bwt_device_noalloc) against the allocating device
entry it varies (mibwt_pfp_bwt_device): byte-identical BWT and primary index on two calls in a
row, the second with a smaller text, plus the size limit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def texts():
    from mibwt import gen
    rng = np.random.default_rng(97)
    big = gen.repetitive_dna((1 << 20) // 16, 17, 0.002, 0.0002, 10, seed=97)[: 1 << 20]
    small = (rng.integers(0, 256, size=150001)).astype(np.uint8)
    return big, small


@pytest.mark.parametrize("W,P", [(5, 16), (10, 31)])
def test_noalloc_matches_device_entry(mibwt, W, P):
    big, small = texts()
    assert len(big) == mibwt.NOALLOC_MAX_LENGTH
    n0 = len(big)
    dT = mibwt.DeviceBuffer(n0)
    dO = mibwt.DeviceBuffer(n0)
    dR = mibwt.DeviceBuffer(n0)
    with mibwt.Storage(n0) as st:
        for T in (big, small):  # the second call is smaller: nothing of the first may leak into it
            n = len(T)
            dT.upload(T)
            idx_ref = st.bwt_device(dT.address, dR.address, n, "pfp", W, P)
            ref = dR.download(count=n)
            idx = mibwt.bwt_device_noalloc(dT.address, dO.address, n, W, P)
            out = dO.download(count=n)
            assert idx == idx_ref
            assert (out == ref).all()
    for b in (dT, dO, dR):
        b.free()


def test_noalloc_rejects_beyond_maximum(mibwt):
    n = mibwt.NOALLOC_MAX_LENGTH + 1
    dT = mibwt.DeviceBuffer(n)
    dO = mibwt.DeviceBuffer(n)
    with pytest.raises(mibwt.BwtError) as e:
        mibwt.bwt_device_noalloc(dT.address, dO.address, n, 5, 16)
    assert e.value.code == mibwt.BAD_PARAMETER
    with pytest.raises(mibwt.BwtError) as e:
        mibwt.bwt_device_noalloc(dT.address, dO.address, 1000, 0, 0)  # no learned choice here
    assert e.value.code == mibwt.BAD_PARAMETER
    dT.free()
    dO.free()
