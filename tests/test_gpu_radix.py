"""The LSD radix sort (upsweep, scan of the tile counts, downsweep) and the timing hooks.

The sort is exercised through `mibwt_time_kernel("radix")`, which sorts uniformly random
(64-bit key, 32-bit value) pairs and checks on the device that the keys come out sorted and
that every value still sits beside its own key; it returns an error otherwise.  Every BWT test
also runs the sort inside the pipelines.
"""
import numpy as np
import pytest

from tests.conftest import load_mibwt


@pytest.mark.gpu
@pytest.mark.parametrize("items", [1, 2, 255, 4095, 4096, 4097, 65536 + 17, 1 << 20, 3_000_001])
@pytest.mark.parametrize("shape", ["8,16,64,64", "8,8,64,64"])
def test_radix_sort_sorts_random_pairs(monkeypatch, items, shape):
    # both tile shapes (16 and 8 items per thread), ragged last tiles, one-tile and many-tile sorts
    mibwt = load_mibwt()
    monkeypatch.setenv("MIBWT_TUNE_FORCE", shape)
    st = mibwt.Storage(1 << 16)
    ms, b = st.time_kernel("radix", 1, float(items))
    assert ms > 0
    assert b == 24.0 * items
    st.close()


@pytest.mark.gpu
def test_radix_timing_rejects_bad_counts():
    mibwt = load_mibwt()
    st = mibwt.Storage(1 << 16)
    for bad in (0.0, -5.0, float(1 << 31)):
        with pytest.raises(mibwt.BwtError):
            st.time_kernel("radix", 1, bad)
    st.close()


@pytest.mark.gpu
def test_scan_timing_counts_the_triggers_and_the_sentinel():
    mibwt = load_mibwt()
    from mibwt import gen
    T = gen.repetitive_dna(1 << 18, 4, 1e-3, 1e-4, 10, seed=7)
    st = mibwt.Storage(len(T))
    st.bwt(T, "pfp", 5, 16)
    t = st.stats()["triggers"]
    ms, b = st.time_kernel("scan", 2)
    assert ms > 0
    assert b == len(T) + 4.0 * (t + 1)
    st.close()
