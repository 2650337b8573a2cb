"""GPU parity of the prefix-free-parsing BWT against the CPU oracle (bit-exact).

Small and edge-case inputs are compared with the SA-IS oracle directly; the 32 MiB
repetitive-DNA workload is compared in full; the 256 MiB workload is checked through
size-independent properties (inverse-BWT round trip, symbol histogram).
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def check(storage, mibwt, T, W=5, P=16):
    T = np.frombuffer(bytes(T), dtype=np.uint8) if isinstance(T, (bytes, bytearray)) else T
    L, idx = storage.bwt(T, "pfp", W, P)
    O, oidx = oracle.bwt(T)
    assert idx == oidx, (len(T), W, P)
    assert (mibwt.cubwt_to_reference(L, idx) == O).all(), (len(T), W, P)


def test_kat_mississippi(storage, mibwt):
    L, idx = storage.bwt(b"mississippi", "pfp", 5, 16)
    assert idx == 5 and bytes(L) == b"ipssmpissii"
    assert bytes(mibwt.cubwt_to_reference(L, idx)) == b"pssmipissii"


def test_libcubwt_entry_min_length(storage, mibwt):
    with pytest.raises(mibwt.BwtError) as e:
        storage.bwt(b"mississippi", "cubwt")
    assert e.value.code == mibwt.BAD_PARAMETER  # libcubwt rejects n < 16 (libcubwt.cu:2418)
    T = b"mississippimississippi"
    L, idx = storage.bwt(T, "cubwt")
    O, oidx = oracle.bwt(T)
    assert idx == oidx and (L == oracle.ref_to_cubwt(O, oidx)).all()


def test_n_above_max_length(storage, mibwt):
    with pytest.raises(mibwt.BwtError) as e:
        storage.bwt(np.zeros(storage.max_length + 1, dtype=np.uint8), "pfp")
    assert e.value.code == mibwt.BAD_PARAMETER


@pytest.mark.parametrize("W,P", [(1, 2), (2, 3), (3, 4), (5, 16), (5, 11), (5, 13), (5, 14),
                                 (5, 23), (10, 31), (35, 81), (64, 997), (4, 1), (6, 65535)])
def test_random_small(storage, mibwt, W, P):
    rng = np.random.default_rng(W * 1000 + P)
    for _ in range(25):
        n = int(rng.integers(1, 3000))
        k = int(rng.integers(1, 6))
        T = (rng.integers(0, k, size=n) * int(rng.integers(1, 64))).astype(np.uint8)
        check(storage, mibwt, T, W, P)


EDGE = {
    "one": b"x",
    "two": b"ba",
    "unary": b"a" * 5000,
    "period2": b"ab" * 3000,
    "period7": b"abcabcd" * 700,
    "all_bytes": bytes(range(256)) * 20,
    "zeros_ones": bytes([0, 1]) * 1000 + bytes([255]) * 10,
    "first_window_repeats": b"ACGTA" * 1000,
    "descending": bytes(range(255, -1, -1)) * 10,
}


@pytest.mark.parametrize("name", sorted(EDGE))
@pytest.mark.parametrize("W,P", [(5, 16), (1, 2), (35, 81)])
def test_edge_cases(storage, mibwt, name, W, P):
    check(storage, mibwt, EDGE[name], W, P)


def test_long_trigger_free_run(storage, mibwt):
    # phrases far longer than the reference's 8192-byte limit (PFPk:581-585)
    rng = np.random.default_rng(4)
    T = np.concatenate([rng.integers(65, 69, size=500), np.full(40000, 67), rng.integers(65, 69, size=500)])
    check(storage, mibwt, T.astype(np.uint8), 5, 16)
    check(storage, mibwt, T.astype(np.uint8), 3, 60000)  # almost no triggers at all


def test_in_place_libcubwt(storage, mibwt):
    rng = np.random.default_rng(11)
    T = rng.choice(np.frombuffer(b"ACGT", dtype=np.uint8), size=100000)
    O, oidx = oracle.bwt(T)
    buf = T.copy()
    import ctypes
    p = buf.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
    idx = mibwt.lib().libcubwt_bwt(storage.handle, p, p, len(buf))  # L aliases T
    assert idx == oidx and (buf == oracle.ref_to_cubwt(O, oidx)).all()


def test_repetitive_collection_various_wp(mibwt):
    from mibwt import gen
    T = gen.repetitive_dna(1 << 16, 16, 0.002, 0.0002, 10, seed=5)
    O, oidx = oracle.bwt(T)
    with mibwt.Storage(len(T)) as st:
        for W, P in [(5, 16), (5, 11), (10, 23), (20, 41), (35, 81)]:
            L, idx = st.bwt(T, "pfp", W, P)
            assert idx == oidx and (mibwt.cubwt_to_reference(L, idx) == O).all(), (W, P)
            s = st.stats()
            assert s["phrases"] == len(oracle.pfp_triggers(T, W, P)) + 1


def test_dna32m_full(mibwt):
    from mibwt import gen
    T = gen.workload("dna32m")
    O, oidx = oracle.bwt(T)
    with mibwt.Storage(len(T)) as st:
        L, idx = st.bwt(T, "pfp", 5, 16)
        s = st.stats()
    assert idx == oidx
    assert (mibwt.cubwt_to_reference(L, idx) == O).all()
    assert s["triggers"] == len(oracle.pfp_triggers(T, 5, 16))


def test_device_entry_reference_format(mibwt):
    rng = np.random.default_rng(3)
    T = rng.choice(np.frombuffer(b"ACGT", dtype=np.uint8), size=1 << 18)
    T = np.concatenate([T, T[1000:], T[:77777]])
    dT = mibwt.DeviceBuffer.from_numpy(T)
    dO = mibwt.DeviceBuffer(len(T))
    with mibwt.Storage(len(T)) as st:
        idx = st.bwt_device(dT.address, dO.address, len(T), "pfp", 5, 16)
    O, oidx = oracle.bwt(T)
    assert idx == oidx and (dO.download() == O).all()
    dT.free()
    dO.free()


def test_dna256m_roundtrip(mibwt):
    from mibwt import gen
    T = gen.workload("dna256m")
    with mibwt.Storage(len(T)) as st:
        L, idx = st.bwt(T, "pfp", 5, 16)
    assert (np.bincount(L, minlength=256) == np.bincount(T, minlength=256)).all()
    assert (oracle.unbwt(L, idx) == T).all()


def test_device_entry_learned_choice(mibwt):
    # W = P = 0: the selector picks (W, P) on the resident text; the BWT is the same
    from mibwt import gen
    T = gen.repetitive_dna(1 << 17, 12, 0.001, 0.0001, 10, seed=13)
    dT = mibwt.DeviceBuffer.from_numpy(T)
    dO = mibwt.DeviceBuffer(len(T))
    with mibwt.Storage(len(T)) as st:
        idx = st.bwt_device(dT.address, dO.address, len(T), "pfp", 0, 0)
        s = st.stats()
    O, oidx = oracle.bwt(T)
    assert idx == oidx and (dO.download() == O).all()
    assert (s["W"], s["P"]) in mibwt.wp_grid()
    dT.free()
    dO.free()
    # the host entry with W = P = 0 makes the same choice and the same BWT
    with mibwt.Storage(len(T)) as st:
        L, hidx = st.bwt(T, "pfp", 0, 0)
        hs = st.stats()
    assert hidx == oidx and (mibwt.cubwt_to_reference(L, hidx) == O).all()
    assert (hs["W"], hs["P"]) == (s["W"], s["P"])


def test_narrow_fingerprint_sort_widens_on_collision(mibwt, monkeypatch):
    # an 8-bit first sort makes different phrases share sort keys: the dictionary must detect
    # it and repeat the sort on all 64 bits, with the same exact result
    rng = np.random.default_rng(17)
    T = rng.integers(0, 256, size=200000).astype(np.uint8)
    monkeypatch.setenv("MIBWT_DICT_KEY_BITS", "8")
    with mibwt.Storage(len(T)) as st:
        L, idx = st.bwt(T, "pfp", 5, 11)
    O, oidx = oracle.bwt(T)
    assert idx == oidx and (mibwt.cubwt_to_reference(L, idx) == O).all()


def test_colliding_fingerprints_reseed(mibwt, monkeypatch):
    # 12-bit first-attempt fingerprints make different phrases collide: the byte comparison
    # of every occurrence with its representative must catch it and re-seed, same exact result
    rng = np.random.default_rng(23)
    T = rng.integers(0, 256, size=200000).astype(np.uint8)
    monkeypatch.setenv("MIBWT_DICT_HASH_BITS", "12")
    with mibwt.Storage(len(T)) as st:
        L, idx = st.bwt(T, "pfp", 5, 11)
    O, oidx = oracle.bwt(T)
    assert idx == oidx and (mibwt.cubwt_to_reference(L, idx) == O).all()


@pytest.mark.parametrize("kind", ["random", "repetitive"])
def test_wide_entries(mibwt, monkeypatch, kind):
    # 64-bit BWT entries (taken when the entry groups exceed 24 bits) carry the primary index
    # through the sort instead of counting it: forced on small texts, same exact result
    rng = np.random.default_rng(29)
    if kind == "random":
        T = rng.integers(0, 256, size=300000).astype(np.uint8)
    else:
        base = rng.integers(0, 4, size=20000).astype(np.uint8)
        T = np.frombuffer(b"ACGT", dtype=np.uint8)[np.tile(base, 12)].copy()
        T[rng.integers(0, len(T), size=200)] = ord("T")
    monkeypatch.setenv("MIBWT_WIDE_ENTRIES", "1")
    with mibwt.Storage(len(T)) as st:
        L, idx = st.bwt(T, "pfp", 5, 11)
    O, oidx = oracle.bwt(T)
    assert idx == oidx and (mibwt.cubwt_to_reference(L, idx) == O).all()


@pytest.mark.parametrize("kind", ["random", "repetitive"])
@pytest.mark.parametrize("payload", ["0", "1", "2", "-5", "-16", None])
def test_dictionary_record_payload(mibwt, monkeypatch, kind, payload):
    # the dictionary positions' records ride above the initial key bits through the sort; with
    # 1 or 2 frequency bits most frequencies do not fit and those positions read their record
    # instead ("0": never carried; "-k": the rest lengths in k bits instead of the records, too
    # short for most suffixes at 5), same exact result on distinct and on repeated phrases
    rng = np.random.default_rng(31)
    if kind == "random":
        T = rng.integers(0, 256, size=250000).astype(np.uint8)
    else:
        base = rng.integers(0, 4, size=15000).astype(np.uint8)
        T = np.frombuffer(b"ACGT", dtype=np.uint8)[np.tile(base, 16)].copy()
        T[rng.integers(0, len(T), size=150)] = ord("G")
    if payload is not None:
        monkeypatch.setenv("MIBWT_DICT_PAYLOAD", payload)
    O, oidx = oracle.bwt(T)
    with mibwt.Storage(len(T)) as st:
        for W, P in [(5, 16), (4, 7)]:
            L, idx = st.bwt(T, "pfp", W, P)
            assert idx == oidx and (mibwt.cubwt_to_reference(L, idx) == O).all(), (W, P)


_LARGE_BWT = {}


@pytest.mark.timeout(600)
@pytest.mark.parametrize("name,W,P", [("dna1g", 5, 16), ("dna1g", 35, 81), ("markov256m", 5, 16),
                                      ("markov256m", 35, 81)])
def test_large_workloads_roundtrip(mibwt, name, W, P):
    # the 1 GiB collection and its 256 MiB low-repetition control: inverse BWT by the host
    # oracle (independent of the library), symbol histogram, and the same BWT for both (W,P)
    from mibwt import gen
    T = gen.workload(name)
    n = len(T)
    dT = mibwt.DeviceBuffer.from_numpy(T)
    dO = mibwt.DeviceBuffer(n)
    with mibwt.Storage(n) as st:
        idx = st.bwt_device(dT.address, dO.address, n, "pfp", W, P)
    dT.free()
    O = dO.download()
    dO.free()
    assert (np.bincount(O, minlength=256) == np.bincount(T, minlength=256)).all()
    prev = _LARGE_BWT.get(name)
    if prev is not None:
        # the other (W,P) of this text was already inverted: its BWT must be the same bytes
        assert prev[0] == idx and (prev[1] == O).all()
        return
    assert (oracle.unbwt_mt(oracle.ref_to_cubwt(O, idx), idx, 16) == T).all()
    _LARGE_BWT.clear()
    _LARGE_BWT[name] = (idx, O)


@pytest.mark.parametrize("W,P", [(5, 16), (1, 2), (10, 31), (35, 81)])
def test_intermediate_phases_match_oracle(storage, mibwt, W, P):
    # phrase ranks (the lexicographic phrase order) and the parse suffix array against the
    # oracle's restatement (PFPk:1990-2164, 616-691, 2342-2412), not only the final BWT
    from mibwt import gen
    rng = np.random.default_rng(W * 7 + P)
    texts = [gen.repetitive_dna(4096, 12, 0.003, 0.0003, 10, seed=W + P),
             rng.integers(0, 4, size=30000).astype(np.uint8) * 60,
             np.frombuffer(b"ACGTA" * 2000, dtype=np.uint8)]
    for T in texts:
        rank, psa, du = storage.pfp_parse(T, W, P)
        orank, opsa, odu = oracle.pfp_parse(T, W, P)
        assert du == odu
        assert (rank.astype(np.int64) == orank).all()
        assert (psa.astype(np.int64) == opsa).all()


def test_device_entry_on_caller_stream(mibwt):
    # a stream created through the library: the build is ordered after the caller's upload on
    # that stream (mibwt_copy_async) and the result is the oracle's
    rng = np.random.default_rng(41)
    T = rng.choice(np.frombuffer(b"ACGT", dtype=np.uint8), size=1 << 17)
    T = np.concatenate([T, T[500:], T[:3000]])
    s = mibwt.Stream()
    dT = mibwt.DeviceBuffer(len(T))
    dO = mibwt.DeviceBuffer(len(T))
    rc = mibwt.lib().mibwt_copy_async(dT.ptr, T.ctypes.data_as(__import__("ctypes").c_void_p), len(T), 0,
                                      s.handle)
    assert rc == mibwt.NO_ERROR
    with mibwt.Storage(len(T)) as st:
        idx = st.bwt_device(dT.address, dO.address, len(T), "pfp", 5, 16, stream=s.address)
        # the caller's stream after a failing call (bad W) stays usable
        with pytest.raises(mibwt.BwtError):
            st.bwt_device(dT.address, dO.address, len(T), "pfp", 99, 16, stream=s.address)
        idx2 = st.bwt_device(dT.address, dO.address, len(T), "pfp", 5, 16, stream=s.address)
    O, oidx = oracle.bwt(T)
    assert idx == oidx == idx2 and (dO.download() == O).all()
    s.close()
    dT.free()
    dO.free()

