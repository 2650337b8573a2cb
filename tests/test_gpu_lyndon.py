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


@pytest.mark.parametrize("bs", ["16", "64", "256"])
def test_lyndon_array_block_sizes(storage, monkeypatch, bs):
    # the local pass over LDS at other block sizes (the merge tree then starts from smaller or
    # larger blocks; 16 makes every workgroup span several blocks of a short text)
    monkeypatch.setenv("MIBWT_LYN_BLOCK", bs)
    for name in ("period5", "repetitive", "long_period", "unary", "descending", "rand3", "rand7"):
        T = TEXTS[name]
        assert (storage.lyndon_array(T).astype(np.int64) == oracle.lyndon_array(T)).all(), name


INDUCTIONS = {
    # grammar induction over distinct labels (the default)
    "grammar": ({}, 1),
    # the position-level induction, asked for directly
    "positions": ({"MIBWT_LYN_INDUCTION": "positions"}, 2),
    # the grammar induction over its item budget from the first round: the position-level one
    # takes over in the same build
    "fallback": ({"MIBWT_LYN_ITEM_BUDGET": "0"}, 2),
}


@pytest.mark.parametrize("mode", sorted(INDUCTIONS))
@pytest.mark.parametrize("name", sorted(TEXTS))
def test_lyndon_bwt(storage, mibwt, monkeypatch, name, mode):
    env, expect = INDUCTIONS[mode]
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    T = np.frombuffer(TEXTS[name], dtype=np.uint8)
    L, idx = storage.bwt(T, "lyndon")
    assert storage.stats()["induction"] == expect
    O, oidx = oracle.bwt(T)
    assert idx == oidx
    assert (mibwt.cubwt_to_reference(L, idx) == O).all()


def _grammar_stats(st):
    s = st.stats()
    return s["induction"], s["retries"], dict(
        labels=s["distinct_phrases"], slots=s["dict_length"], roots=s["triggers"], ups=s["mixed_entries"],
        rounds=s["doubling_rounds_parse"], items=s["phrases"])


@pytest.mark.parametrize("name", sorted(TEXTS))
def test_grammar_sizes_match_oracle(storage, name):
    # the grammar phases against the oracle's restatement with exact string comparisons:
    # distinct labels, child slots, roots, up-context pairs, induction rounds and items
    T = np.frombuffer(TEXTS[name], dtype=np.uint8)
    storage.bwt(T, "lyndon")
    induction, retries, got = _grammar_stats(storage)
    _, _, info = oracle.lyndon_grammar_bwt(T)
    assert induction == 1 and retries == 0
    assert got == info


def test_grammar_sizes_match_oracle_dna(mibwt):
    from mibwt import gen
    T = gen.workload("dna256m", scale=1 / 64)  # 4 MiB: 64 KiB base x 64 copies
    with mibwt.Storage(len(T)) as st:
        L, idx = st.bwt(T, "lyndon")
        induction, retries, got = _grammar_stats(st)
    O, oidx, info = oracle.lyndon_grammar_bwt(T)
    assert idx == oidx and (mibwt.cubwt_to_reference(L, idx) == O).all()
    assert induction == 1 and retries == 0 and got == info
    assert info["items"] < len(T) // 4  # one item per distinct label slot, not per position


def test_lyndon_dna32m(mibwt):
    from mibwt import gen
    T = gen.workload("dna32m")
    O, oidx = oracle.bwt(T)
    with mibwt.Storage(len(T)) as st:
        lam = st.lyndon_array(T[: 1 << 22])
        assert (lam.astype(np.int64) == oracle.lyndon_array(T[: 1 << 22])).all()
        L, idx = st.bwt(T, "lyndon")
    assert idx == oidx and (mibwt.cubwt_to_reference(L, idx) == O).all()


@pytest.mark.parametrize("bits", [1, 2])
def test_fingerprint_collisions_are_caught(mibwt, monkeypatch, bits):
    # narrow first-attempt fingerprints make the suffix comparisons of the Lyndon array and the
    # label order collide: the inversion check of the finished BWT must catch it and rebuild
    # with full fingerprints, giving the exact BWT (and reporting the retry).  Fingerprints are
    # only consulted past 32 equal bytes, so every text repeats long stretches; with 1 or 2
    # bits nearly every such comparison is decided by a collision.
    monkeypatch.setenv("MIBWT_LYN_HASH_BITS", str(bits))
    rng = np.random.default_rng(bits)
    retried = 0
    with mibwt.Storage(1 << 16) as st:
        for k in range(6):
            base = rng.choice(np.frombuffer(b"ACGT", dtype=np.uint8), size=2000)
            T = np.concatenate([base, base[7:], base[:900], base]) if k % 2 else np.tile(base[:300], 10)
            L, idx = st.bwt(T, "lyndon")
            O, oidx = oracle.bwt(T)
            assert idx == oidx and (mibwt.cubwt_to_reference(L, idx) == O).all()
            retried += st.stats()["retries"] > 0
    assert retried > 0


@pytest.mark.timeout(120)
@pytest.mark.parametrize("kind", ["a_run_b", "dna_with_n_run"])
def test_long_runs_are_fast(mibwt, kind):
    # a long run of one symbol followed by a larger one gives Lyndon labels a^j b that all share
    # long prefixes: the label order must not take a round per few symbols
    import time
    rng = np.random.default_rng(5)
    if kind == "a_run_b":
        T = np.frombuffer(b"a" * 300000 + b"b", dtype=np.uint8)
    else:
        dna = rng.choice(np.frombuffer(b"ACGT", dtype=np.uint8), size=400000)
        T = np.concatenate([dna[:200000], np.full(1 << 20, ord("N"), np.uint8), dna[200000:]])
    with mibwt.Storage(len(T)) as st:
        t0 = time.perf_counter()
        L, idx = st.bwt(T, "lyndon")
        dt = time.perf_counter() - t0
        induction, _, got = _grammar_stats(st)
    O, oidx = oracle.bwt(T)
    assert idx == oidx and (mibwt.cubwt_to_reference(L, idx) == O).all()
    assert dt < 10.0, dt
    # the run of equal children is one key (label, run length): no round per symbol
    assert induction == 1 and got["rounds"] < 100, got


@pytest.mark.timeout(600)
def test_lyndon_dna256m_matches_pfp_and_inverts(mibwt):
    # the 256 MiB collection: the Lyndon-grammar BWT is byte for byte the PFP one, and the host
    # oracle's inverse gives the text back
    from mibwt import gen
    T = gen.workload("dna256m")
    n = len(T)
    dT = mibwt.DeviceBuffer.from_numpy(T)
    dO = mibwt.DeviceBuffer(n)
    with mibwt.Storage(n) as st:
        idx_l = st.bwt_device(dT.address, dO.address, n, "lyndon")
        O_l = dO.download()
        idx_p = st.bwt_device(dT.address, dO.address, n, "pfp", 5, 16)
        O_p = dO.download()
    dT.free()
    dO.free()
    assert idx_l == idx_p and (O_l == O_p).all()
    assert (oracle.unbwt_mt(oracle.ref_to_cubwt(O_l, idx_l), idx_l, 16) == T).all()
