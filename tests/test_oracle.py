"""Pins the CPU oracle (CPU only).

The oracle is checked against what the mathematics fixes, not against a retyped formula:
brute-force rotation sorting, the textbook mississippi BWT, inverse-BWT round trips,
an independent vectorised evaluation of the trigger hash, the exhaustively enumerated
trigger density of uniform ACGT text, the zero weight of the first window byte at P=16,
naive sorting of parse suffixes, and the Lyndon-factorisation definition.
"""
import itertools

import numpy as np
import pytest
from hypothesis import given, settings, strategies as st

import oracle


def brute_bwt(T: bytes):
    n = len(T)
    suf = sorted(range(n), key=lambda i: T[i:])
    return bytes(T[(i - 1) % n] for i in suf), suf.index(0) + 1


def test_kat_mississippi():
    # textbook BWT of mississippi$ is ipssm$pissii; the reference format replaces $ by T[n-1]
    O, idx = oracle.bwt(b"mississippi")
    assert bytes(O) == b"pssmipissii"
    assert list(O) == [112, 115, 115, 109, 105, 112, 105, 115, 115, 105, 105]
    assert idx == 5
    L = oracle.ref_to_cubwt(O, idx)
    assert bytes(L) == b"ipssmpissii"


@settings(max_examples=400, deadline=None)
@given(st.binary(min_size=1, max_size=12).map(lambda b: bytes(x % 3 for x in b))
       | st.binary(min_size=1, max_size=12)
       | st.lists(st.sampled_from([0, 1, 2, 255]), min_size=1, max_size=12).map(bytes))
def test_sais_equals_bruteforce(T):
    O, idx = oracle.bwt(T)
    Ob, ib = brute_bwt(T)
    assert bytes(O) == Ob and idx == ib


def test_all_binary_strings_up_to_10():
    for n in range(1, 11):
        for bits in itertools.product(b"ab", repeat=n):
            T = bytes(bits)
            O, idx = oracle.bwt(T)
            assert (bytes(O), idx) == brute_bwt(T)


@pytest.mark.parametrize("seed", range(5))
def test_unbwt_roundtrip(seed):
    rng = np.random.default_rng(seed)
    T = rng.integers(0, 3 + seed * 60, size=5000 + seed * 777).astype(np.uint8)
    O, idx = oracle.bwt(T)
    L = oracle.ref_to_cubwt(O, idx)
    assert (oracle.cubwt_to_ref(L, idx) == O).all()
    assert (oracle.unbwt(L, idx) == T).all()
    # the output is a permutation of the input
    assert (np.bincount(O, minlength=256) == np.bincount(T, minlength=256)).all()


def test_unbwt_rejects_bad_index():
    O, idx = oracle.bwt(b"abracadabra")
    L = oracle.ref_to_cubwt(O, idx)
    with pytest.raises(ValueError):
        oracle.unbwt(L, 0)


def _triggers_closed_form(T: np.ndarray, W: int, P: int) -> np.ndarray:
    """Window value sum_k T[l+k]*2^(W-1-k) mod P evaluated with exact integers, plus the
    equal-to-T[0..W-1] rule, for window starts l in [1, n-W]."""
    n = len(T)
    if n < W + 1:
        return np.zeros(0, dtype=np.int64)
    win = np.lib.stride_tricks.sliding_window_view(T.astype(np.int64), W)[1:]
    weights = np.array([pow(2, W - 1 - k, P) for k in range(W)], dtype=np.int64)
    h = (win * weights).sum(axis=1) % P
    eq = (win == T[:W].astype(np.int64)).all(axis=1)
    starts = np.flatnonzero((h == 0) | eq) + 1
    return starts + W - 1  # window end positions


@pytest.mark.parametrize("W,P", [(1, 2), (2, 3), (5, 16), (5, 11), (5, 13), (5, 14), (5, 23),
                                 (10, 31), (20, 41), (35, 81), (7, 1)])
def test_triggers_closed_form(W, P):
    rng = np.random.default_rng(W * 100 + P)
    for alphabet in (b"ACGT", bytes(range(256)), b"\x00\x01"):
        T = rng.choice(np.frombuffer(alphabet, dtype=np.uint8), size=20000)
        assert (oracle.pfp_triggers(T, W, P) == _triggers_closed_form(T, W, P)).all()


def test_trigger_density_acgt_w5_p16():
    # exhaustive enumeration over the 4^5 ACGT windows: 28 of 1024 have value 0 mod 16
    def h(w):
        return sum(int(c) * 2 ** (4 - k) for k, c in enumerate(w)) % 16
    zero = sum(h(w) == 0 for w in itertools.product(b"ACGT", repeat=5))
    assert zero == 28
    rng = np.random.default_rng(1)
    T = rng.choice(np.frombuffer(b"ACGT", dtype=np.uint8), size=2_000_000)
    # windows equal to T[0..4] add 1/1024 unless that window already has value 0
    expected = zero / 1024 + (0.0 if h(T[:5]) == 0 else 1 / 1024)
    t = len(oracle.pfp_triggers(T, 5, 16))
    assert abs(t / (len(T) - 5) - expected) < 0.0015


def test_p16_first_byte_has_zero_weight():
    rng = np.random.default_rng(3)
    T = rng.choice(np.frombuffer(b"ACGT", dtype=np.uint8), size=4000)
    T[:5] = np.frombuffer(b"TTTTT", dtype=np.uint8)  # a first window that no edit below recreates
    base = set(oracle.pfp_triggers(T, 5, 16).tolist())
    for e in range(200, 3900, 97):
        U = T.copy()
        U[e - 4] = ord("A") if U[e - 4] != ord("A") else ord("C")
        got = set(oracle.pfp_triggers(U, 5, 16).tolist())
        # only windows containing position e-4 at offset > 0 can change; the window ending at e
        # has it at offset 0 (weight 2^4 = 0 mod 16) and keeps its status
        assert (e in base) == (e in got) or (U[e - 4:e + 1] == T[:5]).all()


def _naive_parse(T: bytes, W: int, P: int):
    E = oracle.pfp_triggers(np.frombuffer(T, dtype=np.uint8), W, P).tolist()
    n = len(T)
    ends = E + [n + W - 1]
    ext = list(T) + [-1] * W
    phrases = []
    for k, e in enumerate(ends):
        s = 0 if k == 0 else ends[k - 1] - W + 1
        phrases.append(tuple(ext[s:e + 1]))
    return phrases


@pytest.mark.parametrize("seed", range(8))
def test_pfp_parse_against_naive(seed):
    rng = np.random.default_rng(seed)
    W, P = [(2, 3), (3, 5), (5, 16), (4, 7)][seed % 4]
    base = rng.integers(65, 69, size=300).astype(np.uint8)
    T = bytes(np.concatenate([base, base[10:], base[:150]]))
    phrases = _naive_parse(T, W, P)
    rank, psa, du = oracle.pfp_parse(np.frombuffer(T, dtype=np.uint8), W, P)
    distinct = sorted(set(phrases))
    assert du == len(distinct)
    assert rank.tolist() == [distinct.index(p) for p in phrases]
    parse = rank.tolist()
    naive_sa = sorted(range(len(parse)), key=lambda i: parse[i:])
    assert psa.tolist() == naive_sa
    # consecutive phrases overlap by exactly W symbols
    for a, b in zip(phrases, phrases[1:]):
        assert a[-W:] == b[:W]


CASES = [
    b"mississippi",
    b"a", b"ab", b"aaaaaaaaaaaaaaaaaaaaaaaaa", b"abababababababababababab",
    bytes([0, 1, 255, 0, 1, 255, 0, 0, 255]),
    b"ACGT" * 50 + b"ACGA" + b"ACGT" * 40,
]


@pytest.mark.parametrize("T", CASES)
@pytest.mark.parametrize("W,P", [(1, 2), (2, 3), (3, 4), (5, 16), (5, 11), (5, 14), (35, 81), (4, 1000)])
def test_pfp_restatement_equals_sais(T, W, P):
    O, idx, _ = oracle.pfp_bwt(T, W, P)
    Ob, ib = oracle.bwt(T)
    assert (O == Ob).all() and idx == ib


@pytest.mark.parametrize("seed", range(6))
def test_pfp_restatement_random(seed):
    rng = np.random.default_rng(100 + seed)
    for _ in range(40):
        n = int(rng.integers(1, 400))
        k = int(rng.integers(1, 6))
        T = rng.integers(0, k, size=n).astype(np.uint8) * int(rng.integers(1, 60))
        W = int(rng.integers(1, 9))
        P = int(rng.integers(1, 30))
        O, idx, _ = oracle.pfp_bwt(T, W, P)
        Ob, ib = oracle.bwt(T)
        assert (O == Ob).all() and idx == ib, (T.tolist(), W, P)


def test_pfp_restatement_repetitive_dna():
    rng = np.random.default_rng(9)
    base = rng.choice(np.frombuffer(b"ACGT", dtype=np.uint8), size=50_000)
    copies = []
    for _ in range(8):
        c = base.copy()
        hit = rng.random(len(c)) < 0.001
        c[hit] = rng.choice(np.frombuffer(b"ACGT", dtype=np.uint8), size=hit.sum())
        copies.append(c)
    T = np.concatenate(copies)
    for W, P in [(5, 16), (35, 81), (10, 23)]:
        O, idx, info = oracle.pfp_bwt(T, W, P)
        Ob, ib = oracle.bwt(T)
        assert (O == Ob).all() and idx == ib
        assert info["distinct"] < info["d"]  # repetition shows up as shared phrases


def _is_lyndon(w: bytes) -> bool:
    return all(w < w[i:] + w[:i] for i in range(1, len(w))) if len(w) > 1 else True


@settings(max_examples=300, deadline=None)
@given(st.binary(min_size=1, max_size=40).map(lambda b: bytes(x % 4 for x in b)))
def test_lyndon_factors_definition(T):
    F = oracle.lyndon_factors(T).tolist() + [len(T)]
    words = [T[F[i]:F[i + 1]] for i in range(len(F) - 1)]
    assert b"".join(words) == T
    assert all(_is_lyndon(w) for w in words)
    assert all(words[i] >= words[i + 1] for i in range(len(words) - 1))


def _brute_lyndon_array(T: bytes):
    n = len(T)
    out = []
    for i in range(n):
        out.append(max(l for l in range(1, n - i + 1) if _is_lyndon(T[i:i + l])))
    return out


@settings(max_examples=300, deadline=None)
@given(st.binary(min_size=1, max_size=16).map(lambda b: bytes(x % 3 for x in b)))
def test_lyndon_array_definition(T):
    assert oracle.lyndon_array(T).tolist() == _brute_lyndon_array(T)


def test_lyndon_array_next_smaller_suffix():
    # lam[i] = NSS[i] - i: the first later suffix that is smaller
    rng = np.random.default_rng(12)
    for _ in range(30):
        T = bytes(rng.integers(0, 3, size=int(rng.integers(1, 60))).astype(np.uint8))
        lam = oracle.lyndon_array(T)
        n = len(T)
        for i in range(n):
            nss = next((j for j in range(i + 1, n) if T[j:] < T[i:]), n)
            assert lam[i] == nss - i


def test_lyndon_chain_is_the_factorisation():
    rng = np.random.default_rng(13)
    for _ in range(50):
        T = rng.integers(0, 4, size=int(rng.integers(1, 500))).astype(np.uint8)
        lam = oracle.lyndon_array(T)
        chain, i = [], 0
        while i < len(T):
            chain.append(i)
            i += lam[i]
        assert chain == oracle.lyndon_factors(T).tolist()


def test_lyndon_grouping_principle():
    # suffix order = (longest Lyndon prefix, then the suffix after it)
    rng = np.random.default_rng(14)
    for _ in range(40):
        T = bytes(rng.integers(0, 3, size=int(rng.integers(2, 40))).astype(np.uint8))
        lam = oracle.lyndon_array(T)
        key = {i: (T[i:i + lam[i]],) for i in range(len(T))}
        sa = sorted(range(len(T)), key=lambda i: T[i:])
        for a, b in zip(sa, sa[1:]):
            pa, pb = key[a][0], key[b][0]
            assert pa <= pb  # groups by Lyndon prefix appear in lexicographic order
            if pa == pb:
                assert T[a + lam[a]:] < T[b + lam[b]:]


@settings(max_examples=300, deadline=None)
@given(st.binary(min_size=1, max_size=14).map(lambda b: bytes(x % 4 for x in b))
       | st.lists(st.sampled_from([0, 1, 255]), min_size=1, max_size=14).map(bytes))
def test_lyndon_bwt_restatement_small(T):
    O, idx = oracle.lyndon_bwt(T)
    assert (bytes(O), idx) == brute_bwt(T)


def test_lyndon_bwt_restatement_repetitive():
    rng = np.random.default_rng(15)
    base = rng.choice(np.frombuffer(b"ACGT", dtype=np.uint8), size=30000)
    T = np.concatenate([base, base[11:], base[:9000], base])
    O, idx = oracle.lyndon_bwt(T)
    Ob, ib = oracle.bwt(T)
    assert (O == Ob).all() and idx == ib
    r, g = oracle.lyndon_ranks(T)
    assert 0 < g < len(T)


def _duval(w: bytes):
    """Lyndon factorization of w by Duval's algorithm (starts of the factors)."""
    out, i, n = [], 0, len(w)
    while i < n:
        j, k = i + 1, i
        while j < n and w[k] <= w[j]:
            k = i if w[k] < w[j] else k + 1
            j += 1
        while i <= k:
            out.append(i)
            i += j - k
    return out


def _grammar_sizes(T: bytes):
    """Sizes of the label grammar from the definitions alone: the distinct longest Lyndon
    prefixes (brute force), their children (the Lyndon factorization of the label without its
    first symbol, by Duval) and the roots (the Lyndon factorization of T)."""
    lam = _brute_lyndon_array(T)
    labels = {T[i:i + lam[i]] for i in range(len(T))}
    slots = sum(len(_duval(M[1:])) for M in labels)
    return len(labels), slots, len(_duval(T))


@settings(max_examples=300, deadline=None)
@given(st.binary(min_size=1, max_size=14).map(lambda b: bytes(x % 4 for x in b))
       | st.lists(st.sampled_from([0, 1, 255]), min_size=1, max_size=14).map(bytes))
def test_lyndon_grammar_restatement_small(T):
    O, idx, info = oracle.lyndon_grammar_bwt(T)
    assert (bytes(O), idx) == brute_bwt(T)
    assert (info["labels"], info["slots"], info["roots"]) == _grammar_sizes(T)
    # every position is one child slot of one label occurrence, or a root
    assert info["items"] >= info["slots"] + info["roots"]


def test_lyndon_grammar_restatement_shapes():
    # repetitive DNA (few labels, many rounds), runs of one symbol before a smaller one (long
    # runs of equal children, settled in one round by their run length), descending bytes (many
    # roots), periodic text, all byte values
    rng = np.random.default_rng(16)
    base = rng.choice(np.frombuffer(b"ACGT", dtype=np.uint8), size=20000)
    dna = np.concatenate([base, base[11:], base[:9000], base])
    nrun = np.concatenate([base[:5000], np.full(20000, ord("N"), np.uint8), base[5000:]])
    cases = {"dna": dna, "nrun": nrun, "descending": np.arange(255, -1, -1, dtype=np.uint8).repeat(3),
             "period": np.tile(np.frombuffer(b"ACGTA", dtype=np.uint8), 3000),
             "bytes": rng.integers(0, 256, size=5000).astype(np.uint8)}
    for name, T in cases.items():
        O, idx, info = oracle.lyndon_grammar_bwt(T)
        Ob, ib = oracle.bwt(T)
        assert (O == Ob).all() and idx == ib, name
    _, _, info = oracle.lyndon_grammar_bwt(dna)
    assert info["labels"] < len(dna) // 10 and info["items"] < len(dna)  # the grammar compresses
    _, _, info = oracle.lyndon_grammar_bwt(nrun)
    assert info["rounds"] < 100  # a run of 20000 equal children does not take a round per child


@pytest.mark.parametrize("threads", [1, 2, 5, 16])
def test_unbwt_mt_matches_sequential_inverse(threads):
    # the multi-threaded inverse (segments of the LF cycle walked in parallel) gives the text
    # back exactly like the sequential LF walk, and rejects the same non-BWT inputs
    rng = np.random.default_rng(threads)
    for n in [1, 2, 17, 1000, 65537, 300000]:
        T = (rng.integers(0, 3, size=n) * 40).astype(np.uint8)
        O, idx = oracle.bwt(T)
        L = oracle.ref_to_cubwt(O, idx)
        assert (oracle.unbwt_mt(L, idx, threads) == T).all()
        assert (oracle.unbwt(L, idx) == T).all()
        if n > 2:
            bad = idx % n + 1  # a wrong primary index is not a BWT of this text
            for f in (oracle.unbwt, lambda a, b: oracle.unbwt_mt(a, b, threads)):
                try:
                    U = f(L, bad)
                except ValueError:
                    continue
                assert not (U == T).all() or bad == idx


@pytest.mark.parametrize("W,P", [(5, 16), (1, 2), (3, 4), (10, 31), (35, 81)])
def test_cpu_pfp_baseline_matches_sais(W, P):
    # the multi-threaded C++ restatement (the CPU baseline) gives the SA-IS BWT exactly, for any
    # number of threads, also with bytes 0/1/255 and texts shorter than W
    rng = np.random.default_rng(W * 100 + P)
    for it in range(60):
        n = int(rng.integers(1, 4000))
        k = int(rng.integers(1, 5))
        T = (rng.integers(0, k, size=n) * int(rng.integers(1, 86))).astype(np.uint8)
        if it % 5 == 0:
            T = np.concatenate([T, T, T[: n // 2]])
        O, idx = oracle.bwt(T)
        O2, idx2, d, du = oracle.cpu_pfp_bwt(T, W, P, threads=1 + it % 5)
        assert idx2 == idx and (O2 == O).all(), (n, W, P)
        assert d == len(oracle.pfp_triggers(T, W, P)) + 1
