"""Seeded synthetic texts for the BWT-build workloads (numpy, deterministic).

Repetitive DNA collections: a uniform A/C/G/T base replicated `copies` times, each copy with
independent substitutions (to a different base) and insertions/deletions (50/50, length
uniform in [1, indel_max]).  Low-repetition control: an order-2 Markov chain over the 96
printable bytes 32..127 with Dirichlet(0.5) transition rows.
"""
from __future__ import annotations

import numpy as np

ACGT = np.frombuffer(b"ACGT", dtype=np.uint8)
SEED_BASE = 0xB17C0DE


def mutate(rng: np.random.Generator, base_idx: np.ndarray, sub_rate: float, indel_rate: float,
           indel_max: int, sigma: int = 4) -> np.ndarray:
    """One mutated copy of a base given as symbol indices 0..sigma-1; returns indices."""
    x = base_idx.copy()
    if sub_rate > 0:
        hit = np.flatnonzero(rng.random(len(x)) < sub_rate)
        x[hit] = (x[hit].astype(np.int64) + rng.integers(1, sigma, size=len(hit))) % sigma
    if indel_rate <= 0:
        return x
    nev = rng.binomial(len(x), indel_rate)
    if nev == 0:
        return x
    pos = np.sort(rng.integers(0, len(x), size=nev))
    kind = rng.random(nev) < 0.5
    lens = rng.integers(1, indel_max + 1, size=nev)
    pieces = []
    cur = 0
    for p, ins, ln in zip(pos.tolist(), kind.tolist(), lens.tolist()):
        if p < cur:
            continue
        pieces.append(x[cur:p])
        if ins:
            pieces.append(rng.integers(0, sigma, size=ln).astype(np.uint8))
            cur = p
        else:
            cur = min(len(x), p + ln)
    pieces.append(x[cur:])
    return np.concatenate(pieces)


def repetitive_dna(base_len: int, copies: int, sub_rate: float, indel_rate: float,
                   indel_max: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    base = rng.integers(0, 4, size=base_len).astype(np.uint8)
    parts = [mutate(rng, base, sub_rate, indel_rate, indel_max) for _ in range(copies)]
    return ACGT[np.concatenate(parts)]


def repetitive_bytes(base_len: int, copies: int, sub_rate: float, indel_rate: float,
                     indel_max: int, seed: int) -> np.ndarray:
    """Like repetitive_dna over all 256 byte values: a uniform random byte base, each copy with
    substitutions to another byte and insertions of random bytes."""
    rng = np.random.default_rng(seed)
    base = rng.integers(0, 256, size=base_len).astype(np.uint8)
    parts = [mutate(rng, base, sub_rate, indel_rate, indel_max, sigma=256) for _ in range(copies)]
    return np.concatenate(parts)


def markov2_printable(n: int, seed: int, alpha: float = 0.5, chains: int = 1024) -> np.ndarray:
    """Order-2 Markov text over the 96 printable bytes 32..127: transition rows drawn from
    Dirichlet(alpha) per two-symbol context; `chains` independent chains of equal length are
    advanced together and concatenated, each step sampled by a 1024-bucket inverse-CDF table
    per context (the buckets quantize the row's CDF at multiples of 1/1024)."""
    rng = np.random.default_rng(seed)
    k = 96
    probs = rng.dirichlet(np.full(k, alpha), size=k * k)
    cdf = np.cumsum(probs, axis=1)
    cdf[:, -1] = 1.0
    q = (np.arange(1024) + 0.5) / 1024.0
    table = np.empty((k * k, 1024), dtype=np.uint8)
    for c in range(k * k):
        table[c] = np.minimum(np.searchsorted(cdf[c], q), k - 1)
    steps = -(-n // chains)
    out = np.empty((steps, chains), dtype=np.uint8)
    ctx = rng.integers(0, k * k, size=chains)
    for t in range(steps):
        u = rng.integers(0, 1024, size=chains)
        c = table[ctx, u]
        out[t] = c
        ctx = (ctx % k) * k + c
    return (out.T.reshape(-1)[:n] + 32).astype(np.uint8)


# Workloads as named in the bench/test configuration.
CONFIGS = {
    # 32 MiB: 1 MiB base x 32 copies, 0.1% substitutions + 0.01% indels of length 1..10
    "dna32m": dict(base_len=1 << 20, copies=32, sub_rate=0.001, indel_rate=0.0001, indel_max=10,
                   seed=SEED_BASE + 0),
    # 256 MiB: 4 MiB base x 64 copies, 0.05% substitutions + 0.005% indels of length 1..50
    "dna256m": dict(base_len=4 << 20, copies=64, sub_rate=0.0005, indel_rate=0.00005, indel_max=50,
                    seed=SEED_BASE + 1),
    # 1 GiB: 8 MiB base x 128 copies, 0.02% mutations split 10:1 substitutions:indels
    "dna1g": dict(base_len=8 << 20, copies=128, sub_rate=0.0002 * 10 / 11,
                  indel_rate=0.0002 / 11, indel_max=10, seed=SEED_BASE + 2),
}


def batch_collection(count: int, seed: int = SEED_BASE + 4, threads: int = 16, min_len: int = 64 << 10,
                     max_len: int = 16 << 20, max_copies: int = 64):
    """`count` independent texts shaped like the batched configuration: lengths log-uniform in
    [min_len, max_len], each a random base replicated 2..max_copies times with substitutions at a
    log-uniform rate in [1e-5, 1e-2] (indels at a tenth of it, length 1..10); even texts over
    A/C/G/T, odd texts over all 256 byte values.  The parameters of text i are drawn from one
    seeded stream, the text itself from its own seed, so the result does not depend on
    `threads`."""
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        n_total = int(np.exp(rng.uniform(np.log(min_len), np.log(max_len))))
        copies = int(rng.integers(2, max_copies + 1))
        rate = float(np.exp(rng.uniform(np.log(1e-5), np.log(1e-2))))
        specs.append((i, n_total, copies, rate, int(rng.integers(1 << 31))))

    def make(spec):
        i, n_total, copies, rate, s = spec
        make_text = repetitive_bytes if i % 2 else repetitive_dna
        return make_text(max(1024, n_total // copies), copies, rate, rate / 10, 10, seed=s)

    with ThreadPoolExecutor(max(1, threads)) as ex:
        return list(ex.map(make, specs))


def workload(name: str, scale: float = 1.0) -> np.ndarray:
    """Text of a named workload; `scale` < 1 shrinks the base length (for quick tests)."""
    if name == "markov256m":
        # low-repetition control of the 1 GiB collection: 256 MiB order-2 Markov text
        return markov2_printable(max(1 << 16, int((256 << 20) * scale)), SEED_BASE + 3)
    cfg = dict(CONFIGS[name])
    cfg["base_len"] = max(1024, int(cfg["base_len"] * scale))
    return repetitive_dna(**cfg)
