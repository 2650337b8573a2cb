"""The arithmetic of the single-pass trigger scan, restated on the CPU and pinned.

The kernel replaces `V mod P == 0` (PFPk:468-475 reduces mod P at every step) by exact window
values and a multiply-and-rotate divisibility test.  These tests pin both facts against
direct modular arithmetic: the test itself exhaustively at 16 bits and on random values at
32/64 bits, and the rolled exact value against the reference's per-window formula.
"""
import numpy as np
import pytest

import oracle


def inv_odd(q, bits):
    x = q
    for _ in range(6):
        x = (x * (2 - q * x)) % (1 << bits)
    return x


def divisible(v, P, bits):
    s = (P & -P).bit_length() - 1
    q = P >> s
    mask = (1 << bits) - 1
    if q == 1:
        return v & (P - 1) == 0
    x = (v * inv_odd(q, bits)) & mask
    r = ((x >> s) | (x << (bits - s))) & mask if s else x
    return r <= mask // P


@pytest.mark.parametrize("P", [1, 2, 3, 6, 7, 12, 16, 81, 96, 255, 1024, 4095, 65535])
def test_divisibility_exhaustive_16bit(P):
    bits = 16
    s = (P & -P).bit_length() - 1
    q = P >> s
    vs = np.arange(1 << bits, dtype=np.uint64)
    if q == 1:
        got = (vs & np.uint64(P - 1)) == 0
    else:
        x = (vs * np.uint64(inv_odd(q, bits))) & np.uint64(0xFFFF)
        r = ((x >> np.uint64(s)) | (x << np.uint64(bits - s))) & np.uint64(0xFFFF) if s else x
        got = r <= np.uint64(0xFFFF // P)
    assert (got == (vs % np.uint64(P) == 0)).all()


@pytest.mark.parametrize("P", [3, 11, 13, 17, 23, 31, 41, 53, 61, 71, 81, 255, 4095, 65535])
def test_divisibility_24bit_odd_exhaustive(P):
    # the kernel's test for odd P and window values below 2^24 (W <= 16): a 24 x 24-bit multiply
    # (low 32 bits kept), shifted left by 8 within 32 bits, against floor((2^24 - 1) / P) << 8 | 0xFF;
    # every v < 2^24, so a dropped mask, a wrong inverse width or bound fails here
    inv24 = inv_odd(P, 64) & 0xFFFFFF
    lim = ((0xFFFFFF // P) << 8) | 0xFF
    vs = np.arange(1 << 24, dtype=np.uint64)
    x = (vs * np.uint64(inv24)) & np.uint64(0xFFFFFFFF)
    got = ((x << np.uint64(8)) & np.uint64(0xFFFFFFFF)) <= np.uint64(lim)
    assert (got == (vs % np.uint64(P) == 0)).all()


@pytest.mark.parametrize("bits", [32, 64])
def test_divisibility_random(bits):
    rng = np.random.default_rng(bits)
    for P in list(range(1, 200)) + [256, 1000, 4096, 65535, 65534, 40960]:
        vals = [int(v) for v in rng.integers(0, 1 << 62, size=200)] + [0, P, 2 * P, P - 1, (1 << bits) - 1]
        vals += [P * int(k) for k in rng.integers(0, min(((1 << bits) - 1) // P, 1 << 62), size=50)]
        vals += [((1 << bits) - 1) // P * P, ((1 << bits) - 1) // P * P + 1]
        for v in vals:
            v &= (1 << bits) - 1
            assert divisible(v, P, bits) == (v % P == 0), (v, P, bits)


def rolled_triggers(T, W, P):
    """The kernel's rule: exact rolled value (mod 2^bits), divisibility test, first-window
    prefilter V == V0 confirmed byte by byte."""
    bits = 32 if W <= 24 else 64
    mask = (1 << bits) - 1
    n = len(T)
    t = [int(c) for c in T]
    v0 = 0
    for k in range(min(W, n)):
        v0 = (2 * v0 + t[k]) & mask
    out = []
    v = 0
    for e in range(n):
        v = (2 * v + t[e] - (t[e - W] << W if e >= W else 0)) & mask
        if e < W:
            continue
        hit = divisible(v, P, bits)
        if not hit and v == v0:
            hit = t[e - W + 1:e + 1] == t[:W]
        if hit:
            out.append(e)
    return out


@pytest.mark.parametrize("W,P", [(1, 2), (5, 16), (5, 11), (10, 23), (24, 81), (24, 96), (25, 13),
                                 (35, 81), (56, 65535), (56, 40)])
def test_rolled_exact_value_matches_reference_rule(W, P):
    rng = np.random.default_rng(W * 7 + P)
    T = rng.integers(0, 256, size=4000).astype(np.uint8)
    T[100:100 + W] = T[:W]  # the first window recurs
    assert rolled_triggers(T, W, P) == list(oracle.pfp_triggers(T, W, P))


def test_exact_value_widths():
    # the largest window value, 255 * (2^W - 1), fits the kernel's word for W <= 24 / 56
    assert 255 * ((1 << 24) - 1) < (1 << 32) <= 255 * ((1 << 25) - 1)
    assert 255 * ((1 << 56) - 1) < (1 << 64) <= 255 * ((1 << 57) - 1)
