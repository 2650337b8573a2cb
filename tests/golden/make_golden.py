"""Writes tests/golden/bwt_small.json: brute-force BWTs (rotation sort of all suffixes with a
unique smallest terminator) of small texts, in the reference output format and the libcubwt
convention.  Pure Python, independent of the oracle and of the library."""
import json
import os
import random


def brute(T: bytes):
    n = len(T)
    sa = sorted(range(n), key=lambda i: T[i:])
    O = [T[(i - 1) % n] for i in sa]
    index = sa.index(0) + 1
    L = [T[n - 1]] + O[: index - 1] + O[index:]
    return O, index, L


def main():
    rng = random.Random(20261203)
    cases = [b"mississippi", b"banana", b"abracadabra", b"a", b"aaaaaaaa", b"zyxwvutsrq",
             bytes([0, 1, 255, 1, 0, 255, 255, 0]), b"ACGTACGTTGCA" * 3]
    for _ in range(24):
        n = rng.randint(1, 64)
        k = rng.choice([2, 4, 26, 256])
        cases.append(bytes(rng.randrange(k) for _ in range(n)))
    out = []
    for T in cases:
        O, index, L = brute(T)
        out.append({"text": list(T), "reference_format": O, "primary_index": index, "libcubwt_L": L})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bwt_small.json")
    with open(path, "w") as f:
        json.dump(out, f)
    print(path, len(out))


if __name__ == "__main__":
    main()
