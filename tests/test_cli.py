"""The CLI binaries reproduce the reference programs' file boundary: ./input.txt in,
./out.txt out, "n\\n" then "s " per symbol without a final newline, so that the reference's
compare scripts (cmp of two outputs) keep working."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from conftest import PKG_DIR

BIN = os.path.join(PKG_DIR, "bin")


def reference_text(vals) -> bytes:
    return (f"{len(vals)}\n" + "".join(f"{int(v)} " for v in vals)).encode()


def write_input(d, T):
    with open(os.path.join(d, "input.txt"), "wb") as f:
        f.write(reference_text(T))


def run(d, prog, *args):
    r = subprocess.run([os.path.join(BIN, prog), *args], cwd=d, capture_output=True, timeout=120)
    assert r.returncode == 0, r.stderr.decode()
    return open(os.path.join(d, "out.txt"), "rb").read()


@pytest.mark.gpu
def test_pfp_cli_matches_oracle_format(tmp_path):
    rng = np.random.default_rng(1)
    T = rng.integers(2, 255, size=20000).astype(np.uint8)
    T = np.concatenate([T, T[:5000], T])
    write_input(tmp_path, T)
    O, _ = oracle.bwt(T)
    out = run(tmp_path, "pfp_bwt")
    assert out == reference_text(O)
    assert not out.endswith(b"\n") and out.endswith(b" ")
    out35 = run(tmp_path, "pfp_bwt", "-W", "35", "-P", "81")
    assert out35 == out
    # the learned (W,P) choice replaces the benchmark script's per-file table
    assert run(tmp_path, "pfp_bwt", "--learned") == out


@pytest.mark.gpu
def test_compare_script_semantics(tmp_path):
    # cmp output_<i>.txt output_libcubwt_<i>.txt must report equality (compare_script.sh:10)
    rng = np.random.default_rng(2)
    base = rng.choice(np.frombuffer(b"ACGT", dtype=np.uint8), size=30000)
    T = np.concatenate([base, base[17:], base[:9999]])
    write_input(tmp_path, T)
    a = run(tmp_path, "pfp_bwt")
    b = run(tmp_path, "cubwt_compat")
    assert a == b


@pytest.mark.gpu
def test_cubwt_driver_quirk(tmp_path):
    # primary index == n happens when suffix 0 is the largest suffix, e.g. "zyx...a"-like text
    T = np.array([200] + [3] * 40, dtype=np.uint8)
    O, idx = oracle.bwt(T)
    assert idx == len(T)
    write_input(tmp_path, T)
    full = run(tmp_path, "cubwt_compat")
    assert full == reference_text(O)
    quirk = run(tmp_path, "cubwt_compat", "--driver-quirk")
    L = oracle.ref_to_cubwt(O, idx)
    assert quirk == (f"{len(T)}\n" + "".join(f"{v} " for v in L[1:])).encode()


@pytest.mark.gpu
def test_raw_mode(tmp_path):
    rng = np.random.default_rng(3)
    T = rng.integers(0, 256, size=50000).astype(np.uint8)
    (tmp_path / "in.bin").write_bytes(T.tobytes())
    r = subprocess.run([os.path.join(BIN, "pfp_bwt"), "--raw", "-i", "in.bin", "-o", "out.bin"],
                       cwd=tmp_path, capture_output=True, timeout=120)
    assert r.returncode == 0, r.stderr.decode()
    O, _ = oracle.bwt(T)
    assert (tmp_path / "out.bin").read_bytes() == O.tobytes()


def test_malformed_input_rejected(tmp_path):
    (tmp_path / "input.txt").write_bytes(b"5\n1 2 3")
    r = subprocess.run([os.path.join(BIN, "pfp_bwt")], cwd=tmp_path, capture_output=True, timeout=60)
    assert r.returncode != 0


def test_malformed_counts_rejected(tmp_path):
    # a declared count beyond what the file can hold, and a count beyond the 64-bit range, are
    # malformed input (exit 1 with a message), not an allocation abort
    for text in (b"99999999999999\n1 2 3 ", b"123456789012345678901234567890\n1 "):
        (tmp_path / "input.txt").write_bytes(text)
        r = subprocess.run([os.path.join(BIN, "pfp_bwt")], cwd=tmp_path, capture_output=True, timeout=60)
        assert r.returncode == 1, r.stderr.decode()
        assert b"malformed" in r.stderr
