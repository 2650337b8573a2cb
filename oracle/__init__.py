"""CPU oracle for the BWT-build path (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py``
may import this module, and only as the checker.  The product library never calls
it and has no CPU fallback.

Wraps ``oracle/liboracle.so`` (built from ``oracle.c`` by ``oracle/Makefile``):
SA-IS suffix sorting, the reference-format BWT (Prefix_Free_Parsing_W=5_P=16_kommentiert.cu:4776-4790),
the libcubwt primary-index convention (libcubwt.cu:2242-2254), inverse BWT, and a
restatement of the reference prefix-free-parsing BWT with runtime (W, P).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_CPU_PFP_PATH = os.path.join(_HERE, "libcpupfp.so")
_lib = None
_cpu = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        u8p = ctypes.POINTER(ctypes.c_uint8)
        i32p = ctypes.POINTER(ctypes.c_int32)
        i64p = ctypes.POINTER(ctypes.c_int64)
        i64 = ctypes.c_int64
        L.oracle_sa.argtypes = [u8p, i64, i32p]
        L.oracle_sa.restype = ctypes.c_int
        L.oracle_bwt.argtypes = [u8p, i64, u8p]
        L.oracle_bwt.restype = i64
        L.oracle_ref_to_cubwt.argtypes = [u8p, i64, i64, u8p]
        L.oracle_ref_to_cubwt.restype = None
        L.oracle_cubwt_to_ref.argtypes = [u8p, i64, i64, u8p]
        L.oracle_cubwt_to_ref.restype = None
        L.oracle_unbwt.argtypes = [u8p, i64, i64, u8p]
        L.oracle_unbwt.restype = ctypes.c_int
        L.oracle_unbwt_mt.argtypes = [u8p, i64, i64, u8p, ctypes.c_int]
        L.oracle_unbwt_mt.restype = ctypes.c_int
        L.oracle_pfp_triggers.argtypes = [u8p, i64, ctypes.c_int, ctypes.c_int, i32p]
        L.oracle_pfp_triggers.restype = i64
        L.oracle_pfp_parse.argtypes = [u8p, i64, ctypes.c_int, ctypes.c_int, i32p, i32p, i64p]
        L.oracle_pfp_parse.restype = i64
        L.oracle_pfp_bwt.argtypes = [u8p, i64, ctypes.c_int, ctypes.c_int, u8p,
                                     ctypes.POINTER(PfpInfo)]
        L.oracle_pfp_bwt.restype = i64
        L.oracle_lyndon_factors.argtypes = [u8p, i64, i64p]
        L.oracle_lyndon_factors.restype = i64
        L.oracle_lyndon_array.argtypes = [u8p, i64, i32p]
        L.oracle_lyndon_array.restype = ctypes.c_int
        L.oracle_lyndon_ranks.argtypes = [u8p, i64, i32p]
        L.oracle_lyndon_ranks.restype = i64
        L.oracle_lyndon_bwt.argtypes = [u8p, i64, u8p]
        L.oracle_lyndon_bwt.restype = i64
        L.oracle_lyndon_grammar_bwt.argtypes = [u8p, i64, u8p, ctypes.POINTER(LyndonGrammarInfo)]
        L.oracle_lyndon_grammar_bwt.restype = i64
        _lib = L
    return _lib


class PfpInfo(ctypes.Structure):
    _fields_ = [("t", ctypes.c_int64), ("d", ctypes.c_int64), ("distinct", ctypes.c_int64),
                ("s_size", ctypes.c_int64), ("mixed_intervals", ctypes.c_int64),
                ("mixed_positions", ctypes.c_int64)]


def _u8(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


def _i32(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def _as_text(T) -> np.ndarray:
    if isinstance(T, (bytes, bytearray)):
        return np.frombuffer(bytes(T), dtype=np.uint8).copy()
    return np.ascontiguousarray(T, dtype=np.uint8)


def suffix_array(T) -> np.ndarray:
    T = _as_text(T)
    SA = np.empty(len(T), dtype=np.int32)
    if len(T) and lib().oracle_sa(_u8(T), len(T), _i32(SA)) != 0:
        raise MemoryError("oracle_sa failed")
    return SA


def bwt(T):
    """Reference-format BWT O and the libcubwt primary index of T."""
    T = _as_text(T)
    O = np.empty(len(T), dtype=np.uint8)
    idx = lib().oracle_bwt(_u8(T), len(T), _u8(O)) if len(T) else 0
    if idx < 0:
        raise MemoryError("oracle_bwt failed")
    return O, int(idx)


def ref_to_cubwt(O: np.ndarray, index: int) -> np.ndarray:
    O = _as_text(O)
    L = np.empty_like(O)
    lib().oracle_ref_to_cubwt(_u8(O), len(O), index, _u8(L))
    return L


def cubwt_to_ref(L: np.ndarray, index: int) -> np.ndarray:
    L = _as_text(L)
    O = np.empty_like(L)
    lib().oracle_cubwt_to_ref(_u8(L), len(L), index, _u8(O))
    return O


def unbwt(L: np.ndarray, index: int) -> np.ndarray:
    """Inverse of the libcubwt format (L, primary index)."""
    L = _as_text(L)
    T = np.empty_like(L)
    rc = lib().oracle_unbwt(_u8(L), len(L), index, _u8(T))
    if rc != 0:
        raise ValueError("not a valid BWT / primary index")
    return T


def unbwt_mt(L: np.ndarray, index: int, threads: int = 0) -> np.ndarray:
    """oracle_unbwt with `threads` threads (0: all online CPUs)."""
    L = _as_text(L)
    T = np.empty_like(L)
    rc = lib().oracle_unbwt_mt(_u8(L), len(L), index, _u8(T), threads)
    if rc != 0:
        raise ValueError("not a valid BWT / primary index")
    return T


def pfp_triggers(T, W: int, P: int) -> np.ndarray:
    T = _as_text(T)
    t = lib().oracle_pfp_triggers(_u8(T), len(T), W, P, None)
    E = np.empty(max(t, 1), dtype=np.int32)
    lib().oracle_pfp_triggers(_u8(T), len(T), W, P, _i32(E))
    return E[:t]


def pfp_parse(T, W: int, P: int):
    """(dense lexicographic phrase rank per occurrence, parse suffix array, #distinct)."""
    T = _as_text(T)
    t = lib().oracle_pfp_triggers(_u8(T), len(T), W, P, None)
    d = t + 1
    rank = np.empty(d, dtype=np.int32)
    psa = np.empty(d, dtype=np.int32)
    du = ctypes.c_int64(0)
    lib().oracle_pfp_parse(_u8(T), len(T), W, P, _i32(rank), _i32(psa), ctypes.byref(du))
    return rank, psa, int(du.value)


def pfp_bwt(T, W: int, P: int):
    """Reference-format BWT, primary index and phase counts from the PFP restatement."""
    T = _as_text(T)
    O = np.empty(len(T), dtype=np.uint8)
    info = PfpInfo()
    idx = lib().oracle_pfp_bwt(_u8(T), len(T), W, P, _u8(O), ctypes.byref(info))
    if idx < 0:
        raise RuntimeError(f"oracle_pfp_bwt failed ({idx})")
    return O, int(idx), {k: getattr(info, k) for k, _ in PfpInfo._fields_}


def cpu_pfp_bwt(T, W: int, P: int, threads: int = 0):
    """Reference-format BWT and primary index from the multi-threaded C++ restatement of the
    reference PFP pipeline (oracle/cpu_pfp.cpp; the CPU baseline), with `threads` threads
    (0: all hardware threads).  Returns (O, index, parse length, distinct phrases)."""
    global _cpu
    if _cpu is None:
        if not os.path.exists(_CPU_PFP_PATH):
            build()
        c = ctypes.CDLL(_CPU_PFP_PATH)
        c.cpu_pfp_bwt.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                  ctypes.c_int, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
        c.cpu_pfp_bwt.restype = ctypes.c_int64
        _cpu = c
    T = _as_text(T)
    O = np.empty(len(T), dtype=np.uint8)
    d, du = ctypes.c_int64(), ctypes.c_int64()
    idx = _cpu.cpu_pfp_bwt(T.ctypes.data, len(T), W, P, O.ctypes.data, threads, ctypes.byref(d), ctypes.byref(du))
    if idx < 1:
        raise RuntimeError(f"cpu_pfp_bwt failed ({idx})")
    return O, int(idx), int(d.value), int(du.value)


def lyndon_factors(T) -> np.ndarray:
    T = _as_text(T)
    c = lib().oracle_lyndon_factors(_u8(T), len(T), None)
    F = np.empty(max(c, 1), dtype=np.int64)
    lib().oracle_lyndon_factors(_u8(T), len(T), F.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    return F[:c]


def lyndon_array(T) -> np.ndarray:
    T = _as_text(T)
    lam = np.empty(len(T), dtype=np.int32)
    lib().oracle_lyndon_array(_u8(T), len(T), _i32(lam))
    return lam


def lyndon_ranks(T):
    """(dense rank of each position's longest Lyndon prefix, number of distinct prefixes)."""
    T = _as_text(T)
    r = np.empty(len(T), dtype=np.int32)
    g = lib().oracle_lyndon_ranks(_u8(T), len(T), _i32(r))
    return r, int(g)


class LyndonGrammarInfo(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int64) for k in ("labels", "slots", "roots", "ups", "rounds", "items")]


def lyndon_grammar_bwt(T):
    """Lyndon-grammar BWT by induction over distinct labels: (reference-format BWT, primary
    index, sizes dict with labels, slots, roots, ups, rounds, items)."""
    T = _as_text(T)
    O = np.empty(len(T), dtype=np.uint8)
    info = LyndonGrammarInfo()
    idx = lib().oracle_lyndon_grammar_bwt(_u8(T), len(T), _u8(O), ctypes.byref(info)) if len(T) else 0
    return O, int(idx), {k: int(getattr(info, k)) for k, _ in LyndonGrammarInfo._fields_}


def lyndon_bwt(T):
    """Reference-format BWT and primary index from the Lyndon-grammar restatement."""
    T = _as_text(T)
    O = np.empty(len(T), dtype=np.uint8)
    idx = lib().oracle_lyndon_bwt(_u8(T), len(T), _u8(O)) if len(T) else 0
    return O, int(idx)
