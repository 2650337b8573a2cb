"""mibwt — MI355X-native Burrows-Wheeler transform builder.

Python mirror of the C ABI in include/mibwt.h (ctypes over lib/libmibwt.so).  The library
is the product; this module only binds it.  There is no CPU fallback: if the HIP library is
missing every call raises.

Entry points mirror the reference programs:
  - ``bwt(T, method="pfp", W=5, P=16)`` -> (L, primary_index) in libcubwt format
    (libcubwt.cuh:77, libcubwt.cu:2414);
  - ``bwt_reference_format(...)`` -> O as the reference programs print it
    (Prefix_Free_Parsing_W=5_P=16_kommentiert.cu:4776-4790);
  - ``Storage`` wraps libcubwt_allocate_device_storage / libcubwt_free_device_storage.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libmibwt.so")
BIN_DIR = os.path.join(_HERE, "bin")

NO_ERROR = 0
BAD_PARAMETER = -1
NOT_ENOUGH_MEMORY = -2
GPU_ERROR = -7
GPU_NOT_SUPPORTED = -8
GPU_NOT_ENOUGH_MEMORY = -9
HASH_COLLISION = -20
INDEX_RANGE = -21

METHOD_PFP = 0
METHOD_LYNDON = 1

EXPORTED = [
    "libcubwt_allocate_device_storage",
    "libcubwt_free_device_storage",
    "libcubwt_bwt",
    "libcubwt_bwt_aux",
    "libcubwt_unbwt",
    "mibwt_unbwt_device",
    "mibwt_pfp_bwt",
    "mibwt_pfp_bwt_device",
    "mibwt_pfp_bwt_device_noalloc",
    "mibwt_lyndon_bwt",
    "mibwt_lyndon_bwt_device",
    "mibwt_lyndon_array",
    "mibwt_pfp_parse",
    "mibwt_pfp_triggers",
    "mibwt_last_phase_stats",
    "mibwt_set_kernel_timing",
    "mibwt_time_kernel",
    "mibwt_batch_bwt",
    "mibwt_batch_bwt_device",
    "mibwt_wp_grid",
    "mibwt_wp_features",
    "mibwt_select_wp",
    "mibwt_select_wp_device",
    "mibwt_selector_load",
    "mibwt_mlp_grad",
    "mibwt_mlp_predict",
    "mibwt_mlp_train",
    "mibwt_mlp_train_ensemble",
    "mibwt_autotune",
    "mibwt_tuning_load",
    "mibwt_tuning_use",
    "mibwt_tuning_get",
    "mibwt_device_alloc",
    "mibwt_device_free",
    "mibwt_copy",
    "mibwt_synchronize",
    "mibwt_stream_create",
    "mibwt_stream_destroy",
    "mibwt_copy_async",
]

_lib = None


class BwtError(RuntimeError):
    def __init__(self, code: int, what: str = ""):
        super().__init__(f"mibwt error {code} {what}".strip())
        self.code = code


MAX_PHASES = 16


class PhaseStats(ctypes.Structure):
    _fields_ = [
        ("num_phases", ctypes.c_int32),
        ("name", (ctypes.c_char * 24) * MAX_PHASES),
        ("ms", ctypes.c_double * MAX_PHASES),
        ("bytes", ctypes.c_double * MAX_PHASES),
        ("n", ctypes.c_int64), ("W", ctypes.c_int64), ("P", ctypes.c_int64),
        ("triggers", ctypes.c_int64), ("phrases", ctypes.c_int64),
        ("distinct_phrases", ctypes.c_int64), ("dict_length", ctypes.c_int64),
        ("alpha_groups", ctypes.c_int64), ("mixed_entries", ctypes.c_int64),
        ("doubling_rounds_dict", ctypes.c_int64), ("doubling_rounds_parse", ctypes.c_int64),
        ("total_ms", ctypes.c_double),
        ("kt_launches", ctypes.c_int64), ("kt_ms_total", ctypes.c_double),
        ("kt_bytes_total", ctypes.c_double),
        ("retries", ctypes.c_int64),
        ("device_bytes", ctypes.c_int64),
        ("induction", ctypes.c_int64),
    ]


def lib():
    """Load lib/libmibwt.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run build.py (hipcc, gfx950) first")
        L = ctypes.CDLL(LIB_PATH)
        vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32
        u8p = ctypes.POINTER(ctypes.c_uint8)
        L.libcubwt_allocate_device_storage.argtypes = [ctypes.POINTER(vp), i64]
        L.libcubwt_free_device_storage.argtypes = [vp]
        L.libcubwt_bwt.argtypes = [vp, u8p, u8p, i64]
        L.libcubwt_bwt_aux.argtypes = [vp, u8p, u8p, i64, i64, ctypes.POINTER(ctypes.c_uint32)]
        L.libcubwt_unbwt.argtypes = [vp, u8p, u8p, i64, ctypes.POINTER(ctypes.c_int32), i32]
        L.mibwt_unbwt_device.argtypes = [vp, vp, vp, i64, i64, vp]
        L.mibwt_pfp_bwt.argtypes = [vp, u8p, u8p, i64, i32, i32]
        L.mibwt_pfp_bwt_device.argtypes = [vp, vp, vp, i64, i32, i32, vp]
        L.mibwt_pfp_bwt_device_noalloc.argtypes = [vp, vp, i64, i32, i32, vp]
        L.mibwt_lyndon_bwt.argtypes = [vp, u8p, u8p, i64]
        L.mibwt_lyndon_bwt_device.argtypes = [vp, vp, vp, i64, vp]
        L.mibwt_lyndon_array.argtypes = [vp, u8p, i64, ctypes.POINTER(ctypes.c_uint32)]
        L.mibwt_pfp_parse.argtypes = [vp, u8p, i64, i32, i32, ctypes.POINTER(ctypes.c_uint32),
                                      ctypes.POINTER(ctypes.c_uint32), i64, ctypes.POINTER(ctypes.c_int64)]
        L.mibwt_pfp_triggers.argtypes = [vp, u8p, i64, i32, i32, ctypes.POINTER(ctypes.c_uint32), i64]
        L.mibwt_last_phase_stats.argtypes = [vp, ctypes.POINTER(PhaseStats)]
        L.mibwt_set_kernel_timing.argtypes = [vp, ctypes.c_char_p]
        L.mibwt_time_kernel.argtypes = [vp, ctypes.c_char_p, i32, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_double)]
        fp = ctypes.POINTER(ctypes.c_float)
        i32p = ctypes.POINTER(ctypes.c_int32)
        L.mibwt_batch_bwt.argtypes = [vp, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p),
                                      ctypes.POINTER(ctypes.c_int64), i64, ctypes.POINTER(ctypes.c_int64),
                                      i32, i32, i32, i32]
        L.mibwt_batch_bwt_device.argtypes = [vp, vp, ctypes.POINTER(ctypes.c_int64), i64, vp,
                                             ctypes.POINTER(ctypes.c_int64), i32, i32, vp]
        L.mibwt_wp_grid.argtypes = [i32p, i32p]
        L.mibwt_wp_features.argtypes = [vp, vp, i64, fp]
        L.mibwt_select_wp.argtypes = [vp, i64, i32p, i32p]
        L.mibwt_select_wp_device.argtypes = [vp, vp, i64, i32p, i32p, fp]
        L.mibwt_selector_load.argtypes = [ctypes.c_char_p]
        L.mibwt_mlp_grad.argtypes = [vp, fp, fp, fp, i64, fp, fp]
        L.mibwt_mlp_predict.argtypes = [vp, fp, fp, i64, fp]
        L.mibwt_mlp_train.argtypes = [vp, fp, fp, i64, i32, i32, ctypes.c_float, ctypes.c_uint32,
                                      ctypes.c_char_p, fp]
        L.mibwt_mlp_train_ensemble.argtypes = [vp, fp, fp, i64, i32, i32, ctypes.c_float, ctypes.c_uint32, i32,
                                               ctypes.c_char_p, fp]
        L.mibwt_autotune.argtypes = [vp, i64, i32, ctypes.c_char_p]
        L.mibwt_tuning_load.argtypes = [ctypes.c_char_p]
        L.mibwt_tuning_use.argtypes = [i32]
        L.mibwt_tuning_get.argtypes = [i64, i32p]
        L.mibwt_device_alloc.argtypes = [ctypes.POINTER(vp), i64]
        L.mibwt_device_free.argtypes = [vp]
        L.mibwt_copy.argtypes = [vp, vp, i64, i32]
        L.mibwt_synchronize.argtypes = []
        L.mibwt_stream_create.argtypes = [ctypes.POINTER(vp)]
        L.mibwt_stream_destroy.argtypes = [vp]
        L.mibwt_copy_async.argtypes = [vp, vp, i64, i32, vp]
        for name in EXPORTED:
            getattr(L, name).restype = i64
        _lib = L
    return _lib


def _u8(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


def _fp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


NUM_CAND = 41  # the W x P grid (40) and the reference default W=5,P=16 (last)


def wp_grid():
    """The selector's candidates [(W, P)]: the 4 x 10 grid, then (5, 16)."""
    W = np.zeros(64, dtype=np.int32)
    P = np.zeros(64, dtype=np.int32)
    count = lib().mibwt_wp_grid(W.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                P.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    return list(zip(W[:count].tolist(), P[:count].tolist()))


def read_selector_model(path: str) -> dict:
    """The selector model file: normalisation and the members' fp32 parameter vectors (a single
    network, magic "PLBM", or an ensemble, magic "PLBE" followed by the member count)."""
    raw = np.fromfile(path, dtype=np.uint32)
    magic = int(raw[0])
    members, hdr = (1, 4) if magic == 0x4D424C50 else (int(raw[4]), 5)
    f = raw.view(np.float32)
    xm, xs = f[hdr:hdr + 64], f[hdr + 64:hdr + 128]
    ym, ys = f[hdr + 128:hdr + 131], f[hdr + 131:hdr + 134]
    params = f[hdr + 134:].reshape(members, -1)
    return {"members": members, "xmean": xm, "xstd": xs, "ymean": ym, "ystd": ys, "params": params}


def smooth_along_p(t: np.ndarray, num_p: int = 10) -> np.ndarray:
    """Mean of each grid candidate's predicted time with its P-neighbours of the same W (the
    grid is W-major with num_p values of P; candidates after the grid are left as they are)."""
    s = t.astype(np.float64).copy()
    grid = (len(t) // num_p) * num_p
    for c in range(grid):
        pi = c % num_p
        nb = [t[c]] + ([t[c - 1]] if pi > 0 else []) + ([t[c + 1]] if pi + 1 < num_p else [])
        s[c] = np.mean(nb)
    return s


def select_wp(T):
    """(W, P) the selector chooses for the host text T, and whether a trained model chose
    (mibwt_select_wp)."""
    T = _as_text(T)
    W, P = ctypes.c_int32(), ctypes.c_int32()
    rc = lib().mibwt_select_wp(_u8(T), len(T), ctypes.byref(W), ctypes.byref(P))
    if rc < 0:
        raise BwtError(rc, "mibwt_select_wp")
    return (W.value, P.value), bool(rc)


def selector_load(path: str):
    rc = lib().mibwt_selector_load(path.encode())
    if rc != NO_ERROR:
        raise BwtError(rc, "mibwt_selector_load")


MLP_SHAPES = [("W1", (128, 64)), ("b1", (128,)), ("W2", (128, 128)), ("b2", (128,)),
              ("W3", (3, 128)), ("b3", (3,))]
MLP_PARAMS = sum(int(np.prod(s)) for _, s in MLP_SHAPES)


def _as_text(T) -> np.ndarray:
    if isinstance(T, (bytes, bytearray, memoryview)):
        return np.frombuffer(bytes(T), dtype=np.uint8).copy()
    return np.ascontiguousarray(T, dtype=np.uint8)


class Storage:
    """A libcubwt-style device storage (one HIP stream, grow-only buffers)."""

    def __init__(self, max_length: int):
        self.handle = ctypes.c_void_p()
        rc = lib().libcubwt_allocate_device_storage(ctypes.byref(self.handle), int(max_length))
        if rc != NO_ERROR:
            raise BwtError(rc, "libcubwt_allocate_device_storage")
        self.max_length = int(max_length)

    def close(self):
        if self.handle:
            rc = lib().libcubwt_free_device_storage(self.handle)
            self.handle = ctypes.c_void_p()
            if rc != NO_ERROR:
                raise BwtError(rc, "libcubwt_free_device_storage")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- host-buffer entry points (libcubwt format)
    def bwt(self, T, method: str = "pfp", W: int = 5, P: int = 16):
        T = _as_text(T)
        L = np.empty_like(T)
        if method == "pfp":
            rc = lib().mibwt_pfp_bwt(self.handle, _u8(T), _u8(L), len(T), W, P)
        elif method == "lyndon":
            rc = lib().mibwt_lyndon_bwt(self.handle, _u8(T), _u8(L), len(T))
        elif method == "cubwt":
            rc = lib().libcubwt_bwt(self.handle, _u8(T), _u8(L), len(T))
        else:
            raise ValueError(method)
        if rc < 0:
            raise BwtError(rc, method)
        return L, int(rc)

    # ---- device-buffer entry points (reference format)
    def bwt_device(self, dT_ptr: int, dO_ptr: int, n: int, method: str = "pfp", W: int = 5,
                   P: int = 16, stream: int | None = None) -> int:
        if method == "pfp":
            rc = lib().mibwt_pfp_bwt_device(self.handle, dT_ptr, dO_ptr, n, W, P, stream)
        else:
            rc = lib().mibwt_lyndon_bwt_device(self.handle, dT_ptr, dO_ptr, n, stream)
        if rc < 0:
            raise BwtError(rc, method + "_device")
        return int(rc)

    def unbwt(self, L, index: int) -> np.ndarray:
        """Inverse BWT (libcubwt_unbwt) of L in libcubwt format with its primary index."""
        L = _as_text(L)
        U = np.empty_like(L)
        rc = lib().libcubwt_unbwt(self.handle, _u8(L), _u8(U), len(L), None, int(index))
        if rc != NO_ERROR:
            raise BwtError(rc, "libcubwt_unbwt")
        return U

    def bwt_aux(self, T, r: int):
        """(L, I): the libcubwt BWT and the sampled rows I[j] of text positions j*r."""
        T = _as_text(T)
        L = np.empty_like(T)
        I = np.zeros((len(T) - 1) // r + 1 if r > 0 else 1, dtype=np.uint32)  # r < 4 is rejected
        rc = lib().libcubwt_bwt_aux(self.handle, _u8(T), _u8(L), len(T), int(r),
                                    I.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
        if rc != NO_ERROR:
            raise BwtError(rc, "libcubwt_bwt_aux")
        return L, I

    def unbwt_device(self, dO_ptr: int, dU_ptr: int, n: int, index: int, stream: int | None = None):
        """Inverse of a device BWT in reference format (the *_device entries' output)."""
        rc = lib().mibwt_unbwt_device(self.handle, dO_ptr, dU_ptr, n, int(index), stream)
        if rc != NO_ERROR:
            raise BwtError(rc, "mibwt_unbwt_device")

    def pfp_triggers(self, T, W: int = 5, P: int = 16) -> np.ndarray:
        """Trigger end positions (ascending) of the prefix-free parse of T."""
        T = _as_text(T)
        cap = max(1, len(T))
        E = np.empty(cap, dtype=np.uint32)
        rc = lib().mibwt_pfp_triggers(self.handle, _u8(T), len(T), W, P,
                                      E.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), cap)
        if rc < 0:
            raise BwtError(rc, "mibwt_pfp_triggers")
        return E[:rc].copy()

    def pfp_parse(self, T, W: int = 5, P: int = 16):
        """(lexicographic phrase rank per occurrence, parse suffix array, #distinct phrases)."""
        T = _as_text(T)
        cap = len(T) + 1
        rank = np.empty(cap, dtype=np.uint32)
        psa = np.empty(cap, dtype=np.uint32)
        du = ctypes.c_int64()
        d = lib().mibwt_pfp_parse(self.handle, _u8(T), len(T), W, P,
                                  rank.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                  psa.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), cap, ctypes.byref(du))
        if d < 0:
            raise BwtError(d, "mibwt_pfp_parse")
        return rank[:d].copy(), psa[:d].copy(), int(du.value)

    def lyndon_array(self, T) -> np.ndarray:
        T = _as_text(T)
        lam = np.empty(len(T), dtype=np.uint32)
        rc = lib().mibwt_lyndon_array(self.handle, _u8(T), len(T),
                                      lam.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
        if rc < 0:
            raise BwtError(rc, "mibwt_lyndon_array")
        return lam

    def batch_bwt(self, texts, method: str = "pfp", W: int = 5, P: int = 16, workers: int = 0):
        """[(L, primary_index)] for independent texts; W=0 lets the selector choose per text."""
        arrs = [_as_text(t) for t in texts]
        outs = [np.empty_like(t) for t in arrs]
        k = len(arrs)
        Tp = (ctypes.c_void_p * k)(*[t.ctypes.data for t in arrs])
        Lp = (ctypes.c_void_p * k)(*[o.ctypes.data for o in outs])
        ns = (ctypes.c_int64 * k)(*[len(t) for t in arrs])
        idx = (ctypes.c_int64 * k)()
        m = METHOD_PFP if method == "pfp" else METHOD_LYNDON
        rc = lib().mibwt_batch_bwt(self.handle, Tp, Lp, ns, k, idx, m, W, P, workers)
        if rc < 0:
            raise BwtError(rc, "mibwt_batch_bwt")
        return [(outs[i], int(idx[i])) for i in range(k)]

    def batch_bwt_device(self, dT_ptr: int, offsets, dO_ptr: int, W: int = 5, P: int = 16,
                         stream: int | None = None) -> np.ndarray:
        """Primary indices of the texts dT[offsets[i]:offsets[i+1]] (device); their BWTs in
        reference format go to dO at the same offsets (mibwt_batch_bwt_device)."""
        offs = np.ascontiguousarray(offsets, dtype=np.int64)
        k = len(offs) - 1
        idx = np.zeros(max(k, 1), dtype=np.int64)
        rc = lib().mibwt_batch_bwt_device(self.handle, dT_ptr, offs.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), k,
                                          dO_ptr, idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), W, P, stream)
        if rc < 0:
            raise BwtError(rc, "mibwt_batch_bwt_device")
        return idx[:k]

    def autotune(self, max_bytes: int, reps: int = 2, out_path: str | None = None) -> int:
        """Measure the launch-shape table up to max_bytes on the device (mibwt_autotune)."""
        rc = lib().mibwt_autotune(self.handle, int(max_bytes), int(reps), out_path.encode() if out_path else None)
        if rc < 0:
            raise BwtError(rc, "mibwt_autotune")
        return int(rc)

    # ---- learned (W,P) selector
    def wp_features(self, dT_ptr: int, n: int) -> np.ndarray:
        X = np.zeros((NUM_CAND, 64), dtype=np.float32)
        rc = lib().mibwt_wp_features(self.handle, dT_ptr, n, _fp(X))
        if rc != NO_ERROR:
            raise BwtError(rc, "mibwt_wp_features")
        return X

    def select_wp(self, dT_ptr: int, n: int):
        """((W, P), trained, predicted log-ms-per-byte of the NUM_CAND candidates)."""
        W, P = ctypes.c_int32(), ctypes.c_int32()
        pred = np.zeros(NUM_CAND, dtype=np.float32)
        rc = lib().mibwt_select_wp_device(self.handle, dT_ptr, n, ctypes.byref(W), ctypes.byref(P), _fp(pred))
        if rc < 0:
            raise BwtError(rc, "mibwt_select_wp_device")
        return (W.value, P.value), bool(rc), pred

    def mlp_grad(self, params: np.ndarray, X: np.ndarray, Y: np.ndarray):
        params = np.ascontiguousarray(params, dtype=np.float32)
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        g = np.zeros_like(params)
        loss = ctypes.c_float()
        rc = lib().mibwt_mlp_grad(self.handle, _fp(params), _fp(X), _fp(Y), len(X), _fp(g), ctypes.byref(loss))
        if rc != NO_ERROR:
            raise BwtError(rc, "mibwt_mlp_grad")
        return float(loss.value), g

    def mlp_predict(self, params: np.ndarray, X: np.ndarray) -> np.ndarray:
        params = np.ascontiguousarray(params, dtype=np.float32)
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((len(X), 3), dtype=np.float32)
        rc = lib().mibwt_mlp_predict(self.handle, _fp(params), _fp(X), len(X), _fp(Y))
        if rc != NO_ERROR:
            raise BwtError(rc, "mibwt_mlp_predict")
        return Y

    def mlp_train(self, X: np.ndarray, Y: np.ndarray, epochs: int = 200, batch: int = 512,
                  lr: float = 1e-3, seed: int = 1, out_path: str | None = None, members: int = 1) -> float:
        """Train the selector network (an ensemble of `members` networks when > 1), install it and
        write it to out_path; returns the final training loss."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        loss = ctypes.c_float()
        path = out_path.encode() if out_path else None
        if members == 1:
            rc = lib().mibwt_mlp_train(self.handle, _fp(X), _fp(Y), len(X), epochs, batch, lr, seed, path,
                                       ctypes.byref(loss))
        else:
            rc = lib().mibwt_mlp_train_ensemble(self.handle, _fp(X), _fp(Y), len(X), epochs, batch, lr, seed,
                                                members, path, ctypes.byref(loss))
        if rc != NO_ERROR:
            raise BwtError(rc, "mibwt_mlp_train")
        return float(loss.value)

    def stats(self) -> dict:
        st = PhaseStats()
        rc = lib().mibwt_last_phase_stats(self.handle, ctypes.byref(st))
        if rc != NO_ERROR:
            raise BwtError(rc, "mibwt_last_phase_stats")
        phases = []
        for i in range(st.num_phases):
            phases.append({"name": st.name[i].value.decode(), "ms": st.ms[i], "bytes": st.bytes[i]})
        out = {k: getattr(st, k) for k, _ in PhaseStats._fields_ if k not in ("name", "ms", "bytes")}
        out["phases"] = phases
        return out

    def set_kernel_timing(self, kernel: str | None):
        rc = lib().mibwt_set_kernel_timing(self.handle, kernel.encode() if kernel else None)
        if rc != NO_ERROR:
            raise BwtError(rc, "mibwt_set_kernel_timing")

    def time_kernel(self, name: str, reps: int = 20, nbytes: float = 0.0):
        ms, b = ctypes.c_double(), ctypes.c_double(nbytes)
        rc = lib().mibwt_time_kernel(self.handle, name.encode(), reps, ctypes.byref(ms), ctypes.byref(b))
        if rc != NO_ERROR:
            raise BwtError(rc, "mibwt_time_kernel")
        return ms.value, b.value


class DeviceBuffer:
    """A device allocation made through the library's own HIP runtime."""

    def __init__(self, nbytes: int):
        self.ptr = ctypes.c_void_p()
        self.nbytes = int(nbytes)
        rc = lib().mibwt_device_alloc(ctypes.byref(self.ptr), self.nbytes)
        if rc != NO_ERROR:
            raise BwtError(rc, "mibwt_device_alloc")

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    @property
    def address(self) -> int:
        return int(self.ptr.value)

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        rc = lib().mibwt_copy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 0)
        if rc != NO_ERROR:
            raise BwtError(rc, "mibwt_copy")

    def download(self, dtype=np.uint8, count: int | None = None) -> np.ndarray:
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize if count is None else count, dtype=dtype)
        rc = lib().mibwt_copy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, out.nbytes, 1)
        if rc != NO_ERROR:
            raise BwtError(rc, "mibwt_copy")
        return out

    def free(self):
        if self.ptr:
            lib().mibwt_device_free(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    """A non-blocking HIP stream of the library's runtime (for the ``stream`` arguments)."""

    def __init__(self):
        self.handle = ctypes.c_void_p()
        rc = lib().mibwt_stream_create(ctypes.byref(self.handle))
        if rc != NO_ERROR:
            raise BwtError(rc, "mibwt_stream_create")

    @property
    def address(self) -> int:
        return int(self.handle.value)

    def close(self):
        if self.handle:
            lib().mibwt_stream_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


NOALLOC_MAX_LENGTH = 1 << 20


def bwt_device_noalloc(dT_ptr: int, dO_ptr: int, n: int, W: int = 5, P: int = 16, stream: int | None = None) -> int:
    """mibwt_pfp_bwt_device_noalloc: the device build with fixed (W, P) that allocates nothing
    (reference format, n <= NOALLOC_MAX_LENGTH); returns the primary index."""
    rc = lib().mibwt_pfp_bwt_device_noalloc(dT_ptr, dO_ptr, n, W, P, stream)
    if rc < 0:
        raise BwtError(rc, "mibwt_pfp_bwt_device_noalloc")
    return int(rc)


TUNING_PATH = os.path.join(_HERE, "tuning_table.txt")


def tuning_shape(size: int) -> dict:
    """Launch shapes a build of `size` bytes uses (the autotuner's table or the defaults)."""
    sh = np.zeros(4, dtype=np.int32)
    rc = lib().mibwt_tuning_get(int(size), sh.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    if rc != NO_ERROR:
        raise BwtError(rc, "mibwt_tuning_get")
    return {"scan_sub": int(sh[0]), "sort_items": int(sh[1]), "chain_items": int(sh[2]), "key_bits": int(sh[3])}


def tuning_use(on: bool):
    lib().mibwt_tuning_use(1 if on else 0)


def synchronize():
    rc = lib().mibwt_synchronize()
    if rc != NO_ERROR:
        raise BwtError(rc, "mibwt_synchronize")


def bwt(T, method: str = "pfp", W: int = 5, P: int = 16):
    """(L, primary_index) in libcubwt format."""
    T = _as_text(T)
    with Storage(max(len(T), 16)) as s:
        return s.bwt(T, method, W, P)


def cubwt_to_reference(L: np.ndarray, index: int) -> np.ndarray:
    """O = L[1..i-1] ++ [L[0]] ++ L[i..n-1] (libcubwt.cu:3003-3015, without its i == n quirk)."""
    L = np.asarray(L, dtype=np.uint8)
    return np.concatenate([L[1:index], L[:1], L[index:]])


def reference_to_cubwt(O: np.ndarray, index: int) -> np.ndarray:
    O = np.asarray(O, dtype=np.uint8)
    return np.concatenate([O[index - 1:index], O[:index - 1], O[index:]])


def bwt_reference_format(T, method: str = "pfp", W: int = 5, P: int = 16) -> np.ndarray:
    L, idx = bwt(T, method, W, P)
    return cubwt_to_reference(L, idx)
