"""The (W,P) selector's MLP: the fp64 reference is pinned by finite differences; the device
forward/backward kernels are checked against it; training lowers the loss; the selector's
choice keeps the BWT bit-exact."""
import numpy as np
import pytest

import oracle
from oracle import mlp


def bf16_params(rng, scale=0.2):
    p = rng.normal(0, scale, mlp.COUNT).astype(np.float32)
    return mlp.bf16_round(p)


def test_reference_gradient_finite_differences():
    rng = np.random.default_rng(0)
    p = bf16_params(rng).astype(np.float64)
    X = rng.normal(0, 1, (7, 64))
    Y = rng.normal(0, 1, (7, 3))
    loss, g = mlp.loss_grad(p, X, Y)
    # biases are used unrounded: central differences on them are exact up to O(h^2)
    off = 128 * 64
    for idx in [off + 3, off + 77, off + 128 + 128 * 128 + 5, mlp.COUNT - 1, mlp.COUNT - 3]:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[idx] += h
        pm[idx] -= h
        fd = (mlp.loss_grad(pp, X, Y)[0] - mlp.loss_grad(pm, X, Y)[0]) / (2 * h)
        assert abs(fd - g[idx]) < 1e-6 * max(1, abs(fd))
    # weight gradients: compare against differences of the unrounded network (linear in W)
    def loss_unrounded(q):
        d = mlp.unpack(q)
        a1 = X @ d["W1"].T + d["b1"]
        h1 = np.maximum(a1, 0)
        a2 = h1 @ d["W2"].T + d["b2"]
        h2 = np.maximum(a2, 0)
        y = h2 @ d["W3"].T + d["b3"]
        return ((y - Y) ** 2).mean()
    for idx in [5, 4000, off + 128 + 999, off + 128 + 128 * 128 + 128 + 200]:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[idx] += h
        pm[idx] -= h
        fd = (loss_unrounded(pp) - loss_unrounded(pm)) / (2 * h)
        assert abs(fd - g[idx]) < 1e-5 * max(1, abs(fd))


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.00390625, 1.0078125, -2.5, 3.14159], dtype=np.float32)
    r = mlp.bf16_round(x)
    assert r[0] == 1.0 and r[2] == 1.0078125
    assert r[1] == 1.0  # halfway case rounds to the even mantissa
    assert abs(r[4] - 3.140625) < 1e-7


@pytest.mark.gpu
def test_device_gradient_matches_reference(storage):
    rng = np.random.default_rng(1)
    p = bf16_params(rng)
    X = rng.normal(0, 1, (300, 64)).astype(np.float32)
    Y = rng.normal(0, 1, (300, 3)).astype(np.float32)
    loss_ref, g_ref = mlp.loss_grad(p, X, Y)
    loss, g = storage.mlp_grad(p, X, Y)
    assert abs(loss - loss_ref) < 1e-4 * max(1, loss_ref)
    scale = np.abs(g_ref).max()
    assert np.abs(g - g_ref).max() < 1e-4 * scale
    y_ref, _ = mlp.forward(p, X)
    y = storage.mlp_predict(p, X)
    assert np.abs(y - y_ref).max() < 1e-3 * max(1, np.abs(y_ref).max())


@pytest.mark.gpu
def test_training_fits_a_smooth_target(storage, tmp_path):
    rng = np.random.default_rng(2)
    X = rng.normal(0, 1, (4000, 64)).astype(np.float32)
    Y = np.stack([np.sin(X[:, 0]) + 0.5 * X[:, 1], X[:, 2] * X[:, 3], np.abs(X[:, 4])], 1).astype(np.float32)
    path = str(tmp_path / "m.bin")
    loss = storage.mlp_train(X, Y, epochs=60, batch=256, lr=2e-3, seed=3, out_path=path)
    assert loss < 0.2  # normalised targets start near 1.0
    import os
    assert os.path.getsize(path) > 4 * mlp.COUNT


@pytest.mark.gpu
def test_ensemble_training_writes_members(mibwt, storage, tmp_path):
    # three members from consecutive seeds share one normalisation; each fits the target and
    # they differ (different initialisations and shuffles)
    rng = np.random.default_rng(4)
    X = rng.normal(0, 1, (3000, 64)).astype(np.float32)
    Y = np.stack([np.sin(X[:, 0]), X[:, 2] * X[:, 3], np.abs(X[:, 4])], 1).astype(np.float32)
    path = str(tmp_path / "e.bin")
    loss = storage.mlp_train(X, Y, epochs=40, batch=256, lr=2e-3, seed=5, out_path=path, members=3)
    assert loss < 0.3
    m = mibwt.read_selector_model(path)
    assert m["members"] == 3 and m["params"].shape == (3, mlp.COUNT)
    assert not np.array_equal(m["params"][0], m["params"][1])
    Xn = (X[:200] - m["xmean"]) / m["xstd"]
    for p in m["params"]:
        y = storage.mlp_predict(p, Xn) * m["ystd"] + m["ymean"]
        assert np.corrcoef(y[:, 0], Y[:200, 0])[0, 1] > 0.9


@pytest.mark.gpu
def test_selector_choice_keeps_bwt_exact(mibwt):
    from mibwt import gen
    import os
    T = gen.repetitive_dna(1 << 18, 8, 0.001, 0.0001, 10, seed=9)
    dT = mibwt.DeviceBuffer.from_numpy(T)
    # the shipped model (a model trained earlier in this process would otherwise be active)
    path = os.environ.get("MIBWT_SELECTOR") or os.path.join(os.path.dirname(mibwt.__file__),
                                                            "selector_model.bin")
    if os.path.exists(path):
        mibwt.selector_load(path)
    with mibwt.Storage(len(T)) as st:
        (W, P), trained, pred = st.select_wp(dT.address, len(T))
        assert (W, P) in mibwt.wp_grid() and mibwt.wp_grid()[-1] == (5, 16)
        X = st.wp_features(dT.address, len(T))
        assert np.isfinite(X).all() and X.shape == (41, 64)
        if trained:
            # the device-side choice equals the host path: raw features, normalisation from
            # the model file, plain MLP inference of every member, mean, smoothing along P,
            # argmin
            m = mibwt.read_selector_model(path)
            Xn = (X - m["xmean"]) / m["xstd"]
            t = np.mean([st.mlp_predict(p, Xn)[:, 2] for p in m["params"]], axis=0)
            t = mibwt.smooth_along_p(t * m["ystd"][2] + m["ymean"][2])
            assert np.allclose(t, pred, rtol=0, atol=1e-4)
            assert mibwt.wp_grid()[int(np.argmin(pred))] == (W, P)
        L, idx = st.bwt(T, "cubwt")
    # the host-text entry makes the same choice
    assert mibwt.select_wp(T) == ((W, P), trained)
    O, oidx = oracle.bwt(T)
    assert idx == oidx and (mibwt.cubwt_to_reference(L, idx) == O).all()
    dT.free()


def sample_features(T):
    """CPU restatement of the sampled selector statistics: 64 chunks of 16384 bytes at stride
    n/64 rounded down to a multiple of 16; per grid candidate the fraction of sampled window ends p >= W whose exact window
    value is divisible by P (the hash rule of PFPk:468-475); order-0 entropy of the sample."""
    n = len(T)
    chunk, chunks = 16384, 64
    stride = (n // chunks) & ~15 if n > chunk * chunks else chunk
    pos = np.concatenate([np.arange(c * stride, min(c * stride + chunk, n)) for c in range(chunks)
                          if c * stride < n])
    t = T.astype(np.int64)
    out = {}
    for W in (5, 10, 20, 35):
        ok = pos[pos >= W]
        v = np.zeros(len(ok), dtype=np.int64)  # exact: 255 * (2^35 - 1) < 2^63
        for k in range(W):
            v = v * 2 + t[ok - W + 1 + k]
        for P in (11, 13, 17, 23, 31, 41, 53, 61, 71, 81) + ((16,) if W == 5 else ()):
            out[(W, P)] = int(np.sum(v % P == 0))
    hist = np.bincount(T[pos], minlength=256)
    p = hist[hist > 0] / len(pos)
    return out, len(pos), float(-(p * np.log2(p)).sum())


def _mix64(h):
    h = h ^ (h >> np.uint64(33))
    h = h * np.uint64(0xFF51AFD7ED558CCD)
    h = h ^ (h >> np.uint64(33))
    h = h * np.uint64(0xC4CEB9FE1A85EC53)
    return h ^ (h >> np.uint64(33))


def sketch_counts(T):
    """CPU restatement of the whole-text sketch: positions p with p + 64 <= n whose 8-byte window
    (lo = bytes 0..3, hi = bytes 4..7, little endian) has ((lo ^ hi * 0x85EBCA6B) * 0x9E3779B1)
    >> 22 == 0 (32-bit arithmetic); each contributes the top 32 bits of a chained mix of its
    16 and 64 bytes; returns (selected, distinct 16-byte, distinct 64-byte fingerprints)."""
    n = len(T)
    t = T.astype(np.uint64)

    def word(off, nbytes, pos):
        v = np.zeros(len(pos), dtype=np.uint64)
        for b in range(nbytes):
            v |= t[pos + off + b] << np.uint64(8 * b)
        return v

    pos = np.arange(0, n - 63)
    lo = word(0, 4, pos).astype(np.uint32)
    hi = word(4, 4, pos).astype(np.uint32)
    with np.errstate(over="ignore"):
        h = (lo ^ (hi * np.uint32(0x85EBCA6B))) * np.uint32(0x9E3779B1)
        sel = pos[(h >> np.uint32(22)) == 0]
        f = np.full(len(sel), 0x9E3779B97F4A7C15, dtype=np.uint64)
        f16 = None
        for k in range(8):
            f = _mix64(f ^ word(8 * k, 8, sel))
            if k == 1:
                f16 = f >> np.uint64(32)
    f64 = f >> np.uint64(32)
    return len(sel), len(np.unique(f16)), len(np.unique(f64))


def test_smoothing_along_p(mibwt):
    # every grid candidate averages with its P-neighbours of the same W; the default candidate
    # after the grid is left alone
    t = np.arange(41, dtype=np.float64) ** 2
    s = mibwt.smooth_along_p(t)
    assert s[0] == (0 + 1) / 2 and s[9] == (64 + 81) / 2 and s[10] == (100 + 121) / 2
    assert s[5] == (16 + 25 + 36) / 3 and s[40] == 1600


def test_sketch_restatement_separates_repetitive_from_random():
    # the restatement itself: a text made of 16 copies of one 64 KiB block has about 1/16 as
    # many distinct fingerprints as selected positions; a random text has all distinct
    rng = np.random.default_rng(5)
    block = rng.integers(0, 256, size=1 << 16, dtype=np.uint8)
    m, d16, d64 = sketch_counts(np.tile(block, 16))
    assert m > 500 and abs(d64 / m - 1 / 16) < 0.01 and d16 <= d64
    m, d16, d64 = sketch_counts(rng.integers(0, 256, size=1 << 20, dtype=np.uint8))
    assert m > 500 and d16 == m and d64 == m
    # selection rate is about 1 / 1024 on random bytes
    assert 0.7 < m * 1024 / (1 << 20) < 1.3


@pytest.mark.gpu
def test_sampled_features_match_cpu_restatement(mibwt):
    from mibwt import gen
    T = gen.repetitive_dna(1 << 18, 12, 0.001, 0.0001, 10, seed=21)  # 3 MiB: 64 strided chunks
    trig, npos, H0 = sample_features(T)
    dT = mibwt.DeviceBuffer.from_numpy(T)
    with mibwt.Storage(len(T)) as st:
        X = st.wp_features(dT.address, len(T))
        X2 = st.wp_features(dT.address, len(T))  # deterministic, tables reused across calls
    dT.free()
    assert (X == X2).all()
    grid = mibwt.wp_grid()
    for c, (W, P) in enumerate(grid):
        dens = (trig[(W, P)] + 1.0) / (npos + 1.0)
        assert abs(X[c, 16] - dens) <= 1e-6 * dens, (W, P)
    assert abs(X[0, 2] - H0 / 8.0) < 1e-5
    # repetitiveness: 12 copies of one base leave few distinct 64-byte fingerprints
    assert 0.0 < X[0, 5] < 0.5 and X[0, 4] <= X[0, 5] + 1e-6
    m, d16, d64 = sketch_counts(T)
    assert abs(X[0, 4] - d16 / m) <= 1e-6 and abs(X[0, 5] - d64 / m) <= 1e-6, (m, d16, d64)
