"""What the learned (W,P) selector gains: on held-out texts (never trained on: other seeds than
any training text), the learned device build (selection included in the timing) against the
reference default W=5,P=16, same bytes out.  The learned build may not be slower than the fixed
one beyond 10% (timing noise and the selection pass); the gains are printed."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOLERANCE = 1.10


def best_ms(st, dT, dO, n, W, P, reps=4):
    st.bwt_device(dT.address, dO.address, n, "pfp", W, P)  # warm the buffers
    out = []
    for _ in range(reps):
        t0 = time.perf_counter()
        idx = st.bwt_device(dT.address, dO.address, n, "pfp", W, P)
        out.append((time.perf_counter() - t0) * 1000.0)
    return min(out), idx


@pytest.mark.timeout(600)
def test_learned_choice_gain_over_w5p16(mibwt):
    from mibwt import gen
    texts = {
        "dna_collection_128m": gen.repetitive_dna(1 << 20, 128, 0.0002 * 10 / 11, 0.0002 / 11, 10, seed=987654),
        "dna_collection_64m_1pct": gen.repetitive_dna(1 << 21, 32, 0.009, 0.001, 10, seed=13579),
        "markov_64m": gen.markov2_printable(64 << 20, seed=24680),
    }
    gains = {}
    for name, T in texts.items():
        n = len(T)
        dT = mibwt.DeviceBuffer.from_numpy(T)
        dO = mibwt.DeviceBuffer(n)
        with mibwt.Storage(n) as st:
            ms_fixed, idx_f = best_ms(st, dT, dO, n, 5, 16)
            O_fixed = dO.download()
            ms_learned, idx_l = best_ms(st, dT, dO, n, 0, 0)
            s = st.stats()
            assert idx_l == idx_f and (dO.download() == O_fixed).all()
        gains[name] = (ms_fixed / ms_learned, (int(s["W"]), int(s["P"])), ms_fixed, ms_learned)
        print(f"{name}: n={n} W=5,P=16 {ms_fixed:.2f} ms, learned W={s['W']},P={s['P']} {ms_learned:.2f} ms "
              f"(selection included): gain {ms_fixed / ms_learned:.3f}x")
        dT.free()
        dO.free()
        assert ms_learned <= TOLERANCE * ms_fixed, (name, gains[name])
    geo = float(np.exp(np.mean([np.log(g[0]) for g in gains.values()])))
    print(f"geometric-mean gain over W=5,P=16: {geo:.3f}x")
    assert geo >= 1.0
