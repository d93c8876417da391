"""Pins the CPU oracle of the learned-guess network (oracle/netoracle.c).

* Known answer: identity hidden layers and a one-tap output layer shift the
  first input channel by one node along x (tap convention dx + 3 dy + 9 dz at
  offset d-1, zero padding, weight layout [layer][tap][out][in]).
* Gradients: central differences of the FP64 loss (quant = 0) for weights of
  every layer and for biases.
* The BF16 mode stays close to the FP64 mode.
"""
import numpy as np
import pytest

L, T, C = 5, 27, 16


def _inputs(rng, nx, ny, nz, positive=False):
    N = nx * ny * nz
    f = (lambda: rng.random(N) + 0.1) if positive else (lambda: rng.standard_normal(N))
    x1, x2, src, vx, vy, vz, tgt = (f() for _ in range(7))
    mask = np.zeros(N, dtype=np.int32)
    mask[:nx] = 1  # a few Dirichlet nodes
    return x1, x2, src, vx, vy, vz, tgt, mask


def _identity_shift_params():
    w = np.zeros((L, T, C, C), np.float32)
    for l in range(L - 1):
        w[l, 13] = np.eye(C, dtype=np.float32)  # centre tap
    w[L - 1, 14, 0, 0] = 1.0  # tap dx=2, dy=1, dz=1 -> node x+1
    return w, np.zeros((L, C), np.float32)


def test_tap_convention_known_answer(po):
    nx, ny, nz = 6, 5, 4
    rng = np.random.default_rng(1)
    x1, x2, src, vx, vy, vz, tgt, mask = _inputs(rng, nx, ny, nz, positive=True)
    w, b = _identity_shift_params()
    scale = np.array([2.0, 1, 1, 1, 1, 1], np.float32)
    for quant in (False, True):
        _, _, _, out = po.net_grads(nx, ny, nz, x1, x2, src, vx, vy, vz, tgt, mask, scale, 1.0,
                                    w, b, quant=quant)
        v = (2.0 * x1).reshape(nz, ny, nx)
        if quant:
            v = v.astype(np.float32).view(np.uint32)
            v = ((v + 0x7FFF + ((v >> 16) & 1)) & 0xFFFF0000).view(np.float32).astype(np.float64)
        ref = np.zeros_like(v)
        ref[:, :, :-1] = v[:, :, 1:]
        assert np.array_equal(out.reshape(nz, ny, nx), ref), quant


def test_gradients_match_finite_differences(po):
    nx, ny, nz = 5, 4, 4
    rng = np.random.default_rng(2)
    x1, x2, src, vx, vy, vz, tgt, mask = _inputs(rng, nx, ny, nz)
    w = (rng.standard_normal((L, T, C, C)) * 0.15).astype(np.float32)
    b = (rng.standard_normal((L, C)) * 0.05).astype(np.float32)
    scale = np.array([1.0, 0.7, 1.3, 1, 1, 1], np.float32)
    args = (nx, ny, nz, x1, x2, src, vx, vy, vz, tgt, mask, scale, 0.8)
    loss, dw, db, _ = po.net_grads(*args, w, b, quant=False, extrap=0.5)
    dw = dw.reshape(L, T, C, C)
    db = db.reshape(L, C)
    h = 1e-5  # small enough that no ReLU changes sign
    checked = 0
    for l in range(L):
        for _ in range(6):
            t, co, ci = rng.integers(T), rng.integers(C), rng.integers(6 if l == 0 else C)
            if l == L - 1:
                co = 0
            wp, wm = w.copy(), w.copy()
            wp[l, t, co, ci] += h
            wm[l, t, co, ci] -= h
            lp = po.net_grads(*args, wp, b, quant=False, extrap=0.5)[0]
            lm = po.net_grads(*args, wm, b, quant=False, extrap=0.5)[0]
            fd = (lp - lm) / (float(wp[l, t, co, ci]) - float(wm[l, t, co, ci]))
            assert fd == pytest.approx(dw[l, t, co, ci], rel=2e-3, abs=1e-7), (l, t, co, ci)
            checked += 1
        co = 0 if l == L - 1 else rng.integers(C)
        bp, bm = b.copy(), b.copy()
        bp[l, co] += h
        bm[l, co] -= h
        lp = po.net_grads(*args, w, bp, quant=False, extrap=0.5)[0]
        lm = po.net_grads(*args, w, bm, quant=False, extrap=0.5)[0]
        fd = (lp - lm) / (float(bp[l, co]) - float(bm[l, co]))
        assert fd == pytest.approx(db[l, co], rel=2e-3, abs=1e-7), (l, co)
    assert checked == 30 and loss > 0


def test_bf16_mode_close_to_fp64(po):
    nx, ny, nz = 6, 5, 4
    rng = np.random.default_rng(3)
    x1, x2, src, vx, vy, vz, tgt, mask = _inputs(rng, nx, ny, nz)
    w = (rng.standard_normal((L, T, C, C)) * 0.15).astype(np.float32)
    b = (rng.standard_normal((L, C)) * 0.05).astype(np.float32)
    scale = np.ones(6, np.float32)
    args = (nx, ny, nz, x1, x2, src, vx, vy, vz, tgt, mask, scale, 1.0, w, b)
    l0, dw0, db0, o0 = po.net_grads(*args, quant=False)
    l1, dw1, db1, o1 = po.net_grads(*args, quant=True)
    assert abs(l1 - l0) <= 0.05 * l0
    assert np.linalg.norm(o1 - o0) <= 0.05 * np.linalg.norm(o0)
    assert np.linalg.norm(dw1 - dw0) <= 0.1 * np.linalg.norm(dw0)


def test_gpu_bars_catch_a_dropped_tap(po):
    """The GPU gradient test holds each layer's dW to 5e-3 relative L2, the bias
    gradient to 1e-3 and the prediction to 2e-3 of this oracle (BF16 roundings
    that fall on the other side of a rounding boundary from the FP64-summed
    oracle give 5e-5..2e-3 for dW, 2e-4 for the bias, 5e-4 for the prediction).
    A kernel that dropped one tap of one layer, or mirrored a tap, fails at
    least one of those bars: planted here in the oracle's weights, on the GPU
    test's inputs."""
    nx, ny, nz = 20, 17, 9
    N = nx * ny * nz
    rng = np.random.default_rng(2)
    vec = {k: rng.standard_normal(N) for k in ("x1", "x2", "src", "vx", "vy", "vz", "tgt")}
    mask = np.zeros(N, dtype=np.int32)
    rng = np.random.default_rng(3)
    w = (rng.standard_normal((L, T, C, C)) * 0.12).astype(np.float32)
    b = (rng.standard_normal((L, C)) * 0.05).astype(np.float32)
    scale = np.array([1.0, 0.8, 1.2, 1, 1, 1], np.float32)
    args = (nx, ny, nz, vec["x1"], vec["x2"], vec["src"], vec["vx"], vec["vy"], vec["vz"],
            vec["tgt"], mask, scale, 0.7)
    _, dw, db, out = po.net_grads(*args, w, b, quant=True, extrap=1.0)
    dw = dw.reshape(L, -1)

    def worst(w2):
        _, dw2, db2, out2 = po.net_grads(*args, w2, b, quant=True, extrap=1.0)
        dw2 = dw2.reshape(L, -1)
        r = max(np.linalg.norm(dw2[l] - dw[l]) / np.linalg.norm(dw[l]) / 5e-3 for l in range(L))
        r = max(r, np.linalg.norm(db2 - db) / np.linalg.norm(db) / 1e-3)
        return max(r, np.linalg.norm(out2 - out) / np.linalg.norm(out) / 2e-3)

    for layer in range(L):
        for t in (0, 13, 26):
            w2 = w.copy()
            w2[layer, t] = 0.0
            assert worst(w2) > 1.0, ("dropped", layer, t)
        w2 = w.copy()
        w2[layer] = w[layer, ::-1]  # taps mirrored (forward read as backward)
        assert worst(w2) > 1.0, ("mirrored", layer)
