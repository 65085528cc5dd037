"""The CPU oracle of the warm-start network (oracle/cnn_oracle.hpp), pinned
without the GPU:

* its backward pass is the gradient of its loss: central finite differences in
  double precision on random parameters of every layer;
* the residual form: with layer-4 weights and bias zero the output is the
  centre of the input tile;
* linearity of layer 4: scaling W4 and b4 scales y - x;
* bf16 emulation changes results only at bf16 precision.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O

N_PARAMS = 14737
OFF = dict(W1=0, b1=432, W2=448, b2=7360, W3=7376, b3=14288, W4=14304, b4=14736)


def lib():
    L = O.load("oracle")
    f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
    L.orc_cnn_forward.argtypes = [f32p, f32p, C.c_int, C.c_int, C.c_int, f32p]
    L.orc_cnn_grad.argtypes = [f32p, f32p, f32p, C.c_int, C.c_int, C.c_int, f64p]
    L.orc_cnn_grad.restype = C.c_double
    return L


def params(seed, scale_last=0.05):
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS, np.float32)
    p[OFF["W1"]:OFF["b1"]] = rng.normal(0, np.sqrt(2 / 27), 432)
    p[OFF["b1"]:OFF["W2"]] = rng.normal(0, 0.05, 16)
    p[OFF["W2"]:OFF["b2"]] = rng.normal(0, np.sqrt(2 / 432), 6912)
    p[OFF["b2"]:OFF["W3"]] = rng.normal(0, 0.05, 16)
    p[OFF["W3"]:OFF["b3"]] = rng.normal(0, np.sqrt(2 / 432), 6912)
    p[OFF["b3"]:OFF["W4"]] = rng.normal(0, 0.05, 16)
    p[OFF["W4"]:OFF["b4"]] = rng.normal(0, scale_last, 432)
    p[OFF["b4"]] = 0.01
    return p


def data(seed, T, S):
    rng = np.random.default_rng(seed)
    x = rng.random((T, (S + 8) ** 3), dtype=np.float32) * 2.0
    t = rng.random((T, S ** 3), dtype=np.float32) * 2.0
    return x, t


def loss_grad(L, p, x, t, T, S, emulate=0):
    g = np.zeros(N_PARAMS, np.float64)
    loss = L.orc_cnn_grad(p, x.reshape(-1), t.reshape(-1), T, S, emulate, g)
    return loss, g


def test_gradient_matches_finite_differences():
    L = lib()
    T, S = 2, 2
    p = params(0, scale_last=0.3)
    x, t = data(1, T, S)
    loss, g = loss_grad(L, p, x, t, T, S)
    rng = np.random.default_rng(2)
    picks = []
    for name, (a, b) in {"W1": (0, 432), "b1": (432, 448), "W2": (448, 7360), "b2": (7360, 7376),
                         "W3": (7376, 14288), "b3": (14288, 14304), "W4": (14304, 14736), "b4": (14736, 14737)}.items():
        picks += [int(i) for i in rng.integers(a, b, size=min(6, b - a))]
    checked = 0
    for i in picks:
        if abs(g[i]) < 1e-7:
            continue
        errs = []
        for rel_h in (1e-4, 1e-5):   # a step may cross a ReLU kink; one of two must agree
            h = rel_h * max(1.0, abs(float(p[i])))
            pp, pm = p.copy(), p.copy()
            pp[i] += np.float32(h)
            pm[i] -= np.float32(h)
            hh = (float(pp[i]) - float(pm[i])) / 2
            lp, _ = loss_grad(L, pp, x, t, T, S)
            lm, _ = loss_grad(L, pm, x, t, T, S)
            errs.append(abs((lp - lm) / (2 * hh) - g[i]))
        assert min(errs) <= 1e-5 * abs(g[i]) + 1e-8, (i, errs, g[i])
        checked += 1
    assert checked >= 30


def test_residual_identity():
    L = lib()
    T, S = 1, 3
    p = params(3)
    p[OFF["W4"]:] = 0
    x, _ = data(4, T, S)
    y = np.zeros(S ** 3, np.float32)
    L.orc_cnn_forward(p, x.reshape(-1), T, S, 0, y)
    centre = x.reshape(S + 8, S + 8, S + 8)[4:4 + S, 4:4 + S, 4:4 + S].reshape(-1)
    assert np.array_equal(y, centre)


def test_last_layer_linear():
    L = lib()
    T, S = 1, 3
    p = params(5)
    x, _ = data(6, T, S)
    y1 = np.zeros(S ** 3, np.float32)
    y2 = np.zeros(S ** 3, np.float32)
    L.orc_cnn_forward(p, x.reshape(-1), T, S, 0, y1)
    q = p.copy()
    q[OFF["W4"]:] *= 2
    L.orc_cnn_forward(q, x.reshape(-1), T, S, 0, y2)
    c = x.reshape(S + 8, S + 8, S + 8)[4:4 + S, 4:4 + S, 4:4 + S].reshape(-1).astype(np.float64)
    np.testing.assert_allclose(y2 - c, 2 * (y1 - c), rtol=1e-5, atol=1e-6)


def test_bf16_emulation_close():
    L = lib()
    T, S = 2, 3
    p = params(7)
    x, t = data(8, T, S)
    y0 = np.zeros(T * S ** 3, np.float32)
    y1 = np.zeros(T * S ** 3, np.float32)
    L.orc_cnn_forward(p, x.reshape(-1), T, S, 0, y0)
    L.orc_cnn_forward(p, x.reshape(-1), T, S, 1, y1)
    d = np.abs(y1 - y0).max() / np.abs(y0).max()
    assert 0 < d < 3e-2
    _, g0 = loss_grad(L, p, x, t, T, S, 0)
    _, g1 = loss_grad(L, p, x, t, T, S, 1)
    assert np.linalg.norm(g1 - g0) / np.linalg.norm(g0) < 5e-2
