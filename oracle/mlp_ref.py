"""Test infrastructure: CPU restatement of the MLP cooling surrogate
(forward, loss, backward).  Used only by tests to check the GPU kernels
(libplutomx pmx_mlp_*); the surrogate has no counterpart in the reference.

Network: 4 -> 64 -> 64 -> 64 -> 1, SiLU, loss L = sum_s (y_s - t_s)^2 / B.
Parameter vector (float32, 8705 entries): W1[64][4] b1[64] W2[64][64] b2[64]
W3[64][64] b3[64] w4[64] b4.

``rounded=True`` reproduces the GPU's mixed precision: weights, inputs,
activations and the back-propagated dZ operands rounded to BF16 (round to
nearest even), FP32 accumulation, FP32 biases.  ``rounded=False`` with
float64 is the exact function, which the tests pin against finite
differences.
"""
from __future__ import annotations

import numpy as np

NP = 8705
OFF = {"W1": 0, "b1": 256, "W2": 320, "b2": 4416, "W3": 4480, "b3": 8576, "w4": 8640,
       "b4": 8704}
HULL_LO = np.array([-16.0, 4.0, 0.0, -2.0])
HULL_HI = np.array([-10.0, 8.0, 6.0, 2.0])
YS = 2.0


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to the nearest BF16 (ties to even)."""
    x = np.asarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32)


def unpack(P: np.ndarray) -> dict:
    P = np.asarray(P)
    return {"W1": P[0:256].reshape(64, 4), "b1": P[256:320],
            "W2": P[320:4416].reshape(64, 64), "b2": P[4416:4480],
            "W3": P[4480:8576].reshape(64, 64), "b3": P[8576:8640],
            "w4": P[8640:8704], "b4": P[8704]}


def pack(g: dict) -> np.ndarray:
    return np.concatenate([g["W1"].ravel(), g["b1"], g["W2"].ravel(), g["b2"],
                           g["W3"].ravel(), g["b3"], g["w4"], np.atleast_1d(g["b4"])])


def normalize(X: np.ndarray, dtype) -> np.ndarray:
    c = 0.5 * (HULL_LO + HULL_HI)
    h = 0.5 * (HULL_HI - HULL_LO)
    return ((np.asarray(X, dtype=dtype) - c.astype(dtype)) / h.astype(dtype)).astype(dtype)


def _sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def forward(P, X, rounded: bool = True, dtype=np.float32, keep: bool = False):
    """Network output y (before the YS scale) for raw features X[n][4]."""
    r = bf16 if rounded else (lambda a: a)
    p = {k: (np.asarray(v, dtype=dtype) if k != "b4" else dtype(v)) for k, v in unpack(P).items()}
    Xb = r(normalize(X, dtype))
    Z1 = Xb @ r(p["W1"]).T + p["b1"]
    H1 = r(Z1 * _sig(Z1))
    Z2 = H1 @ r(p["W2"]).T + p["b2"]
    H2 = r(Z2 * _sig(Z2))
    Z3 = H2 @ r(p["W3"]).T + p["b3"]
    H3 = r(Z3 * _sig(Z3))
    y = H3 @ r(p["w4"]) + p["b4"]
    if keep:
        return y, dict(Xb=Xb, Z1=Z1, H1=H1, Z2=Z2, H2=H2, Z3=Z3, H3=H3, p=p)
    return y


def loss_grad(P, X, T, rounded: bool = True, dtype=np.float32):
    """(loss, gradient) of L = sum (y - T/YS)^2 / B over the batch, T the
    labels log10(T_final / T_0)."""
    r = bf16 if rounded else (lambda a: a)
    y, c = forward(P, X, rounded, dtype, keep=True)
    p = c["p"]
    B = len(T)
    res = y - (np.asarray(T, dtype=dtype) * dtype(1.0 / YS)).astype(dtype)
    loss = float(np.sum(res.astype(np.float64) ** 2) / B)
    dy = (2.0 * res / B).astype(dtype)

    def dsilu(z):
        s = _sig(z)
        return s * (1.0 + z * (1.0 - s))

    g = {}
    g["w4"] = c["H3"].T @ dy
    g["b4"] = dy.sum()
    dZ3 = dy[:, None] * r(p["w4"])[None, :] * dsilu(c["Z3"])
    g["b3"] = dZ3.sum(0)
    dZ3b = r(dZ3)
    g["W3"] = dZ3b.T @ c["H2"]
    dZ2 = (dZ3b @ r(p["W3"])) * dsilu(c["Z2"])
    g["b2"] = dZ2.sum(0)
    dZ2b = r(dZ2)
    g["W2"] = dZ2b.T @ c["H1"]
    dZ1 = (dZ2b @ r(p["W2"])) * dsilu(c["Z1"])
    g["b1"] = dZ1.sum(0)
    g["W1"] = r(dZ1).T @ c["Xb"]
    return loss, pack(g).astype(dtype)
