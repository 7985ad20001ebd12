"""CPU oracle of the learned-initial-guess network (TEST INFRASTRUCTURE ONLY).

Restates, in numpy float64 arithmetic, the forward pass and the gradient of
mean((y - t)^2) of the 16 -> 64 -> 64 -> 1 ReLU network the HIP kernels
evaluate, rounding to BF16 at the same points (inputs, weights, hidden
activations, back-propagated pre-activation gradients), so that only the
summation order of the FP32 accumulation differs.  The network has no
counterpart in the reference; it is pinned by finite differences of this
restatement (tests/test_oracle_mlp.py).
"""
import numpy as np

NF, NH = 16, 64
P_W1 = 0
P_B1 = P_W1 + NF * NH
P_W2 = P_B1 + NH
P_B2 = P_W2 + NH * NH
P_W3 = P_B2 + NH
P_B3 = P_W3 + NH
NPARAM = P_B3 + 1


def bf16_bits(x):
    """round-to-nearest-even float32 -> bf16 bit patterns (uint16)"""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def bf16(x):
    """value of x rounded to bf16, as float64"""
    b = bf16_bits(x).astype(np.uint32) << 16
    return b.view(np.float32).astype(np.float64)


def from_bits(bits):
    return (np.asarray(bits, np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(P):
    P = np.asarray(P, np.float64)
    return (P[P_W1:P_B1].reshape(NF, NH), P[P_B1:P_W2], P[P_W2:P_B2].reshape(NH, NH), P[P_B2:P_W3],
            P[P_W3:P_B3], P[P_B3])


def forward(P, Xbits, rounding=True):
    """rounding=False evaluates the same network in exact float64 arithmetic
    (used to check the back-propagation against finite differences)"""
    r = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    W1, b1, W2, b2, W3, b3 = unpack(P)
    X = from_bits(Xbits) if np.asarray(Xbits).dtype == np.uint16 else np.asarray(Xbits, np.float64)
    z1 = X @ r(W1) + b1
    h1 = r(np.maximum(z1, 0.0))
    z2 = h1 @ r(W2) + b2
    h2 = r(np.maximum(z2, 0.0))
    y = h2 @ r(W3) + b3
    return y, (X, h1, h2)


def gradient(P, Xbits, t, rounding=True):
    """d/dP of mean((y - t)^2); returns (grad[NPARAM], summed squared error)"""
    r = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    W1, b1, W2, b2, W3, b3 = unpack(P)
    y, (X, h1, h2) = forward(P, Xbits, rounding)
    n = y.size
    d = y - np.asarray(t, np.float64)
    g = 2.0 * d / n
    gb = r(g)
    gh2 = np.outer(gb, r(W3))
    gz2 = np.where(h2 > 0, gh2, 0.0)
    gZ2 = r(gz2)
    gh1 = gZ2 @ r(W2).T
    gz1 = np.where(h1 > 0, gh1, 0.0)
    gZ1 = r(gz1)
    grad = np.zeros(NPARAM)
    grad[P_W1:P_B1] = (X.T @ gZ1).ravel()
    grad[P_B1:P_W2] = gz1.sum(axis=0)
    grad[P_W2:P_B2] = (h1.T @ gZ2).ravel()
    grad[P_B2:P_W3] = gz2.sum(axis=0)
    grad[P_W3:P_B3] = h2.T @ g
    grad[P_B3] = g.sum()
    return grad, float((d * d).sum())


def loss(P, Xbits, t, rounding=True):
    y, _ = forward(P, Xbits, rounding)
    return float(np.mean((y - t) ** 2))
