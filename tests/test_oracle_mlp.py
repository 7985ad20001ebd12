"""Pins of the network oracle: its back-propagation against central finite
differences of its own loss (exact arithmetic mode), and the BF16 rounding
helper against the definition of round-to-nearest-even."""
import numpy as np

import mlp_oracle as M


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.14159265, 1e-30, 0.0], np.float32)
    got = M.bf16(x)
    # 8 significant bits: 1 + 2^-8 is a tie -> even (1.0); 1 + 3*2^-9 rounds up to 1 + 2^-7
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.0 + 2 ** -7
    assert got[3] == -2.5
    assert abs(got[4] - 3.140625) < 1e-12
    assert np.all(np.abs(got - x) <= np.abs(x) * 2 ** -8)


def test_backprop_matches_finite_differences():
    rng = np.random.default_rng(0)
    P = rng.standard_normal(M.NPARAM) * 0.3
    X = rng.standard_normal((40, M.NF))
    t = rng.standard_normal(40)
    g, _ = M.gradient(P, X, t, rounding=False)
    idx = rng.choice(M.NPARAM, 200, replace=False)
    idx = np.concatenate([idx, [M.P_B3, M.P_W3, M.P_B2, M.P_B1, M.P_W1]])
    h = 1e-6
    for k in idx:
        Pp = P.copy(); Pp[k] += h
        Pm = P.copy(); Pm[k] -= h
        fd = (M.loss(Pp, X, t, False) - M.loss(Pm, X, t, False)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])
