"""The closure network's CPU restatement (oracle/mlp_oracle.py), pinned by
mathematics: its manual back-propagation against central finite differences
of the float64 loss, and the BF16 rounding against exact bit patterns."""
import numpy as np

import mlp_oracle as M


def test_bf16_rounding_is_nearest_even():
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa (1) wins;
    # 1 + 3*2^-8 ties between 1 + 2^-7 and 1 + 2^-6: even mantissa (1 + 2^-6) wins
    x = np.array([1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -8 + 2.0 ** -20, -2.5], dtype=np.float32)
    np.testing.assert_array_equal(M.bf16(x), [1.0, 1.0 + 2.0 ** -6, 1.0 + 2.0 ** -7, -2.5])


def test_backprop_matches_finite_differences():
    rng = np.random.default_rng(5)
    P = M.init_params(3).astype(np.float64)
    P += rng.normal(0, 0.05, P.shape)          # nonzero biases too
    x = rng.normal(0, 1, (40, M.NIN))
    t = rng.uniform(0, 2, 40)
    g = M.backward(P, x.astype(np.float32), t, quantize=False)
    # a sample of parameters from every tensor
    idx = np.concatenate([rng.choice(np.arange(o, o + s), min(s, 12), replace=False) for o, s in M.OFF.values()])
    for p in idx:
        h = 1e-6
        Pp, Pm = P.copy(), P.copy()
        Pp[p] += h
        Pm[p] -= h
        fd = (M.loss(Pp, x.astype(np.float32), t) - M.loss(Pm, x.astype(np.float32), t)) / (2 * h)
        assert abs(fd - g[p]) <= 1e-6 + 1e-5 * abs(fd), (p, fd, g[p])


def test_quantised_forward_stays_close():
    rng = np.random.default_rng(6)
    P = M.init_params(4)
    x = rng.normal(0, 1, (256, M.NIN)).astype(np.float32)
    yq, _, _ = M.forward(P, x, True)
    y, _, _ = M.forward(P, x, False)
    assert np.max(np.abs(yq - y)) < 0.05 * (1 + np.max(np.abs(y)))
