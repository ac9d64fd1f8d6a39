"""Helpers for the learned eddy-viscosity closure (fv3d_mlp_* of libfv3d)."""
from __future__ import annotations

import numpy as np

N_IN, N_HIDDEN, N_PARAMS = 6, 64, 8833


def init_params(seed: int = 0) -> np.ndarray:
    """Glorot-uniform weights and zero biases in the library's parameter
    layout: W0[64][6] b0[64] W1[64][64] b1[64] W2[64][64] b2[64] W3[64] b3."""
    rng = np.random.default_rng(seed)

    def glorot(n_out, n_in):
        lim = np.sqrt(6.0 / (n_in + n_out))
        return rng.uniform(-lim, lim, size=(n_out, n_in))

    h = N_HIDDEN
    p = np.concatenate([glorot(h, N_IN).ravel(), np.zeros(h), glorot(h, h).ravel(), np.zeros(h),
                        glorot(h, h).ravel(), np.zeros(h), glorot(1, h)[0], [0.0]]).astype(np.float32)
    assert p.size == N_PARAMS
    return p


def gate_bounds(x: np.ndarray, t: np.ndarray, out_factor: float = 1.5) -> np.ndarray:
    """Gate bounds of a training set: min and max of every feature, then the
    largest output the closure may give (out_factor times the largest target)."""
    return np.concatenate([np.stack([x.min(0), x.max(0)], 1).ravel(), [t.max() * out_factor]]).astype(np.float32)
