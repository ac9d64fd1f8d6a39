"""Converged-field error of both GPU forms against the reference's FP64
converged state.

tests/golden/c0_converged.npz holds C0 (the 96x32x32 bump channel, Euler, Roe
+ MUSCL3 primitive, LU-SGS local dt at CFL 5, single grid) iterated to
convergence by the C restatement, which is bit for bit the compiled reference
(tests/golden/gen_converged_c0.py).  Plain FP64 and the FP32-operator form run
the same number of iterations from the same start on the GPU; every
conservative field is compared cell by cell (max and L2 error relative to the
field's scale).  Plain FP64 differs from the reference per
iteration only in the last bit (device pow and sqrt), the FP32 form only in
its implicit operator, so both must land on the reference's fixed point."""
import json
import os

import numpy as np
import pytest

from fv3d import cases
from fv3d.solver import GpuSolver

pytestmark = pytest.mark.gpu

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c0_converged.npz")
NAMES = ("rho", "rho u", "rho v", "rho w", "rho E")


def field_errors(cv, ref):
    """Max and L2 error of each conservative field relative to its scale: the
    largest density, the largest momentum magnitude |rho V| for the three
    momentum components (the channel is spanwise uniform, so rho u alone is
    round-off-sized and no scale of its own), the largest total energy."""
    mom = np.sqrt((ref[..., 1:4] ** 2).sum(axis=-1)).max()
    scales = (np.abs(ref[..., 0]).max(), mom, mom, mom, np.abs(ref[..., 4]).max())
    out = {}
    for q, name in enumerate(NAMES):
        d = cv[..., q] - ref[..., q]
        out[name] = (float(np.abs(d).max() / scales[q]), float(np.sqrt(np.mean(d * d)) / scales[q]))
    return out


# C0 on one grid does not settle below a drop of about 2e-7: the residual
# wanders between 1e-7 and 1e-6 of its first value from about 21,000
# iterations on (the MUSCL3 limiter switching near the bump), and two runs
# whose arithmetic differs in the last bit sit at different points of that
# wandering after the same iteration count.  The bar is set by it.
@pytest.mark.parametrize("precision,bar", [(0, 1e-5), (1, 1e-5)], ids=["fp64", "fp32-operator"])
def test_converged_c0_against_reference_state(precision, bar):
    z = np.load(FIX)
    meta = json.loads(bytes(z["meta"]).decode())
    n = int(meta["iterations"])
    case = cases.bump_channel(max_iter=n + 10, **meta["kwargs"])
    print(f"C0 converged, precision {precision}, {n} iterations (reference drop {meta['drop']:.2e})", flush=True)
    g = GpuSolver(case, precision=precision)
    done = 0
    while done < n:
        k = min(5000, n - done)
        h = g.iterate(k, graph=True)
        done += k
    cv = g.get(1, "cv")[1:-1, 1:-1, 1:-1]
    g.close()
    err = field_errors(cv, z["cv"])
    for name, (mx, l2) in err.items():
        print(f"  {name:6s} max {mx:.2e}  L2 {l2:.2e}", flush=True)
    assert all(mx <= bar for mx, _ in err.values())
