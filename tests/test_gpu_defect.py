"""FP32 defect correction (precision mode FV3D_DEFECT_FP32) against the plain
FP64 path.

The LU-SGS operator (both sweeps) runs in FP32; the residual, the update and
the norms stay FP64.  So the residual kernel must be bitwise the plain one,
one iteration's dQ differs only at FP32 rounding, and the converged state is
the FP64 fixed point: the two runs, iterated until the density residual is at
round-off, agree to that level in every field (the errors are printed per
field)."""
import numpy as np
import pytest

from fv3d import cases
from fv3d.solver import GpuSolver

pytestmark = pytest.mark.gpu

FP64, FP32 = 0, 1


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("kind", ["euler", "bl_passage"])
def test_residual_bitwise_and_first_step(kind):
    if kind == "euler":
        case = cases.bump_channel(ni=19, nj=13, nk=29, seed=81, max_iter=3)
    else:
        case = cases.with_wall_shear(cases.annular_passage(ni=11, nj=13, nk_blocks=(9, 21, 7), seed=82, rpm=5000.0,
                                                           blade_blocks=(1,), camber=0.1, viscous=True,
                                                           len_ref=1e-3, max_iter=3))
        case.cfg.turbulence = 100
    print(f"case defect correction first step {kind}", flush=True)
    a = GpuSolver(case, precision=FP64)
    b = GpuSolver(case, precision=FP32)
    a.call("resflux")
    b.call("resflux")
    for blk in range(1, len(case.blocks) + 1):
        np.testing.assert_array_equal(a.get(blk, "res"), b.get(blk, "res"))
    a.call("lusgs")
    b.call("lusgs")
    for blk in range(1, len(case.blocks) + 1):
        e = relerr(b.get(blk, "dq"), a.get(blk, "dq"))
        print(f"  block {blk} dQ FP32 vs FP64 rel err {e:.3e}")
        assert 0.0 < e <= 1e-4


def test_converged_fields_equal_fp64_fixed_point():
    case = cases.bump_channel(ni=8, nj=6, nk=16, seed=1, max_iter=3000)
    print("case defect correction fixed point 16x6x8, 3000 iterations", flush=True)
    a = GpuSolver(case, precision=FP64)
    b = GpuSolver(case, precision=FP32)
    ha = a.iterate(3000, graph=True)
    hb = b.iterate(3000, graph=True)
    print(f"  final density residual FP64 {ha[-1, 1]:.3e} FP32 {hb[-1, 1]:.3e}")
    assert ha[-1, 1] < 1e-11 and hb[-1, 1] < 1e-11
    for name in ("cv", "pv"):
        x, y = a.get(1, name)[1:-1, 1:-1, 1:-1], b.get(1, name)[1:-1, 1:-1, 1:-1]
        for c in range(x.shape[-1]):
            d = np.abs(y[..., c] - x[..., c])
            sc = max(np.max(np.abs(x[..., c])), 1e-300)
            print(f"  {name}[{c}] max rel err {np.max(d) / sc:.3e}  L2 rel err {np.sqrt(np.mean(d ** 2)) / sc:.3e}")
        assert relerr(y, x) <= 1e-9
    # performance integrals of the converged states
    np.testing.assert_allclose(hb[-1, 2:6], ha[-1, 2:6], rtol=1e-9)
