"""Tile shapes of the residual kernels and the autotuner: the shape changes
the schedule, never the result (every face is evaluated once and every cell
sums its faces in a fixed order), so any shape, and whatever the autotuner
picks, gives the default shape's bits.
"""
import numpy as np
import pytest

from fv3d import cases
from fv3d.solver import GpuSolver

pytestmark = pytest.mark.gpu

SHAPES = range(len(GpuSolver.TILE_SHAPES))


def sa_passage():
    c = cases.with_sa(cases.with_wall_shear(cases.annular_passage(ni=10, nj=6, nk_blocks=(5, 9, 6), seed=91,
                                                                  rpm=5000.0, blade_blocks=(1,), camber=0.1,
                                                                  viscous=True, len_ref=1e-3, max_iter=6)))
    return cases.turn_block(c, 1, 1)


@pytest.mark.parametrize("q", list(SHAPES))
def test_every_shape_gives_the_same_bits(q):
    case = sa_passage()
    print(f"case SA passage tiles {GpuSolver.TILE_SHAPES[q]}", flush=True)
    ref, g = GpuSolver(case), GpuSolver(case)
    g.set_tiles([q, q, q])
    hr, hg = ref.iterate(5), g.iterate(5, graph=True)
    np.testing.assert_array_equal(hr, hg)
    for b in range(1, len(case.blocks) + 1):
        for f in ("cv", "ev", "res", "resev"):
            np.testing.assert_array_equal(ref.get(b, f), g.get(b, f))


def test_autotune_keeps_the_bits():
    case = sa_passage()
    print("case SA passage autotune", flush=True)
    ref, g = GpuSolver(case), GpuSolver(case)
    hr = ref.iterate(6)
    h1 = g.iterate(2, graph=True)
    chosen = g.autotune(2)
    print(f"  chosen {[GpuSolver.conv_shape(chosen[0])] + [GpuSolver.TILE_SHAPES[s] for s in chosen[1:]]}")
    # the convective kernel may take any of its configurations, the others a tile shape
    assert 0 <= chosen[0] <= GpuSolver.CONV_FACES
    assert all(0 <= s < len(GpuSolver.TILE_SHAPES) for s in chosen[1:])
    h2 = g.iterate(4, graph=True)
    np.testing.assert_array_equal(hr, np.concatenate([h1, h2]))
    for b in range(1, len(case.blocks) + 1):
        np.testing.assert_array_equal(ref.get(b, "cv"), g.get(b, "cv"))
        np.testing.assert_array_equal(ref.get(b, "ev"), g.get(b, "ev"))


def test_autotune_skips_kernels_not_in_use():
    case = cases.bump_channel(ni=8, nj=8, nk=16, seed=92, max_iter=4)
    print("case Euler bump autotune", flush=True)
    g = GpuSolver(case)
    g.iterate(1)
    chosen = g.autotune(1)
    assert chosen[0] >= 0 and chosen[1] == -1 and chosen[2] == -1


def test_batch_with_mixed_shapes():
    cs = [cases.with_sa(cases.with_wall_shear(cases.annular_passage(ni=8, nj=6, nk_blocks=nk, seed=s, rpm=rpm,
                                                                   blade_blocks=(1,), camber=0.1, viscous=True,
                                                                   len_ref=1e-3, max_iter=5)))
          for s, nk, rpm in ((93, (5, 8, 4), 4000.0), (94, (4, 11, 6), 5500.0))]
    print("case SA batch, mixed tile shapes", flush=True)
    ref, g = GpuSolver(cs), GpuSolver(cs)
    g.set_tiles([4, 1, 2])
    np.testing.assert_array_equal(ref.iterate(5, graph=True), g.iterate(5, graph=True))
