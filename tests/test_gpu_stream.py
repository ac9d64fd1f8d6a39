"""The streaming convective residual (k_conv_stream, convective configurations
5-9) and the face-parallel one (k_conv_faces + k_conv_sum, configuration 10)
against the per-tile kernel (k_convective, configurations 0-4): all evaluate every
face once with the same face code and sum each cell's six faces in the same
order, so the residual must agree bit for bit, for every reconstruction
variable, both flux schemes, rotating blocks with rotating walls, partial
tiles, multi-block passages with blade walls and turned blocks, and every
streaming configuration the autotuner may pick."""
import numpy as np
import pytest

from fv3d import cases
from fv3d.solver import GpuSolver

pytestmark = pytest.mark.gpu

STREAM0 = 5   # first streaming configuration


def _variants():
    out = {}
    out["bump_prim_partial"] = cases.bump_channel(ni=19, nj=13, nk=29, seed=7, max_iter=3)
    out["bump_cons_sw_nnd"] = cases.bump_channel(ni=9, nj=17, nk=21, seed=8, flux=200, recon_scheme=100,
                                                 recon_var=100, max_iter=3)
    out["bump_char_muscl2o"] = cases.bump_channel(ni=11, nj=7, nk=35, seed=9, recon_scheme=400, recon_var=300,
                                                  max_iter=3)
    c = cases.bump_channel(ni=6, nj=21, nk=18, seed=10, max_iter=3)
    c.nodes[0] = c.nodes[0] + np.array([1.0, 0.3, 0.0])
    c.blocks[0].rpm, c.blocks[0].nblade = 3000.0, 22
    for p in c.patches:
        if p.kind == "wall" and p.face in (3, 1):
            p.rotate = 1
    out["bump_rotating"] = c
    out["passage_blade_turned"] = cases.turn_block(
        cases.with_wall_shear(cases.annular_passage(ni=10, nj=18, nk_blocks=(6, 34, 9), seed=11, rpm=6000.0,
                                                    blade_blocks=(1,), camber=0.1, viscous=True, len_ref=1e-3,
                                                    max_iter=3)), 2, 1)
    return out


VARIANTS = _variants()


@pytest.mark.parametrize("name", list(VARIANTS))
def test_stream_equals_tiled(name):
    case = VARIANTS[name]
    print(f"case {name}", flush=True)
    a = GpuSolver(case)
    a.set_tiles((STREAM0, 0, 0))
    b = GpuSolver(case)
    for g in (a, b):
        g.call("bc_apply")
        g.call("convective")
    for blk in range(1, len(case.blocks) + 1):
        np.testing.assert_array_equal(a.get(blk, "res"), b.get(blk, "res"), err_msg=f"{name} block {blk}")
    # a few iterations: identical histories and states
    ha, hb = a.iterate(3), b.iterate(3)
    np.testing.assert_array_equal(ha, hb)
    for blk in range(1, len(case.blocks) + 1):
        np.testing.assert_array_equal(a.get(blk, "cv"), b.get(blk, "cv"))
    a.close()
    b.close()


@pytest.mark.parametrize("q", [0, 1, 2, 3, 4])
def test_stream_configurations(q):
    case = VARIANTS["passage_blade_turned"]
    print(f"stream configuration {q}", flush=True)
    a = GpuSolver(case)
    b = GpuSolver(case)
    a.set_tiles((STREAM0 + q, 0, 0))
    for g in (a, b):
        g.call("bc_apply")
        g.call("convective")
    for blk in range(1, len(case.blocks) + 1):
        np.testing.assert_array_equal(a.get(blk, "res"), b.get(blk, "res"))
    a.close()
    b.close()


@pytest.mark.parametrize("name", list(VARIANTS))
def test_face_form_equals_tiled(name):
    case = VARIANTS[name]
    print(f"case {name}, face-parallel form", flush=True)
    a = GpuSolver(case)
    a.set_tiles((GpuSolver.CONV_FACES, 0, 0))
    b = GpuSolver(case)
    for g in (a, b):
        g.call("bc_apply")
        g.call("convective")
    for blk in range(1, len(case.blocks) + 1):
        np.testing.assert_array_equal(a.get(blk, "res"), b.get(blk, "res"), err_msg=f"{name} block {blk}")
    ha, hb = a.iterate(3), b.iterate(3)
    np.testing.assert_array_equal(ha, hb)
    for blk in range(1, len(case.blocks) + 1):
        np.testing.assert_array_equal(a.get(blk, "cv"), b.get(blk, "cv"))
    a.close()
    b.close()
