"""Mixing plane on the GPU (k_mixing_plane) against the oracle's restatement
of CMesh::BndCondMixingPlane (source/BndCondMixingPlane.cpp:287-522), itself
checked bit for bit against the compiled reference in test_oracle_vs_ref.py.

Tolerances: FP64 with the reference's expression order and the reference's
j order in the pitch-wise sums; the library's other kernels hold 1e-12 for
one iteration and 1e-9 over a history (device pow and exp round differently
from libm in the last bit)."""
import numpy as np
import pytest

from fv3d import cases
from fv3d.solver import GpuSolver
import oracle as O

pytestmark = pytest.mark.gpu


def euler_case():
    c = cases.annular_passage(ni=6, nj=8, nk_blocks=(5, 6, 4), max_iter=6, seed=74, rpm=5000.0)
    return cases.with_mixing_plane(c, 2, 3)


def sa_case():
    c = cases.with_sa(cases.with_wall_shear(cases.annular_passage(ni=8, nj=6, nk_blocks=(5, 8, 4), seed=75,
                                                                  rpm=6000.0, blade_blocks=(1,), camber=0.1,
                                                                  viscous=True, len_ref=1e-3, max_iter=8)))
    return cases.with_mixing_plane(c, 1, 2)


def mg_case():
    c = cases.with_wall_shear(cases.annular_passage(ni=8, nj=8, nk_blocks=(6, 8, 6), seed=76, rpm=3000.0,
                                                    blade_blocks=(1,), camber=0.1, viscous=True, len_ref=1e-3,
                                                    max_iter=8))
    c.cfg.turbulence = 100
    c.cfg.use_multigrid, c.cfg.reuse_solution, c.cfg.coarse_iter, c.cfg.initial = "Y", "N", 2, "I"
    c.cfg.outlet_angle = (0.0, 0.2, 1.0)
    return cases.with_mixing_plane(c, 2, 3)


def stage_case():
    # rotor (8 cells pitchwise, 22 blades, 3000 rpm) and stator (5 cells, 30
    # blades, stationary): the two sides of the mixing plane differ in cell
    # count, pitch and rotation
    c = cases.rotor_stator_stage(ni=6, nj_rotor=8, nj_stator=5, rotor_nk=(4, 6, 4), stator_nk=(3, 7, 4), max_iter=8)
    c.cfg.use_multigrid = "N"
    return c


def stage_mg_case():
    return cases.rotor_stator_stage(ni=6, nj_rotor=8, nj_stator=6, rotor_nk=(6, 8, 6), stator_nk=(6, 10, 6),
                                    coarse_iter=2, max_iter=8)


@pytest.mark.parametrize("make", [euler_case, sa_case, stage_case], ids=["euler", "sa", "stage"])
def test_ghosts_after_bc(make):
    case = make()
    print(f"case mixing plane bc {make.__name__}", flush=True)
    g, o = GpuSolver(case), O.Oracle(case)
    g.call("bc_apply")
    o.call("bc")
    fields = ("cv", "pv", "ev") if case.cfg.turbulence == 200 else ("cv", "pv")
    for b in range(1, len(case.blocks) + 1):
        for f in fields:
            np.testing.assert_allclose(g.get(b, f), o.get(b, f), rtol=1e-13, atol=1e-15, err_msg=f"block {b} {f}")


@pytest.mark.parametrize("make", [euler_case, sa_case, mg_case, stage_case, stage_mg_case],
                         ids=["euler", "sa", "two-grid", "stage", "stage-two-grid"])
def test_history_matches_oracle(make):
    case = make()
    n = case.cfg.max_iter
    print(f"case mixing plane history {make.__name__}, {n} iterations", flush=True)
    g, o = GpuSolver(case), O.Oracle(case)
    hg, ho = g.iterate(n), o.iterate(n)
    fin = np.isfinite(ho)
    assert np.array_equal(fin, np.isfinite(hg))
    np.testing.assert_allclose(hg[fin], ho[fin], rtol=1e-9, atol=1e-14)
    for b in range(1, len(case.blocks) + 1):
        np.testing.assert_allclose(g.get(b, "cv"), o.get(b, "cv"), rtol=1e-9, atol=1e-12, err_msg=f"block {b}")
