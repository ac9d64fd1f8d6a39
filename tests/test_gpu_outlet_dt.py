"""The radial-equilibrium (axial) outlet and the global time step of the HIP
library against the CPU oracle.

Both oracle paths are pinned bit for bit to the compiled reference
(test_oracle_vs_ref.py: the axial outlet on a bump channel, and both on a
two-grid passage).  Tolerance as elsewhere: the device atan/sin/cos/sqrt of
the outlet's radial equilibrium and the device pow of the other terms round
differently from libm, so fields agree to 1e-12 of their magnitude after one
iteration and histories to 1e-10 over a few.
"""
import numpy as np
import pytest

from fv3d import cases
from fv3d.solver import GpuSolver
import oracle as O

pytestmark = pytest.mark.gpu


def relerr(a, b):
    scale = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale)


def passage(**kw):
    args = dict(ni=8, nj=6, nk_blocks=(4, 8, 4), seed=61, rpm=4000.0, max_iter=8)
    args.update(kw)
    return cases.annular_passage(**args)


def variant(name):
    if name == "axial_passage":
        c = passage()
        c.cfg.turbo_type = "A"
    elif name == "global_dt_bump":
        c = cases.bump_channel(ni=7, nj=6, nk=11, seed=62, max_iter=8)
        c.cfg.timestep = "G"
    elif name == "global_dt_axial_ns_passage":
        c = passage(seed=63, viscous=True, len_ref=1e-3)
        c.cfg.timestep, c.cfg.turbo_type = "G", "A"
    elif name == "two_grid_global_dt_axial":
        c = passage(ni=8, nj=8, nk_blocks=(6, 8, 6), seed=64)
        c.cfg.timestep, c.cfg.turbo_type = "G", "A"
        c.cfg.use_multigrid, c.cfg.reuse_solution, c.cfg.coarse_iter, c.cfg.initial = "Y", "N", 2, "U"
    else:
        raise KeyError(name)
    return c


@pytest.mark.parametrize("name", ["axial_passage", "global_dt_bump", "global_dt_axial_ns_passage"])
def test_first_iteration(name):
    case = variant(name)
    print(f"case {name} first iteration", flush=True)
    o, g = O.Oracle(case), GpuSolver(case)
    o.call("lusgs")
    g.call("lusgs")
    for b in range(1, len(case.blocks) + 1):
        for f in ("dt", "res", "dq", "cv", "pv", "dKP_cv", "dKP_pv"):
            a, r = g.get(b, f), o.get(b, f)
            assert relerr(a, r) <= 1e-12, f"{name} block {b} {f}: {relerr(a, r):.3e}"
    if case.cfg.timestep == "G":
        for b in range(1, len(case.blocks) + 1):
            dt = g.get(b, "dt")
            assert np.all(dt == dt.flat[0]), "one time step per block"


@pytest.mark.parametrize("name", ["axial_passage", "global_dt_bump", "global_dt_axial_ns_passage",
                                  "two_grid_global_dt_axial"])
def test_history(name):
    case = variant(name)
    n = case.cfg.max_iter
    print(f"case {name} history n={n}", flush=True)
    o, g = O.Oracle(case), GpuSolver(case)
    ho, hg = o.iterate(n), g.iterate(n)
    for col in range(1, 7):
        assert relerr(hg[:, col], ho[:, col]) <= 1e-10, f"{name} history col {col}"
    for b in range(1, len(case.blocks) + 1):
        assert relerr(g.get(b, "cv"), o.get(b, "cv")) <= 1e-10
    h2 = GpuSolver(case).iterate(n, graph=True)
    np.testing.assert_array_equal(hg, h2)
