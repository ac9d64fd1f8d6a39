"""N-level V-cycle (cfg.mg_levels 3-5): the reference's two-grid cycle
(source/Solver.cpp:47-140, two levels at most by source/UserInput.cpp:1057-1058)
extended to more levels with full approximation storage below the first coarse
level.  Checks: the level count follows the block sizes, the two-level path is
the unchanged reference cycle, graph replay equals eager execution, and the
V-cycle converges to a fixed point of the fine level's own iteration (the fine
residual of its converged state is round-off and plain FP64 single-grid LU-SGS
started there does not move it) in both the FP64 and the FP32-operator forms."""
import numpy as np
import pytest

from fv3d import cases
from fv3d.solver import GpuSolver

pytestmark = pytest.mark.gpu


def bump(levels, ni=24, nj=24, nk=48, coarse_iter=4, seed=5):
    c = cases.bump_channel(ni=ni, nj=nj, nk=nk, seed=seed, max_iter=100000)
    c.cfg.use_multigrid, c.cfg.reuse_solution, c.cfg.initial = "Y", "N", "U"
    c.cfg.coarse_iter, c.cfg.mg_levels = coarse_iter, levels
    return c


@pytest.mark.parametrize("levels,want", [(0, 2), (2, 2), (3, 3), (4, 4), (5, 4)])
def test_level_count(levels, want):
    # 24 x 24 x 48 cells: 12/6/3 cells across at levels 1-3; a fifth level
    # would have fewer than three cells and is not built
    g = GpuSolver(bump(levels))
    assert g.levels == want
    g.close()


def test_two_levels_is_reference_cycle():
    a, b = GpuSolver(bump(0)), GpuSolver(bump(2))
    np.testing.assert_array_equal(a.iterate(12), b.iterate(12))
    np.testing.assert_array_equal(a.get(1, "cv"), b.get(1, "cv"))


@pytest.mark.parametrize("levels", [3, 4])
def test_vcycle_graph_equals_eager(levels):
    c = bump(levels)
    a, b = GpuSolver(c), GpuSolver(c)
    ha = a.iterate(14)
    hb = np.concatenate([b.iterate(5, graph=True), b.iterate(9, graph=True)])
    np.testing.assert_array_equal(ha, hb)
    np.testing.assert_array_equal(a.get(1, "cv"), b.get(1, "cv"))


def _converge(c, precision, tol=1e-11, max_iter=40000, chunk=2000):
    g = GpuSolver(c, precision=precision)
    h = g.iterate(c.cfg.coarse_iter + 1, graph=True)
    r0 = h[-1, 1]
    it = c.cfg.coarse_iter + 1
    while it < max_iter:
        h = g.iterate(chunk, graph=True)
        it += chunk
        if h[-1, 1] <= tol * r0:
            break
    return g, it, h[-1, 1] / r0


def _fixed_point_pair(case2, casev, precision):
    ref, it2, r2 = _converge(case2, 0, tol=1e-13)
    g, it, r = _converge(casev, precision, tol=1e-13)
    print(f"two-grid FP64: {it2} iterations to {r2:.2e}; V-cycle: {it} iterations to {r:.2e}", flush=True)
    g.call("resflux")
    res = np.abs(g.get(1, "res")).max(axis=(0, 1, 2))
    ref.call("resflux")
    res2 = np.abs(ref.get(1, "res")).max(axis=(0, 1, 2))
    print("max |R| V-cycle:", res, " two-grid:", res2, flush=True)
    pv, cv = g.get(1, "pv")[1:-1, 1:-1, 1:-1].copy(), g.get(1, "cv")[1:-1, 1:-1, 1:-1].copy()
    cv2 = ref.get(1, "cv")[1:-1, 1:-1, 1:-1]
    err = np.abs(cv - cv2).max(axis=(0, 1, 2)) / np.abs(cv2).max()
    print("max field difference from the two-grid fixed point (of max |q|):", err, flush=True)
    g.close()
    ref.close()
    assert r <= 1e-13 and r2 <= 1e-13
    return res, res2, err, pv, cv


def _plain_drift(case1, pv, cv, n=2000):
    """Plain single-grid FP64 LU-SGS started from a state: how far it moves."""
    st = case1.states[0]
    st.update(rho=pv[..., 0].copy(), u=pv[..., 1].copy(), v=pv[..., 2].copy(), w=pv[..., 3].copy(),
              p=pv[..., 4].copy(), T=pv[..., 5].copy(), rhoE=cv[..., 4].copy())
    g = GpuSolver(case1)
    g.iterate(n, graph=True)
    d = np.abs(g.get(1, "cv")[1:-1, 1:-1, 1:-1] - cv).max() / np.abs(cv).max()
    g.close()
    return d


@pytest.mark.parametrize("levels,precision", [(3, 0), (4, 1)])
def test_vcycle_reaches_fine_fixed_point(levels, precision):
    # The V-cycle's converged state is a fixed point of the plain FP64
    # single-grid iteration (the fine residual is round-off there and 2000
    # plain iterations do not move it).  The discrete equations (MUSCL3 with
    # the van Albada-type limiter) have more than one steady solution at the
    # 1e-6 level: single-grid and two-grid runs land on one of them to 1e-15,
    # the V-cycle on a neighbouring one near the bump's trailing edge.
    print(f"V-cycle {levels} levels precision {precision}, laminar bump", flush=True)

    def visc(lv):
        c = bump(lv, ni=16, nj=16, nk=32, seed=6)
        c.cfg.equation, c.cfg.len_ref = "N", 1e-4
        return c

    res, res2, err, pv, cv = _fixed_point_pair(visc(2), visc(levels), precision)
    assert np.all(res <= 1e-15) and np.all(res2 <= 1e-15)
    assert np.all(err <= 1e-5)
    plain = visc(1)
    plain.cfg.use_multigrid, plain.cfg.reuse_solution = "N", "Y"
    drift = _plain_drift(plain, pv, cv)
    print(f"plain FP64 single-grid drift from the V-cycle state over 2000 iterations: {drift:.2e}", flush=True)
    assert drift <= 1e-10


def test_vcycle_euler_fixed_point():
    res, res2, err, pv, cv = _fixed_point_pair(bump(2, seed=6), bump(3, seed=6), 0)
    assert np.all(res <= 1e-15) and np.all(res2 <= 1e-15)
    assert np.all(err <= 1e-6)
    plain = bump(1, seed=6)
    plain.cfg.use_multigrid, plain.cfg.reuse_solution = "N", "Y"
    drift = _plain_drift(plain, pv, cv)
    print(f"plain FP64 single-grid drift from the V-cycle state over 2000 iterations: {drift:.2e}", flush=True)
    assert drift <= 1e-10
