"""Two-grid cycle of the HIP library against the CPU oracle and the reference's
own fixtures.

The oracle's cycle (coarse start-up phase from CMesh::Initial, state
interpolation at Coarse_iter, restriction, forced coarse LU-SGS on frozen
coarse time steps, corner treatment and prolonged correction) is pinned bit
for bit to the compiled reference (test_oracle_vs_ref.py,
test_oracle_golden.py).  Here the device path is compared with it.  Tolerance
as for the single-grid path: FP contraction is off on both sides, and only the
device pow/exp/sqrt round differently from libm, so fields agree to 1e-12 of
their magnitude over a few iterations and to 1e-9 over the fixture runs.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from fv3d import cases
from fv3d._lib import Fv3dError
from fv3d.solver import GpuSolver
import oracle as O

pytestmark = pytest.mark.gpu

RUN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "3d-fluid-numerical-simulation-software_amd", "lib", "fv3d_run")


def relerr(a, b):
    scale = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale)


def mg(case, coarse_iter, initial="U", outlet=(0.0, 0.0, 1.0)):
    case.cfg.use_multigrid, case.cfg.reuse_solution = "Y", "N"
    case.cfg.coarse_iter, case.cfg.initial, case.cfg.outlet_angle = coarse_iter, initial, outlet
    return case


def bump(**kw):
    args = dict(ni=8, nj=8, nk=16, seed=51, max_iter=20)
    args.update(kw)
    return mg(cases.bump_channel(**args), 3)


def bl_passage(turn=0):
    c = cases.with_wall_shear(cases.annular_passage(ni=8, nj=8, nk_blocks=(6, 10, 6), seed=52, rpm=3000.0,
                                                    blade_blocks=(1,), camber=0.1, viscous=True, len_ref=1e-3,
                                                    max_iter=10))
    c.cfg.turbulence = 100
    if turn:
        c = cases.turn_block(c, 1, turn)
    c.cfg.flux_scheme = 200
    return mg(c, 2, "I", (0.0, 0.2, 1.0))


def compare_levels(g, o, case, tol, fine=("cv", "pv"), coarse=("cv", "pv"), what=""):
    for b in range(1, len(case.blocks) + 1):
        for f in fine:
            a, r = g.get(b, f), o.get(b, f)
            assert relerr(a, r) <= tol, f"{what} fine block {b} {f}: {relerr(a, r):.3e}"
        for f in coarse:
            a, r = g.get(b, f, level=2), o.get(b, f, level=2)
            assert relerr(a, r) <= tol, f"{what} coarse block {b} {f}: {relerr(a, r):.3e}"


@pytest.mark.parametrize("which", ["bump", "bump_sw_limiter", "bl_passage", "bl_passage_turn1"])
def test_phases_against_oracle(which):
    """Field by field after each phase: the coarse start-up iterations, the
    state interpolation at Coarse_iter, and two cycles."""
    case = {"bump": lambda: bump(),
            "bump_sw_limiter": lambda: bump(seed=53, flux=200, recon_scheme=200),
            "bl_passage": lambda: bl_passage(),
            "bl_passage_turn1": lambda: bl_passage(1)}[which]()
    if which == "bump_sw_limiter":
        case.cfg.limiter = "Y"
    print(f"case two-grid phases {which}", flush=True)
    g, o = GpuSolver(case), O.Oracle(case)
    assert g.levels == o.levels() == 2
    ci = case.cfg.coarse_iter
    coarse_fields = ("cv", "pv", "res", "dq", "dt", "lamcI", "lamcJ", "lamcK")
    # coarse start-up phase (the last of its iterations interpolates onto the fine level)
    hg, ho = g.iterate(ci - 1), o.iterate(ci - 1)
    assert relerr(hg[:, 1], ho[:, 1]) <= 1e-12
    compare_levels(g, o, case, 1e-12, fine=(), coarse=coarse_fields, what="start-up")
    hg, ho = g.iterate(1), o.iterate(1)
    compare_levels(g, o, case, 1e-12, fine=("cv", "pv"), coarse=coarse_fields, what="Coarse_iter")
    # two cycles: fine LU-SGS + forced coarse correction
    for n in range(2):
        hg, ho = g.iterate(1), o.iterate(1)
        assert relerr(hg[:, 1:6], ho[:, 1:6]) <= 1e-11, f"cycle {n} history"
        fine = ("cv", "pv", "res", "dq", "dt", "lamcI") + (("mut", "lamvI") if case.cfg.turbulence else ())
        compare_levels(g, o, case, 1e-11, fine=fine, coarse=("cv", "pv", "res", "dq", "dt"), what=f"cycle {n}")


@pytest.mark.parametrize("name", ["tiny_mg_bump", "tiny_mg_passage3_bl"])
def test_against_reference_fixture(golden_dir, name):
    z = np.load(os.path.join(golden_dir, name + ".npz"))
    meta = json.loads(bytes(z["meta"]).decode())
    case = cases.from_meta(meta)
    n = meta["iterations"]
    print(f"case two-grid fixture {name} n={n}", flush=True)
    g = GpuSolver(case)
    h = g.iterate(n)
    ref = z["history"]
    np.testing.assert_array_equal(h[:, 0], ref[:, 0])
    for col in range(1, 7):
        assert relerr(h[:, col], ref[:, col]) <= 1e-9, f"history col {col}"
    for b in range(1, len(case.blocks) + 1):
        for f in ("cv", "res", "dq", "dt"):
            want = z[f"final_b{b}_{f}"]
            got = g.get(b, f).reshape(want.shape)
            sl = (slice(1, -1),) * 3 if f == "cv" else (slice(None),) * 3
            assert relerr(got[sl], want[sl]) <= 1e-9, f"fine block {b} {f}"
            want = z[f"final_c{b}_{f}"]
            got = g.get(b, f, level=2).reshape(want.shape)
            assert relerr(got[sl], want[sl]) <= 1e-9, f"coarse block {b} {f}"


def test_graph_replay_matches_eager():
    """Graphs of the start-up phase and of the cycle, with the eager iteration
    between them, replay bit for bit what eager iteration computes."""
    case = bl_passage()
    case.cfg.coarse_iter = 3
    print("case two-grid graph replay", flush=True)
    a = GpuSolver(case)
    ha = a.iterate(8)
    b = GpuSolver(case)
    hb = np.concatenate([b.iterate(2, graph=True), b.iterate(3, graph=True), b.iterate(3, graph=True)])
    np.testing.assert_array_equal(ha, hb)
    for blk in range(1, len(case.blocks) + 1):
        np.testing.assert_array_equal(a.get(blk, "cv"), b.get(blk, "cv"))
        np.testing.assert_array_equal(a.get(blk, "cv", level=2), b.get(blk, "cv", level=2))


def test_single_level_when_ranges_do_not_coarsen():
    case = mg(cases.bump_channel(ni=7, nj=8, nk=16, seed=54, max_iter=6), 2)
    print("case two-grid not coarsenable", flush=True)
    g, o = GpuSolver(case), O.Oracle(case)
    assert g.levels == o.levels() == 1
    hg, ho = g.iterate(6), o.iterate(6)
    assert relerr(hg[:, 1:6], ho[:, 1:6]) <= 1e-11
    # near-zero cross-flow momenta carry the device pow's last-bit differences
    assert relerr(g.get(1, "cv"), o.get(1, "cv")) <= 1e-10


def test_cycle_from_restart_is_rejected():
    case = bump()
    case.cfg.reuse_solution = "Y"
    print("case two-grid from a restart file", flush=True)
    g = GpuSolver(case)
    assert g.levels == 2
    with pytest.raises(Fv3dError, match="fv3d_initial"):
        g.iterate(1)


def test_fp32_operator_reaches_fp64_fixed_point():
    """FP32 LU-SGS operator on both levels, FP64 residual, restriction and
    prolongation: the converged two-grid state is the FP64 one."""
    case = bump(ni=8, nj=6, nk=16, seed=55, max_iter=3000)
    print("case two-grid defect correction 16x6x8, 3000 iterations", flush=True)
    a, b = GpuSolver(case, precision=0), GpuSolver(case, precision=1)
    ha, hb = a.iterate(3000, graph=True), b.iterate(3000, graph=True)
    print(f"  final density residual FP64 {ha[-1, 1]:.3e} FP32 {hb[-1, 1]:.3e}")
    assert ha[-1, 1] < 1e-11 and hb[-1, 1] < 1e-11
    for name in ("cv", "pv"):
        x, y = a.get(1, name)[1:-1, 1:-1, 1:-1], b.get(1, name)[1:-1, 1:-1, 1:-1]
        for c in range(x.shape[-1]):
            d = np.abs(y[..., c] - x[..., c])
            sc = max(np.max(np.abs(x[..., c])), 1e-300)
            print(f"  {name}[{c}] max rel err {np.max(d) / sc:.3e}  L2 rel err {np.sqrt(np.mean(d ** 2)) / sc:.3e}")
        assert relerr(y, x) <= 1e-9


def test_cpp_runner_two_grid(tmp_path):
    """The reference's own case directory with UseMultigrid and no restart
    file, through the C++ class layer: same history as the Python mirror."""
    case = bl_passage()
    d = str(tmp_path / "case")
    case.write(d)
    print("case two-grid C++ runner", flush=True)
    hist = str(tmp_path / "h.txt")
    subprocess.run([RUN, d, "8", "--history", hist], check=True, timeout=120)
    h = np.loadtxt(hist, ndmin=2)
    g = GpuSolver(case)
    hg = g.iterate(8, graph=True)
    np.testing.assert_array_equal(h[:, :7], hg)
