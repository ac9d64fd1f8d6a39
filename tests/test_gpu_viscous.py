"""Navier-Stokes path (laminar and Baldwin-Lomax) of the HIP library against
the CPU oracle.

The oracle's viscous and turbulence terms are pinned bit for bit to the
compiled reference (test_oracle_golden.py, test_oracle_vs_ref.py) and to closed
forms (test_oracle_viscous_laws.py).  Here the HIP kernels (metric
coefficients, face gradients, viscous flux, viscous spectral radii, vorticity,
the viscous terms of the LU-SGS sweeps and the Baldwin-Lomax eddy viscosity)
are compared with it.  Tolerance: FP contraction is off in both, so values
agree to the last bits except where the device pow and exp (Sutherland's law,
the wall distance, the Baldwin-Lomax damping, the MUSCL3 limiter) round
differently from libm's; the bound is 1e-12 of each field's magnitude for one
iteration and 1e-9 for twenty.
"""
import json
import os

import numpy as np
import pytest

from fv3d import cases
from fv3d.solver import GpuSolver
import oracle as O

pytestmark = pytest.mark.gpu


def relerr(a, b):
    scale = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale)


def ns_case(**kw):
    args = dict(ni=6, nj=5, nk=10, seed=41, viscous=True, len_ref=2e-5, max_iter=20)
    args.update(kw)
    return cases.bump_channel(**args)


def rotating_ns_case():
    c = ns_case(ni=7, nj=6, nk=11, seed=42, recon_scheme=400, len_ref=5e-5)
    c.nodes[0] = c.nodes[0] + np.array([1.0, 0.3, 0.0])
    c.blocks[0].rpm = 3000.0
    c.blocks[0].nblade = 22
    for p in c.patches:
        if p.kind == "wall" and p.face in (3, 1):
            p.rotate = 1
    return c


@pytest.mark.parametrize("which", ["roe_muscl3", "sw_nnd_cons", "rotating"])
def test_residual_stage_fields(which):
    case = {"roe_muscl3": lambda: ns_case(),
            "sw_nnd_cons": lambda: ns_case(seed=43, flux=200, recon_scheme=100, recon_var=100, len_ref=1e-4),
            "rotating": rotating_ns_case}[which]()
    print(f"case NS residual {which}", flush=True)
    o = O.Oracle(case)
    o.call("resflux")
    g = GpuSolver(case, debug=1)
    g.call("resflux")
    g.call("curl")
    for f in ("lamvI", "lamvJ", "lamvK", "dt", "gradI", "gradJ", "gradK", "fvI", "fvJ", "fvK", "fv", "res", "omega"):
        a, b = g.get(1, f), o.get(1, f)
        assert relerr(a, b) <= 1e-12, f"{which} {f}: rel err {relerr(a, b):.3e}"


def test_first_iteration():
    case = ns_case()
    print("case NS first iteration", flush=True)
    o = O.Oracle(case)
    o.call("lusgs")
    g = GpuSolver(case)
    g.call("lusgs")
    for f in ("dq", "cv", "pv", "lamvI", "lamvJ", "lamvK"):
        a, b = g.get(1, f), o.get(1, f)
        assert relerr(a, b) <= 1e-12, f"{f}: rel err {relerr(a, b):.3e}"


def test_history_against_reference_fixture(golden_dir):
    z = np.load(os.path.join(golden_dir, "tiny_ns_roe_muscl3_prim.npz"))
    meta = json.loads(bytes(z["meta"]).decode())
    n = meta["iterations"]
    case = cases.bump_channel(max_iter=n, **meta["kwargs"])
    print(f"case NS fixture n={n}", flush=True)
    g = GpuSolver(case)
    h = g.iterate(n)
    ref = z["history"]
    np.testing.assert_array_equal(h[:, 0], ref[:, 0])
    for col in range(1, 7):
        assert relerr(h[:, col], ref[:, col]) <= 1e-9, f"history col {col}"
    assert relerr(g.get(1, "cv")[1:-1, 1:-1, 1:-1], z["final_cv"][1:-1, 1:-1, 1:-1]) <= 1e-9


def test_row_sweep_equals_plane_sweep():
    """Partial 4x16 sweep tiles and 4x4x16 residual tiles: the default sweep and
    the per-hyperplane sweep give the same Gauss-Seidel result exactly."""
    case = ns_case(ni=19, nj=13, nk=29, seed=44, max_iter=3)
    print("case NS row vs plane sweep 19x13x29", flush=True)
    a = GpuSolver(case)
    b = GpuSolver(case, debug=2)
    a.call("lusgs")
    b.call("lusgs")
    np.testing.assert_array_equal(a.get(1, "dq"), b.get(1, "dq"))
    np.testing.assert_array_equal(a.iterate(3), b.iterate(3))
    o = O.Oracle(case)
    o.iterate(4)
    assert relerr(a.get(1, "cv"), o.get(1, "cv")) <= 1e-11


def test_rotating_iterations():
    case = rotating_ns_case()
    print("case NS rotating 4 iterations", flush=True)
    g = GpuSolver(case)
    h = g.iterate(4)
    o = O.Oracle(case)
    ho = o.iterate(4)
    assert relerr(h[:, 1], ho[:, 1]) <= 1e-10
    assert relerr(g.get(1, "cv"), o.get(1, "cv")) <= 1e-11


def test_graph_replay_matches_eager():
    case = ns_case(seed=45)
    print("case NS graph replay", flush=True)
    a = GpuSolver(case)
    ha = a.iterate(6)
    b = GpuSolver(case)
    hb = np.concatenate([b.iterate(2, graph=True), b.iterate(4, graph=True)])
    np.testing.assert_array_equal(ha, hb)
    np.testing.assert_array_equal(a.get(1, "cv"), b.get(1, "cv"))


# ------------------------------------------------------------- Baldwin-Lomax
def bl_case(**kw):
    args = dict(ni=7, nj=6, nk=10, seed=46, viscous=True, len_ref=1e-3, max_iter=20)
    args.update(kw)
    c = cases.with_wall_shear(cases.bump_channel(**args))
    c.cfg.turbulence = 100
    return c


def bl_kwall_case():
    from fv3d.refio import Patch
    ni, nj, nk = 6, 8, 8
    c = bl_case(ni=ni, nj=nj, nk=nk, seed=47, flux=200, recon_scheme=200)
    pt = []
    for p in c.patches:
        if p.kind == "inlet":
            pt.append(Patch("inlet", 1, 5, (1, ni, 1, nj - 2, 1, 1)))
            pt.append(Patch("wall", 1, 5, (1, ni, nj - 1, nj, 1, 1)))
        else:
            pt.append(p)
    c.patches = pt
    return c


@pytest.mark.parametrize("which", ["ij_walls", "k_wall"])
def test_bl_first_iteration(which):
    case = bl_case() if which == "ij_walls" else bl_kwall_case()
    print(f"case BL first iteration {which}", flush=True)
    o = O.Oracle(case)
    o.call("lusgs")
    g = GpuSolver(case)
    g.call("lusgs")
    assert np.max(np.abs(o.get(1, "mut"))) > 0.0
    for f in ("mut", "lamvI", "lamvJ", "lamvK", "dt", "res", "dq", "cv", "pv"):
        a, b = g.get(1, f), o.get(1, f)
        assert relerr(a, b) <= 1e-12, f"{which} {f}: rel err {relerr(a, b):.3e}"


def test_bl_history_against_reference_fixture(golden_dir):
    z = np.load(os.path.join(golden_dir, "tiny_bl_roe_muscl3_prim.npz"))
    meta = json.loads(bytes(z["meta"]).decode())
    n = meta["iterations"]
    case = cases.from_meta(meta)
    print(f"case BL fixture n={n}", flush=True)
    g = GpuSolver(case)
    h = g.iterate(n)
    ref = z["history"]
    for col in range(1, 7):
        assert relerr(h[:, col], ref[:, col]) <= 1e-9, f"history col {col}"
    assert relerr(g.get(1, "cv")[1:-1, 1:-1, 1:-1], z["final_cv"][1:-1, 1:-1, 1:-1]) <= 1e-9
    assert relerr(g.get(1, "mut").ravel(), z["final_mut"].ravel()) <= 1e-9


def test_bl_row_sweep_equals_plane_sweep():
    case = bl_case(ni=19, nj=13, nk=29, seed=48, max_iter=3)
    print("case BL row vs plane sweep 19x13x29", flush=True)
    a = GpuSolver(case)
    b = GpuSolver(case, debug=2)
    np.testing.assert_array_equal(a.iterate(3), b.iterate(3))
    np.testing.assert_array_equal(a.get(1, "mut"), b.get(1, "mut"))
    o = O.Oracle(case)
    o.iterate(3)
    assert relerr(a.get(1, "cv"), o.get(1, "cv")) <= 1e-11
    assert relerr(a.get(1, "mut"), o.get(1, "mut")) <= 1e-10
