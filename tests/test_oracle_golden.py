"""The C restatement against fixtures written by the compiled reference.

The fixtures (tests/golden/*.npz) hold the reference's own outputs for seeded
synthetic cases; the restatement must reproduce them bit for bit at one
thread.  Cases cover Roe and Steger-Warming, MUSCL3 / NND / MUSCL2u / MUSCL2o
reconstruction on primitive and conservative variables, the inlet,
general-outlet and wall ghost fills, the laminar Navier-Stokes terms, the
Baldwin-Lomax eddy viscosity, three-block rotating passages with periodic
and interior patches (Baldwin-Lomax, and Spalart-Allmaras with the wall
distance, the nu~ gradients, fluxes and source, and its LU-SGS equation), and the two-grid cycle started from CMesh::Initial on
the coarse level (uniform start on the bump channel, interpolated start on the
Baldwin-Lomax passage).
"""
import json
import os

import numpy as np
import pytest

from fv3d import cases
import oracle as O

TINY = ["tiny_roe_muscl3_prim", "tiny_sw_nnd_prim", "tiny_roe_muscl3_cons", "tiny_roe_muscl2o_prim",
        "tiny_sw_muscl2u_cons", "tiny_ns_roe_muscl3_prim", "tiny_bl_roe_muscl3_prim", "tiny_passage3_bl",
        "tiny_passage3_sa"]
SA_FIELDS = ("dwall", "gevI", "gevJ", "gevK", "cfctur", "cfvtur", "resev", "dev", "ev")
NS_FIELDS = ("omega", "gradI", "gradJ", "gradK", "fvI", "fvJ", "fvK", "fv", "lamvI", "lamvJ", "lamvK")


def _load(golden_dir, name):
    z = np.load(os.path.join(golden_dir, name + ".npz"))
    meta = json.loads(bytes(z["meta"]).decode())
    return z, meta


@pytest.mark.parametrize("name", TINY)
def test_first_iteration_bitexact(golden_dir, name):
    z, meta = _load(golden_dir, name)
    case = cases.from_meta(meta)
    o = O.Oracle(case)
    o.call("lusgs")
    fields = ("res", "dq", "dt", "lamcI", "lamcJ", "lamcK", "fcI", "fcJ", "fcK", "cv", "pv",
              "dIS_cv", "dIP_cv", "dJS_cv", "dJP_cv", "dKS_cv", "dKP_cv", "dKS_pv", "dKP_pv", "dIS_pv")
    if meta["kwargs"].get("viscous"):
        fields += NS_FIELDS
    if meta.get("turbulence"):
        fields += ("mut",)
    if meta.get("turbulence") == 200:
        fields += SA_FIELDS
    for b in range(1, len(case.blocks) + 1):
        tag = "" if b == 1 else f"b{b}_"
        for f in fields:
            ref = z["it1_" + tag + f]
            got = o.get(b, f).reshape(ref.shape)
            assert np.array_equal(got, ref), f"{name} block {b}: {f} differs, max {np.max(np.abs(got - ref))}"


@pytest.mark.parametrize("name", TINY)
def test_history_bitexact(golden_dir, name):
    z, meta = _load(golden_dir, name)
    case = cases.from_meta(meta)
    o = O.Oracle(case)
    h = o.iterate(meta["iterations"])
    ref = z["history"]
    # columns: iter res_dens mass_in mass_out PR eff CFL
    assert np.array_equal(h[:, :7], ref[:, :7])
    for b in range(1, len(case.blocks) + 1):
        tag = "" if b == 1 else f"b{b}_"
        assert np.array_equal(o.get(b, "cv"), z["final_" + tag + "cv"])
        assert np.array_equal(o.get(b, "pv"), z["final_" + tag + "pv"])
        if meta.get("turbulence"):
            assert np.array_equal(o.get(b, "mut").ravel(), z["final_" + tag + "mut"].ravel())
        if meta.get("turbulence") == 200:
            assert np.array_equal(o.get(b, "ev").ravel(), z["final_" + tag + "ev"].ravel())


def test_c0_first_iterations(golden_dir):
    path = os.path.join(golden_dir, "c0_bump_roe_muscl3.npz")
    if not os.path.exists(path):
        pytest.skip("C0 fixture not generated")
    z, meta = _load(golden_dir, "c0_bump_roe_muscl3")
    case = cases.bump_channel(max_iter=meta["iterations"], **meta["kwargs"])
    o = O.Oracle(case)
    h = o.iterate(3)
    assert np.array_equal(h[:, :7], z["history"][:3, :7])


MG = ["tiny_mg_bump", "tiny_mg_passage3_bl"]


@pytest.mark.parametrize("name", MG)
def test_two_grid_bitexact(golden_dir, name):
    """History and the final fields of both levels: restriction, forced coarse
    LU-SGS with frozen coarse time steps, corner treatment and prolongation."""
    z, meta = _load(golden_dir, name)
    case = cases.from_meta(meta)
    o = O.Oracle(case)
    assert o.levels() == 2
    h = o.iterate(meta["iterations"])
    assert np.array_equal(h[:, :7], z["history"][:, :7])
    for b in range(1, len(case.blocks) + 1):
        for f in ("cv", "pv", "res", "dq", "dt") + (("mut",) if meta.get("turbulence") else ()):
            ref = z[f"final_b{b}_{f}"]
            assert np.array_equal(o.get(b, f).reshape(ref.shape), ref), f"{name} fine block {b}: {f}"
        for f in ("cv", "res", "dq", "dt"):
            ref = z[f"final_c{b}_{f}"]
            assert np.array_equal(o.get(b, f, level=2).reshape(ref.shape), ref), f"{name} coarse block {b}: {f}"
