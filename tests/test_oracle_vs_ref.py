"""The C restatement against the compiled reference itself (oracle/_ref).

Runs only where the reference harness has been built (make -C oracle ref, in
a container that has the reference sources); elsewhere these tests skip and
the committed fixtures (test_oracle_golden.py) carry the comparison.  The
cases cover the paths the fixtures leave out: MUSCL2c reconstruction, the
axial (radial-equilibrium) and centrifugal outlets, a rotating block with
rotating walls, laminar Navier-Stokes with Steger-Warming/NND and with a
rotating block, Baldwin-Lomax with a K-face wall and with a rotating block,
and three-block passages whose interior and periodic patches use all four
orientations of the reference's shadow-array transfer, the flipped periodic
pitch angle, and cambered blade walls.  The reference runs on one thread (its face loops race on a
shared variable with more, source/Upwind_Convective.cpp:7,29).
"""
import os
import tempfile

import numpy as np
import pytest

from fv3d import cases
from fv3d.refio import Patch
import oracle as O

pytestmark = pytest.mark.skipif(not os.path.exists(O.REF_BIN), reason="reference harness not built here")

FIELDS = ["res", "dq", "dt", "lamcI", "lamcJ", "lamcK", "cv", "pv", "fcI", "fcJ", "fcK",
          "dIS_cv", "dIP_cv", "dJS_cv", "dJP_cv", "dKS_cv", "dKP_cv"]
# laminar Navier-Stokes: vorticity, face gradients, viscous face and cell
# fluxes, viscous spectral radii
NS_FIELDS = ["omega", "gradI", "gradJ", "gradK", "fvI", "fvJ", "fvK", "fv", "lamvI", "lamvJ", "lamvK"]
# Spalart-Allmaras: wall distance, face gradients of nu~, the cell sums of its
# convective and viscous fluxes, its residual, increment and value
SA_FIELDS = ["dwall", "gevI", "gevJ", "gevK", "cfctur", "cfvtur", "resev", "dev", "ev"]


def _shifted(case, dx=1.0, dy=0.3):
    # keep every centroid away from the rotation axis (the outlets and the
    # rotating frame use r and theta of the face and cell centroids)
    case.nodes[0] = case.nodes[0] + np.array([dx, dy, 0.0])
    return case


def variant(name):
    if name == "muscl2c_axial":
        c = _shifted(cases.bump_channel(ni=6, nj=5, nk=10, seed=21, recon_scheme=300, max_iter=12))
        c.cfg.turbo_type = "A"
    elif name == "nnd_cons_sw_centrifugal":
        c = _shifted(cases.bump_channel(ni=6, nj=5, nk=10, seed=22, flux=200, recon_scheme=100, recon_var=100,
                                        max_iter=12))
        c.cfg.turbo_type = "C"
    elif name == "rotating_roe_muscl3":
        c = _shifted(cases.bump_channel(ni=6, nj=5, nk=10, seed=23, max_iter=12))
        c.blocks[0].rpm = 3000.0
        c.blocks[0].nblade = 22
        for p in c.patches:
            if p.kind == "wall" and p.face in (3, 1):
                p.rotate = 1
    elif name == "ns_sw_nnd_cons":
        c = cases.bump_channel(ni=7, nj=6, nk=9, seed=24, flux=200, recon_scheme=100, recon_var=100, max_iter=12,
                               viscous=True, len_ref=1e-4)
    elif name == "ns_rotating_roe_muscl2o":
        c = _shifted(cases.bump_channel(ni=6, nj=5, nk=10, seed=25, recon_scheme=400, max_iter=12, viscous=True,
                                        len_ref=5e-5))
        c.blocks[0].rpm = 3000.0
        c.blocks[0].nblade = 22
        for p in c.patches:
            if p.kind == "wall" and p.face in (3, 1):
                p.rotate = 1
    elif name == "bl_sw_muscl2u_kwall":
        # Baldwin-Lomax with walls on I, J and part of the K = 1 face (the
        # reference's K-wall line length is JM, so JM == KM here)
        ni, nj, nk = 6, 8, 8
        c = cases.with_wall_shear(cases.bump_channel(ni=ni, nj=nj, nk=nk, seed=26, flux=200, recon_scheme=200,
                                                     max_iter=10, viscous=True, len_ref=1e-3))
        c.cfg.turbulence = 100
        pt = []
        for p in c.patches:
            if p.kind == "inlet":
                pt.append(Patch("inlet", 1, 5, (1, ni, 1, nj - 2, 1, 1)))
                pt.append(Patch("wall", 1, 5, (1, ni, nj - 1, nj, 1, 1)))
            else:
                pt.append(p)
        c.patches = pt
    elif name == "bl_rotating_roe_muscl3":
        c = _shifted(cases.with_wall_shear(cases.bump_channel(ni=7, nj=6, nk=11, seed=27, max_iter=10, viscous=True,
                                                              len_ref=1e-3)))
        c.cfg.turbulence = 100
        c.blocks[0].rpm = 3000.0
        c.blocks[0].nblade = 22
        for p in c.patches:
            if p.kind == "wall" and p.face in (3, 1):
                p.rotate = 1
    elif name.startswith("passage_turn"):
        # three-block passage, middle block re-indexed: the interfaces use
        # orientations (3, 1), (4, 4) or (1, 3); turn 2 also makes the middle
        # block's periodic pair run from the higher angle (the flipped pitch)
        q = int(name[-1])
        c = cases.annular_passage(ni=6, nj=8, nk_blocks=(5, 6, 4), max_iter=6, seed=30 + q, rpm=5000.0)
        cases.turn_block(c, 2, q)
    elif name == "passage_bl_blade":
        c = cases.with_wall_shear(cases.annular_passage(ni=7, nj=6, nk_blocks=(4, 7, 5), max_iter=6, seed=34,
                                                        rpm=8000.0, blade_blocks=(1,), camber=0.15, viscous=True,
                                                        len_ref=1e-3, periodic_from_high=True))
        c.cfg.turbulence = 100
    elif name == "sa_bump":
        c = cases.with_sa(cases.with_wall_shear(cases.bump_channel(ni=7, nj=6, nk=10, seed=71, viscous=True,
                                                                   len_ref=1e-3, max_iter=8)))
    elif name == "sa_passage_turn1":
        c = cases.with_sa(cases.with_wall_shear(cases.annular_passage(ni=8, nj=6, nk_blocks=(5, 8, 4), seed=72,
                                                                      rpm=6000.0, blade_blocks=(1,), camber=0.1,
                                                                      viscous=True, len_ref=1e-3, max_iter=8)))
        c = cases.turn_block(c, 1, 1)
    elif name == "sa_sw_nnd_limiter_centrifugal":
        c = cases.with_sa(cases.with_wall_shear(cases.bump_channel(ni=6, nj=7, nk=9, seed=73, viscous=True,
                                                                   len_ref=1e-3, max_iter=8, flux=200,
                                                                   recon_scheme=100, recon_var=100)))
        c.cfg.limiter, c.cfg.turbo_type = "Y", "C"
    elif name == "mixing_plane_euler":
        # the interface between the second and third blocks as a mixing plane
        c = cases.annular_passage(ni=6, nj=8, nk_blocks=(5, 6, 4), max_iter=6, seed=74, rpm=5000.0)
        c = cases.with_mixing_plane(c, 2, 3)
    elif name == "mixing_plane_sa":
        c = cases.with_sa(cases.with_wall_shear(cases.annular_passage(ni=8, nj=6, nk_blocks=(5, 8, 4), seed=75,
                                                                      rpm=6000.0, blade_blocks=(1,), camber=0.1,
                                                                      viscous=True, len_ref=1e-3, max_iter=8)))
        c = cases.with_mixing_plane(c, 1, 2)
    elif name == "limiter_low_pressure":
        # two neighbouring cells at p = 1e-7: MUSCL3 reconstructs a negative
        # pressure next to them (first-order fallback), Limit_Vars replaces
        # them by the mean of their admissible neighbours, and the time-step
        # damping (Flag_low_dt = 10, decremented per TimeStep) runs twice
        c = cases.bump_channel(ni=6, nj=5, nk=10, seed=41, max_iter=12)
        c.cfg.limiter = "Y"
        c = cases.with_cell_values(c, 1, (2, 2, slice(4, 6)), p=1e-7)
    elif name == "limiter_dense_corner":
        # a 2x2x2 block corner at rho = 1e4: the corner cell has no
        # admissible neighbour (the last-neighbour clamp of kn == 0)
        c = cases.bump_channel(ni=6, nj=5, nk=10, seed=41, max_iter=12)
        c.cfg.limiter = "Y"
        c = cases.with_cell_values(c, 1, (slice(0, 2), slice(0, 2), slice(0, 2)), rho=1e4)
    elif name == "limiter_sa_high":
        # nu~ above 1000: Limit_SA's neighbour mean, and at the block's first
        # cell (all neighbours above 1000) its 1000 clamp
        c = cases.with_sa(cases.with_wall_shear(cases.bump_channel(ni=7, nj=6, nk=10, seed=71, viscous=True,
                                                                   len_ref=1e-3, max_iter=8)))
        c.cfg.limiter = "Y"
        c = cases.with_cell_values(c, 1, (slice(0, 2), slice(0, 2), slice(0, 2)), nut=1e5)
        c = cases.with_cell_values(c, 1, (0, 0, 0), nut=3000.0)
        c = cases.with_cell_values(c, 1, (4, 3, 6), nut=1e5)
    elif name == "stage_mixing_plane":
        # single grid: rotor and stator rows of different pitchwise cell
        # counts, blade counts and speeds at a mixing plane, Spalart-Allmaras
        c = cases.rotor_stator_stage(ni=6, nj_rotor=8, nj_stator=5, rotor_nk=(4, 6, 4), stator_nk=(3, 7, 4),
                                     max_iter=8)
        c.cfg.use_multigrid = "N"
    elif name == "char_muscl3_roe":
        c = _shifted(cases.bump_channel(ni=6, nj=5, nk=10, seed=42, recon_var=300, max_iter=12))
    elif name == "char_nnd_sw_ns":
        c = cases.with_wall_shear(cases.bump_channel(ni=7, nj=6, nk=9, seed=43, flux=200, recon_scheme=100,
                                                     recon_var=300, max_iter=12, viscous=True, len_ref=1e-4))
    elif name == "char_muscl2u_low_pressure":
        # characteristic MUSCL2u next to a low-pressure pair: the conservative
        # positivity fallback of CharacterReconstruct, then Limit_Vars
        c = cases.bump_channel(ni=6, nj=5, nk=10, seed=44, recon_scheme=200, recon_var=300, max_iter=12)
        c.cfg.limiter = "Y"
        c = cases.with_cell_values(c, 1, (2, 2, slice(4, 6)), p=1e-7)
    elif name == "char_muscl2o_rotating_passage":
        c = cases.annular_passage(ni=6, nj=8, nk_blocks=(5, 6, 4), max_iter=6, seed=45, rpm=5000.0)
        c.cfg.recon_scheme, c.cfg.recon_var = 400, 300
        cases.turn_block(c, 2, 1)
    elif name == "char_muscl2c_bl":
        c = cases.with_wall_shear(cases.annular_passage(ni=7, nj=6, nk_blocks=(4, 7, 5), max_iter=6, seed=46,
                                                        rpm=8000.0, blade_blocks=(1,), camber=0.15, viscous=True,
                                                        len_ref=1e-3))
        c.cfg.turbulence, c.cfg.recon_scheme, c.cfg.recon_var, c.cfg.flux_scheme = 100, 300, 300, 200
    else:
        raise KeyError(name)
    return c


# repair branches each case must take in the restatement (oracle counters)
MUST_FIRE = {
    "limiter_low_pressure": ("recon_fallback", "limit_vars", "low_dt"),
    "limiter_dense_corner": ("recon_fallback", "limit_vars", "limit_vars_kn0", "low_dt"),
    "limiter_sa_high": ("limit_sa", "limit_sa_kn0"),
    "char_muscl2u_low_pressure": ("recon_fallback", "limit_vars", "low_dt"),
}


@pytest.mark.parametrize("name", ["muscl2c_axial", "nnd_cons_sw_centrifugal", "rotating_roe_muscl3",
                                  "ns_sw_nnd_cons", "ns_rotating_roe_muscl2o", "bl_sw_muscl2u_kwall",
                                  "bl_rotating_roe_muscl3", "passage_turn1", "passage_turn2", "passage_turn3",
                                  "passage_bl_blade", "sa_bump", "sa_passage_turn1",
                                  "sa_sw_nnd_limiter_centrifugal", "mixing_plane_euler", "mixing_plane_sa",
                                  "limiter_low_pressure", "limiter_dense_corner", "limiter_sa_high",
                                  "char_muscl3_roe", "char_nnd_sw_ns", "char_muscl2u_low_pressure",
                                  "char_muscl2o_rotating_passage", "char_muscl2c_bl", "stage_mixing_plane"])
def test_restatement_matches_reference(name):
    case = variant(name)
    n = case.cfg.max_iter
    fields = FIELDS + (NS_FIELDS if case.cfg.equation == "N" else []) + (["mut"] if case.cfg.turbulence else [])
    if case.cfg.turbulence == 200:
        fields = fields + SA_FIELDS
    with tempfile.TemporaryDirectory() as td:
        cdir, odir = os.path.join(td, "case"), os.path.join(td, "out")
        case.write(cdir)
        ref = O.run_reference(cdir, n, odir, threads=1, dump_first=True, timeout=300)
        o = O.Oracle(case)
        o.call("lusgs")
        for b in range(1, len(case.blocks) + 1):
            for f in fields:
                want = np.load(os.path.join(odir, f"it1_b{b}_{f}.npy"))
                got = o.get(b, f).reshape(want.shape)
                assert np.array_equal(got, want), f"{name} block {b}: {f}, max diff {np.max(np.abs(got - want))}"
        o2 = O.Oracle(case)
        O.stats()
        h = o2.iterate(n)
        fired = O.stats()
        for branch in MUST_FIRE.get(name, ()):
            assert fired[branch] > 0, f"{name}: the {branch} branch never ran ({fired})"
        # (the rotating case reverses the outlet mass flow, which makes the
        # mass-averaged pressure ratio NaN on both sides)
        assert np.array_equal(h[:, :7], ref["history"][:, :7], equal_nan=True), name
        for b in range(1, len(case.blocks) + 1):
            want = np.load(os.path.join(odir, f"final_b{b}_cv.npy"))
            assert np.array_equal(o2.get(b, "cv"), want)
            if case.cfg.turbulence:
                want = np.load(os.path.join(odir, f"final_b{b}_mut.npy"))
                assert np.array_equal(o2.get(b, "mut").ravel(), want.ravel())
            if case.cfg.turbulence == 200:
                want = np.load(os.path.join(odir, f"final_b{b}_ev.npy"))
                assert np.array_equal(o2.get(b, "ev").ravel(), want.ravel())
        if case.cfg.turbulence == 200:
            assert o2.res_calculate()[5] == ref["history"][-1, 7], "Res_ev"


# ------------------------------------------------------------ two-grid cycle
def mg_variant(name):
    """Two-grid cases (reuse_solution = N: CMesh::Initial on the coarse level,
    Coarse_iter coarse iterations, then the cycle of source/Solver.cpp:86-103)."""
    if name == "bump_uniform":
        c = cases.bump_channel(ni=8, nj=8, nk=16, seed=31, max_iter=12)
        c.cfg.coarse_iter, c.cfg.initial = 5, "U"
    elif name == "bump_interpolate_sw_limiter":
        c = cases.bump_channel(ni=8, nj=6, nk=16, seed=32, flux=200, recon_scheme=200, max_iter=12)
        c.cfg.coarse_iter, c.cfg.initial, c.cfg.outlet_angle = 4, "I", (0.1, 0.0, 1.0)
        c.cfg.limiter = "Y"
    elif name == "bl_bump":
        c = cases.with_wall_shear(cases.bump_channel(ni=8, nj=8, nk=16, seed=33, viscous=True, len_ref=1e-3,
                                                     max_iter=12))
        c.cfg.turbulence = 100
        c.cfg.coarse_iter, c.cfg.initial = 4, "U"
    elif name in ("bl_passage", "bl_passage_turn1"):
        c = cases.with_wall_shear(cases.annular_passage(ni=8, nj=8, nk_blocks=(6, 10, 6), seed=34, rpm=3000.0,
                                                        blade_blocks=(1,), camber=0.1, viscous=True, len_ref=1e-3,
                                                        max_iter=10))
        c.cfg.turbulence = 100
        if name.endswith("turn1"):
            c = cases.turn_block(c, 1, 1)
        c.cfg.coarse_iter, c.cfg.initial, c.cfg.outlet_angle = 3, "I", (0.0, 0.2, 1.0)
        c.cfg.flux_scheme = 200
    elif name == "global_dt_axial_passage":
        # global time step (per-block minimum) and the radial-equilibrium outlet
        # on a three-block passage
        c = cases.annular_passage(ni=8, nj=6, nk_blocks=(4, 8, 4), seed=36, rpm=4000.0, max_iter=8)
        c.cfg.timestep, c.cfg.turbo_type = "G", "A"
        c.cfg.coarse_iter, c.cfg.initial = 2, "U"
    elif name == "sa_passage_global_dt_axial":
        c = cases.with_wall_shear(cases.annular_passage(ni=8, nj=8, nk_blocks=(6, 10, 6), seed=74, rpm=3000.0,
                                                        blade_blocks=(1,), camber=0.1, viscous=True, len_ref=1e-3,
                                                        max_iter=10))
        c.cfg.turbulence = 200
        c.cfg.coarse_iter, c.cfg.initial, c.cfg.outlet_angle = 3, "I", (0.0, 0.2, 1.0)
        c.cfg.timestep, c.cfg.turbo_type = "G", "A"
    elif name == "bl_passage_mixing_plane":
        c = cases.with_wall_shear(cases.annular_passage(ni=8, nj=8, nk_blocks=(6, 8, 6), seed=76, rpm=3000.0,
                                                        blade_blocks=(1,), camber=0.1, viscous=True, len_ref=1e-3,
                                                        max_iter=8))
        c.cfg.turbulence = 100
        c.cfg.coarse_iter, c.cfg.initial, c.cfg.outlet_angle = 2, "I", (0.0, 0.2, 1.0)
        c = cases.with_mixing_plane(c, 2, 3)
    elif name == "stage_mixing_plane_sa":
        # a rotor row (8 cells pitchwise, 22 blades, 3000 rpm) and a stator row
        # (6 cells, 30 blades, stationary) joined by a mixing plane whose two
        # sides differ in cell count, pitch and rotation
        c = cases.rotor_stator_stage(ni=6, nj_rotor=8, nj_stator=6, rotor_nk=(6, 8, 6), stator_nk=(6, 10, 6),
                                     coarse_iter=2, max_iter=8)
    elif name == "not_coarsenable":
        # 7 cells along i: the patch ranges end on an even node, one level only
        c = cases.bump_channel(ni=7, nj=8, nk=16, seed=35, max_iter=8)
        c.cfg.coarse_iter, c.cfg.initial = 3, "U"
    else:
        raise KeyError(name)
    c.cfg.use_multigrid, c.cfg.reuse_solution = "Y", "N"
    return c


MG_FINE = ["cv", "pv", "res", "dq", "dt", "lamcI", "lamcJ", "lamcK", "dIS_cv", "dIP_cv", "dJS_cv", "dJP_cv",
           "dKS_cv", "dKP_cv"]
MG_COARSE = ["cv", "pv", "res", "dq", "dt", "lamcI", "lamcJ", "lamcK"]


@pytest.mark.parametrize("name", ["bump_uniform", "bump_interpolate_sw_limiter", "bl_bump", "bl_passage",
                                  "bl_passage_turn1", "global_dt_axial_passage", "sa_passage_global_dt_axial",
                                  "not_coarsenable", "bl_passage_mixing_plane", "stage_mixing_plane_sa"])
def test_two_grid_matches_reference(name):
    case = mg_variant(name)
    n = case.cfg.max_iter
    with tempfile.TemporaryDirectory() as td:
        cdir, odir = os.path.join(td, "case"), os.path.join(td, "out")
        case.write(cdir)
        ref = O.run_reference(cdir, n, odir, threads=1, dump_coarse=True, timeout=300)
        o = O.Oracle(case)
        assert o.levels() == ref["levels"] == (1 if name == "not_coarsenable" else 2)
        h = o.iterate(n)
        assert np.all(np.isfinite(h[:, :7]))
        assert np.array_equal(h[:, :7], ref["history"][:, :7]), name
        visc = case.cfg.equation == "N"
        for b in range(1, len(case.blocks) + 1):
            for f in MG_FINE + (["lamvI", "lamvJ", "lamvK"] if visc else []) + (["mut"] if case.cfg.turbulence else []):
                want = np.load(os.path.join(odir, f"final_b{b}_{f}.npy"))
                got = o.get(b, f).reshape(want.shape)
                assert np.array_equal(got, want), f"{name} fine block {b}: {f}"
            if ref["levels"] == 2:
                for f in MG_COARSE + (["lamvI", "lamvJ", "lamvK"] if visc else []):
                    want = np.load(os.path.join(odir, f"final_c{b}_{f}.npy"))
                    got = o.get(b, f, level=2).reshape(want.shape)
                    assert np.array_equal(got, want), f"{name} coarse block {b}: {f}"
