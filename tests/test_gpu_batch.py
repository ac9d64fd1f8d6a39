"""Batched instances: independent passages advanced together by every launch
of one context (fv3d_instance_add), against each passage run alone.

Each instance must behave exactly as its own run: the same cells, patches,
norms (its cells start on a whole 256-cell group, so its norm partials are
those of the lone run), performance sections, CFL ramp and outlet pressure.
The bar is bit equality in FP64 and in the FP32 operator form, eager and
replayed from a HIP graph, single grid and two-grid, Baldwin-Lomax and
Spalart-Allmaras.
"""
import numpy as np
import pytest

from fv3d import cases
from fv3d._lib import Fv3dError
from fv3d.solver import GpuSolver

pytestmark = pytest.mark.gpu


def passage(seed, ni, nj, nk_blocks, rpm, pratio, turbulence=100, mg=False):
    c = cases.with_wall_shear(cases.annular_passage(ni=ni, nj=nj, nk_blocks=nk_blocks, seed=seed, rpm=rpm,
                                                    blade_blocks=(1,), camber=0.1, viscous=True, len_ref=1e-3,
                                                    max_iter=8))
    c.cfg.p_out = c.cfg.p_out * pratio
    if turbulence == 200:
        c = cases.with_sa(c)
    else:
        c.cfg.turbulence = turbulence
    if mg:
        c.cfg.use_multigrid, c.cfg.reuse_solution, c.cfg.coarse_iter, c.cfg.initial = "Y", "N", 2, "I"
        c.cfg.outlet_angle = (0.0, 0.2, 1.0)
    return c


def family(turbulence=100, mg=False):
    # uneven sizes, speeds and back pressures
    if mg:
        shapes = [(81, 8, 8, (6, 8, 6), 3000.0, 1.0), (82, 8, 6, (6, 10, 6), 2500.0, 1.05),
                  (83, 10, 8, (6, 12, 6), 3500.0, 0.97)]
    else:
        shapes = [(81, 8, 6, (5, 8, 4), 6000.0, 1.0), (82, 6, 8, (4, 6, 5), 4500.0, 1.08),
                  (83, 10, 6, (6, 9, 5), 5200.0, 0.96)]
    return [passage(s, ni, nj, nk, rpm, pr, turbulence, mg) for s, ni, nj, nk, rpm, pr in shapes]


def compare(batch, alone, cs, n, graph, fields):
    hb = batch.iterate(n, graph=graph)
    assert hb.shape == (n, len(cs), 7)
    for w, (c, a) in enumerate(zip(cs, alone)):
        ha = a.iterate(n, graph=graph)
        np.testing.assert_array_equal(hb[:, w], ha, err_msg=f"instance {w + 1} history")
        for b in range(1, len(c.blocks) + 1):
            for f in fields:
                np.testing.assert_array_equal(batch.get(b, f, inst=w + 1), a.get(b, f),
                                              err_msg=f"instance {w + 1} block {b} {f}")
        rb, pb = batch.instance_results(w + 1)
        ra = a.res_calculate()
        np.testing.assert_array_equal(rb, ra)
        np.testing.assert_array_equal(pb, a.perf())


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("precision", [0, 1])
def test_baldwin_lomax_batch_matches_alone(precision, graph):
    cs = family(100)
    print(f"case batch BL precision={precision} graph={graph}", flush=True)
    batch = GpuSolver(cs, precision=precision)
    assert int(batch.L.fv3d_instances(batch.h)) == 3
    alone = [GpuSolver(c, precision=precision) for c in cs]
    assert batch.n_cells == sum(a.n_cells for a in alone)
    compare(batch, alone, cs, 6, graph, ("cv", "pv", "mut", "dt", "res"))


def test_spalart_allmaras_batch_matches_alone():
    cs = family(200)
    print("case batch SA", flush=True)
    batch = GpuSolver(cs)
    alone = [GpuSolver(c) for c in cs]
    for w, (c, a) in enumerate(zip(cs, alone)):
        for b in range(1, len(c.blocks) + 1):
            np.testing.assert_array_equal(batch.get(b, "dwall", inst=w + 1), a.get(b, "dwall"))
    compare(batch, alone, cs, 6, True, ("cv", "ev", "mut", "dev"))


def test_two_grid_batch_matches_alone():
    cs = family(100, mg=True)
    print("case batch two-grid", flush=True)
    batch = GpuSolver(cs)
    assert batch.levels == 2
    alone = [GpuSolver(c) for c in cs]
    compare(batch, alone, cs, 6, False, ("cv", "pv", "mut"))
    for w, (c, a) in enumerate(zip(cs, alone)):
        for b in range(1, len(c.blocks) + 1):
            np.testing.assert_array_equal(batch.get(b, "cv", level=2, inst=w + 1), a.get(b, "cv", level=2))


def test_two_grid_batch_graph_matches_eager():
    cs = family(100, mg=True)
    print("case batch two-grid graph replay", flush=True)
    a, b = GpuSolver(cs), GpuSolver(cs)
    ha = a.iterate(7)
    hb = np.concatenate([b.iterate(1, graph=True), b.iterate(2, graph=True), b.iterate(4, graph=True)])
    np.testing.assert_array_equal(ha, hb)


def test_instances_before_blocks():
    """An instance is declared before its blocks; the library says so."""
    cs = family(100)
    g = GpuSolver(cs[0])
    L = g.L
    import ctypes as C
    from fv3d.solver import make_config
    cfg = make_config(cs[0].cfg)
    h = C.c_void_p()
    assert L.fv3d_ctx_create(C.byref(cfg), C.byref(h)) == 0
    bi = cs[0].blocks[0]
    a = np.ascontiguousarray(cs[0].nodes[0])
    assert L.fv3d_block_add(h, bi.im, bi.jm, bi.km, a.ctypes.data, bi.nblade, bi.rpm) == 1
    assert L.fv3d_instance_add(h, cs[0].cfg.p_out) < 0
    assert b"before their blocks" in L.fv3d_last_error(h)
    L.fv3d_ctx_destroy(h)


def test_batch_needs_shared_config():
    cs = family(100)
    cs[1].cfg.cfl = cs[0].cfg.cfl * 2
    with pytest.raises(ValueError, match="share"):
        GpuSolver(cs)


def test_instances_must_agree_on_levels():
    cs = family(100, mg=True)
    odd = passage(84, 7, 8, (6, 8, 6), 3000.0, 1.0, mg=True)   # 7 cells along i: one level
    print("case batch with a non-coarsening instance", flush=True)
    with pytest.raises(Fv3dError, match="do not all coarsen"):
        GpuSolver(cs + [odd])
