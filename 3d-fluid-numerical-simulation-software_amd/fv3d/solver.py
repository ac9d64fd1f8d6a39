"""Host-side mirror of the reference's CMesh / CSolver surface over libfv3d.

``GpuSolver`` loads a case (fv3d.cases.Case, i.e. the reference's config /
grid / boundary-definition / restart content) into a libfv3d context and
exposes the stages of CMesh::ResFlux and CMesh::LU_SGS_Marching
(source/ComputeAllBlockVars.cpp:223-255, source/LU_SGS_Marching.cpp:10-291)
and the iteration of CSolver::Solver (source/Solver.cpp:47-135), single grid
or the two-grid cycle.  Without a restart (``reuse_solution = N``) the start
field comes from CMesh::Initial on the level the reference starts from.

Given a list of cases, ``GpuSolver`` runs them as a batch: one context, each
case an instance (fv3d_instance_add) advanced by the same launches.  The cases
share every config value except the outlet pressure; rotation rates are per
block anyway.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .refio import Config as CaseConfig

_TURBO = {"A": 0, "C": 1}
_KIND = {"inlet": 0, "outlet": 1, "wall": 2}


def make_config(cfg: CaseConfig, device: int = 0, precision: int = 0) -> _lib.Config:
    s = _lib.Config()
    s.rgas, s.gamma, s.pr_lam, s.pr_tur = cfg.rgas, cfg.gamma, cfg.pr_lam, cfg.pr_tur
    s.pt_in, s.tt_in, s.p_out, s.ma0 = cfg.pt_in, cfg.tt_in, cfg.p_out, cfg.ma0
    s.len_ref, s.ev_ratio = cfg.len_ref, cfg.ev_ratio
    for c in range(3):
        s.inlet_dir[c] = cfg.inlet_angle[c]
    s.cfl, s.c_timestep, s.relax = cfg.cfl, cfg.c_timestep, cfg.relax_factor
    s.turbo_type = _TURBO.get(cfg.turbo_type.upper(), 2)
    s.viscous = 0 if cfg.equation.upper() == "E" else 1
    s.recon_var, s.recon_scheme, s.flux_scheme = cfg.recon_var, cfg.recon_scheme, cfg.flux_scheme
    s.turbulence = cfg.turbulence
    s.limiter = 1 if cfg.limiter.upper() == "Y" else 0
    s.global_dt = 0 if cfg.timestep.upper() == "L" else 1
    s.device = device
    s.precision = precision
    s.multigrid = 1 if cfg.use_multigrid.upper() == "Y" else 0
    s.coarse_iter = cfg.coarse_iter
    s.initial_type = 0 if cfg.initial.upper() == "U" else 1
    s.p21ratio = cfg.p21_ratio
    for c in range(3):
        s.outlet_dir[c] = cfg.outlet_angle[c]
    s.mg_levels = cfg.mg_levels
    return s


def _shared_config(cfg: _lib.Config) -> tuple:
    """The config values the instances of a batch must share (all but p_out)."""
    out = []
    for name, _ in _lib.Config._fields_:
        if name == "p_out":
            continue
        v = getattr(cfg, name)
        out.append(tuple(v) if hasattr(v, "__len__") else v)
    return tuple(out)


class GpuSolver:
    """One libfv3d context holding one case, or a batch of cases."""

    def __init__(self, case, device: int = 0, precision: int = 0, debug: int = 0):
        self.L = _lib.load()
        self.batch = isinstance(case, (list, tuple))
        self.cases = list(case) if self.batch else [case]
        self.case = self.cases[0]
        self._cfg = make_config(self.case.cfg, device, precision)
        ref = _shared_config(self._cfg)
        for cs in self.cases[1:]:
            if _shared_config(make_config(cs.cfg, device, precision)) != ref:
                raise ValueError("the cases of a batch share every config value but the outlet pressure")
            if cs.cfg.reuse_solution.upper() != self.case.cfg.reuse_solution.upper():
                raise ValueError("the cases of a batch all start from a restart file, or none does")
        h = C.c_void_p()
        rc = self.L.fv3d_ctx_create(C.byref(self._cfg), C.byref(h))
        self.h = h
        self._check(rc, "ctx_create")
        self.offsets = []   # global number of each instance's first block, less one
        nblock = 0
        for cs in self.cases:
            if self.batch:
                self._check(self.L.fv3d_instance_add(self.h, cs.cfg.p_out), "instance_add")
            off = nblock
            self.offsets.append(off)
            for b, xyz in enumerate(cs.nodes):
                bi = cs.blocks[b]
                a = np.ascontiguousarray(xyz, dtype=np.float64)
                rc = self.L.fv3d_block_add(self.h, bi.im, bi.jm, bi.km, a.ctypes.data, bi.nblade, bi.rpm)
                self._check(rc, "block_add")
            nblock += len(cs.nodes)
            for p in cs.patches:
                if p.kind == "mixing":
                    r1 = np.ascontiguousarray(p.rng, dtype=np.int32)
                    r2 = np.ascontiguousarray(p.rng2, dtype=np.int32)
                    self._check(self.L.fv3d_mixing_plane_add(self.h, p.block + off, r1.ctypes.data, p.block2 + off,
                                                             r2.ctypes.data), "mixing_plane_add")
                    continue
                if p.kind in ("interior", "periodic"):
                    r1 = np.ascontiguousarray(p.rng, dtype=np.int32)
                    r2 = np.ascontiguousarray(p.rng2, dtype=np.int32)
                    self._check(self.L.fv3d_connection_add(self.h, 0 if p.kind == "interior" else 1, p.orient,
                                                           p.block + off, p.face, r1.ctypes.data, p.block2 + off,
                                                           p.face2, r2.ctypes.data), "connection_add")
                    continue
                if p.kind not in _KIND:
                    raise NotImplementedError(f"patch kind {p.kind} not in this build")
                r = np.ascontiguousarray(p.rng, dtype=np.int32)
                self._check(self.L.fv3d_patch_add(self.h, _KIND[p.kind], p.block + off, p.face, r.ctypes.data,
                                                  p.rotate), "patch_add")
        self._check(self.L.fv3d_setup(self.h), "setup")
        if debug:
            self._check(self.L.fv3d_set_debug(self.h, debug), "set_debug")
        if self.case.cfg.reuse_solution.upper() == "Y":
            for cs, off in zip(self.cases, self.offsets):
                for b, st in enumerate(cs.states):
                    pv = np.ascontiguousarray(np.stack([st["rho"], st["u"], st["v"], st["w"], st["p"], st["T"]], -1))
                    e = np.ascontiguousarray(st["rhoE"])
                    self._check(self.L.fv3d_state_upload(self.h, off + b + 1, pv.ctypes.data, e.ctypes.data),
                                "state_upload")
                    if cs.cfg.turbulence == 200 and "nut" in st:
                        ev = np.ascontiguousarray(st["nut"], dtype=np.float64)
                        self._check(self.L.fv3d_nut_upload(self.h, off + b + 1, ev.ctypes.data), "nut_upload")
        else:
            self._check(self.L.fv3d_initial(self.h), "initial")

    def _check(self, rc, what):
        if rc < 0:
            msg = self.L.fv3d_last_error(self.h).decode() if self.h else ""
            raise _lib.Fv3dError(f"fv3d {what} failed ({rc}): {msg}")
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.L.fv3d_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # -- stages
    def call(self, stage: str) -> None:
        self._check(getattr(self.L, "fv3d_" + stage)(self.h), stage)

    def iterate(self, n: int, graph: bool = False) -> np.ndarray:
        """History rows (n, 7); for a batch (n, instances, 7)."""
        k = len(self.cases)
        h = np.zeros((max(n, 1), k, 7))
        fn = self.L.fv3d_iterate_graph if graph else self.L.fv3d_iterate
        self._check(fn(self.h, n, h.ctypes.data), "iterate")
        return h[:n] if self.batch else h[:n, 0]

    TILE_SHAPES = [(4, 4, 16), (2, 8, 16), (4, 8, 8), (8, 4, 8), (8, 8, 4)]   # fv3d_set_tiles numbering
    # the convective residual's streaming configurations, numbered after
    # TILE_SHAPES (5 + n): column tile (TJ, TK) and i-chunk length
    STREAM_SHAPES = [(16, 16, 4), (16, 16, 2), (8, 32, 4), (8, 32, 2), (16, 16, 8)]
    CONV_FACES = 10   # the convective residual one face per thread (the default on small coarse levels)

    @classmethod
    def conv_shape(cls, q: int):
        """Name of convective configuration q: a tile shape, ("stream", TJ, TK, ci) or ("faces",)."""
        if q == cls.CONV_FACES:
            return ("faces",)
        return cls.TILE_SHAPES[q] if q < len(cls.TILE_SHAPES) else ("stream",) + cls.STREAM_SHAPES[q - 5]

    def set_tiles(self, shapes):
        """Configuration numbers of the convective (TILE_SHAPES, then
        STREAM_SHAPES from 5 on, then CONV_FACES), viscous and nu~ (TILE_SHAPES)
        residual kernels."""
        a = np.ascontiguousarray(shapes, dtype=np.int32)
        self._check(self.L.fv3d_set_tiles(self.h, a.ctypes.data), "set_tiles")

    def autotune(self, reps: int = 3):
        """Times every tile shape of the residual kernels and keeps the fastest;
        returns the chosen shape numbers (-1 for a kernel not in use)."""
        a = np.zeros(3, dtype=np.int32)
        self._check(self.L.fv3d_autotune(self.h, reps, a.ctypes.data), "autotune")
        return [int(x) for x in a]

    # -- learned closure (fv3d_mlp_*)
    MLP_PARAMS = 8833

    def mlp_set(self, params, bounds=None):
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == self.MLP_PARAMS
        b = None if bounds is None else np.ascontiguousarray(bounds, dtype=np.float32)
        self._check(self.L.fv3d_mlp_set(self.h, p.ctypes.data, None if b is None else b.ctypes.data), "mlp_set")

    def mlp_get(self) -> np.ndarray:
        p = np.zeros(self.MLP_PARAMS, dtype=np.float32)
        self._check(self.L.fv3d_mlp_get(self.h, p.ctypes.data), "mlp_get")
        return p

    def mlp_features(self, block: int, inst: int = 1) -> np.ndarray:
        bi = self.cases[inst - 1].blocks[block - 1]
        out = np.zeros((bi.im - 1, bi.jm - 1, bi.km - 1, 6), dtype=np.float32)
        n = self.L.fv3d_mlp_features(self.h, self.offsets[inst - 1] + block, out.ctypes.data)
        self._check(n, "mlp_features")
        return out

    def mlp_targets(self, block: int, inst: int = 1) -> np.ndarray:
        bi = self.cases[inst - 1].blocks[block - 1]
        out = np.zeros((bi.im - 1, bi.jm - 1, bi.km - 1), dtype=np.float32)
        n = self.L.fv3d_mlp_targets(self.h, self.offsets[inst - 1] + block, out.ctypes.data)
        self._check(n, "mlp_targets")
        return out

    def mlp_forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros(x.shape[0], dtype=np.float32)
        self._check(self.L.fv3d_mlp_forward(self.h, x.ctypes.data, x.shape[0], y.ctypes.data), "mlp_forward")
        return y

    def mlp_grad(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(self.MLP_PARAMS, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        self._check(self.L.fv3d_mlp_grad(self.h, x.ctypes.data, t.ctypes.data, x.shape[0], g.ctypes.data,
                                         loss.ctypes.data), "mlp_grad")
        return g, float(loss[0])

    def mlp_train(self, x, t, steps: int, batch: int, lr: float) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        h = np.zeros(steps, dtype=np.float32)
        self._check(self.L.fv3d_mlp_train(self.h, x.ctypes.data, t.ctypes.data, x.shape[0], steps, batch,
                                          C.c_float(lr), h.ctypes.data), "mlp_train")
        return h

    def mlp_closure(self, on: bool):
        self._check(self.L.fv3d_mlp_closure(self.h, int(on)), "mlp_closure")

    def mlp_relax(self, omega: float):
        self._check(self.L.fv3d_mlp_relax(self.h, float(omega)), "mlp_relax")

    def mlp_tripped(self) -> int:
        return self._check(int(self.L.fv3d_mlp_tripped(self.h)), "mlp_tripped")

    def instance_results(self, inst: int):
        """(res6, perf5) of instance inst (1-based) after the last iteration."""
        r, p = np.zeros(6), np.zeros(5)
        self._check(self.L.fv3d_instance_results(self.h, inst, r.ctypes.data, p.ctypes.data), "instance_results")
        return r, p

    def res_calculate(self) -> np.ndarray:
        r = np.zeros(6)
        self._check(self.L.fv3d_res_calculate(self.h, r.ctypes.data), "res_calculate")
        return r

    def perf(self) -> np.ndarray:
        r = np.zeros(5)
        self._check(self.L.fv3d_perf(self.h, r.ctypes.data), "perf")
        return r

    def sync(self):
        self._check(self.L.fv3d_sync(self.h), "sync")

    def set_timing(self, on: bool):
        self._check(self.L.fv3d_set_timing(self.h, int(on)), "set_timing")

    TIMING_SLOTS = 12   # FV3D_TIMING_SLOTS

    @property
    def levels(self) -> int:
        return self._check(self.L.fv3d_levels(self.h), "levels")

    def timing(self):
        ms = np.zeros(self.TIMING_SLOTS)
        n = np.zeros(self.TIMING_SLOTS, dtype=np.int64)
        self._check(self.L.fv3d_timing(self.h, ms.ctypes.data, n.ctypes.data), "timing")
        return ms, n

    @property
    def n_cells(self) -> int:
        return int(self.L.fv3d_cell_count(self.h))

    def get(self, block: int, name: str, level: int = 1, inst: int = 1) -> np.ndarray:
        """A field of one block (of instance inst); level 2 reads the coarse
        level of the two-grid cycle."""
        bi = self.cases[inst - 1].blocks[block - 1]
        gblock = self.offsets[inst - 1] + block
        im, jm, km = bi.im, bi.jm, bi.km
        if level == 2:
            im, jm, km = (im + 1) // 2, (jm + 1) // 2, (km + 1) // 2
        ni, nj, nk = im - 1, jm - 1, km - 1
        if name in ("cv", "pv"):
            shape = (im + 1, jm + 1, km + 1, 5 if name == "cv" else 6)
        elif name in ("mut", "ev"):
            shape = (im + 1, jm + 1, km + 1)
        elif name in ("res", "dq", "fv"):
            shape = (ni, nj, nk, 5)
        elif name in ("fcI", "fcJ", "fcK", "fvI", "fvJ", "fvK"):
            shape = {"I": (im, nj, nk, 5), "J": (ni, jm, nk, 5), "K": (ni, nj, km, 5)}[name[2]]
        elif name in ("gradI", "gradJ", "gradK"):
            shape = {"I": (im, nj, nk, 12), "J": (ni, jm, nk, 12), "K": (ni, nj, km, 12)}[name[4]]
        elif name in ("dt", "lamcI", "lamcJ", "lamcK", "lamvI", "lamvJ", "lamvK", "omega", "vol", "resev", "dev", "dwall"):
            shape = (ni, nj, nk)
        elif name.startswith("d") and name[1] in "IJK":
            c = 6 if name.endswith("_pv") else 5
            shape = {"I": (2, nj, nk, c), "J": (ni, 2, nk, c), "K": (ni, nj, 2, c)}[name[1]]
        else:
            raise KeyError(name)
        out = np.zeros(shape)
        n = self.L.fv3d_level_field_download(self.h, level, gblock, name.encode(), out.ctypes.data)
        self._check(n, "field_download " + name)
        if n != out.size:
            raise _lib.Fv3dError(f"field {name}: got {n} values, expected {out.size}")
        return out
