"""ctypes wrapper of the C restatement (oracle/_build/liboracle_fv3d.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, never by the product.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import subprocess
from typing import Dict, List, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle_fv3d.so")
REF_BIN = os.path.join(HERE, "_ref", "fv3d_ref")


class _Cfg(C.Structure):
    _fields_ = [("rgas", C.c_double), ("gamma", C.c_double), ("pr_lam", C.c_double), ("pr_tur", C.c_double),
                ("pt_in", C.c_double), ("tt_in", C.c_double), ("p_out", C.c_double), ("ma0", C.c_double),
                ("len_ref", C.c_double), ("ev_ratio", C.c_double), ("inlet_dir", C.c_double * 3),
                ("cfl", C.c_double), ("c_timestep", C.c_double), ("relax", C.c_double),
                ("turbo_type", C.c_int), ("viscous", C.c_int), ("recon_var", C.c_int),
                ("recon_scheme", C.c_int), ("flux_scheme", C.c_int), ("turbulence", C.c_int),
                ("limiter", C.c_int), ("global_dt", C.c_int), ("nthreads", C.c_int),
                ("multigrid", C.c_int), ("coarse_iter", C.c_int), ("initial_type", C.c_int),
                ("p21ratio", C.c_double), ("outlet_dir", C.c_double * 3)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library not built: {LIB_PATH} (run make -C oracle restatement)")
        L = C.CDLL(LIB_PATH)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(_Cfg)]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_block_add.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_double]
        L.orc_patch_add.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int]
        L.orc_mixing_plane_add.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p]
        L.orc_connection_add.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int,
                                         C.c_void_p]
        L.orc_set_state.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        for fn in ("orc_geometry", "orc_bc", "orc_spectral", "orc_timestep", "orc_convective", "orc_resflux",
                   "orc_lusgs"):
            getattr(L, fn).argtypes = [C.c_void_p]
        L.orc_res_calculate.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_perf.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_cfl.argtypes = [C.c_void_p]
        L.orc_cfl.restype = C.c_double
        L.orc_iterate.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.orc_get.argtypes = [C.c_void_p, C.c_int, C.c_char_p, C.c_void_p]
        L.orc_get.restype = C.c_long
        L.orc_test_flux.argtypes = [C.c_int, C.c_double, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p]
        L.orc_test_recon.argtypes = [C.c_int, C.c_int, C.c_double] + [C.c_void_p] * 6
        L.orc_test_viscous.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.orc_test_sutherland.argtypes = [C.c_void_p, C.c_double]
        L.orc_test_sutherland.restype = C.c_double
        L.orc_levels.argtypes = [C.c_void_p]
        L.orc_coarse.argtypes = [C.c_void_p]
        L.orc_coarse.restype = C.c_void_p
        L.orc_initial.argtypes = [C.c_void_p]
        L.orc_set_nut.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.orc_stats.argtypes = [C.c_void_p]
        L.orc_set_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


_TURBO = {"A": 0, "C": 1}
_KIND = {"inlet": 0, "outlet": 1, "wall": 2}


def config_struct(cfg) -> _Cfg:
    s = _Cfg()
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
    s.nthreads = 1
    s.multigrid = 1 if cfg.use_multigrid.upper() == "Y" else 0
    s.coarse_iter = cfg.coarse_iter
    s.initial_type = 0 if cfg.initial.upper() == "U" else 1
    s.p21ratio = cfg.p21_ratio
    for c in range(3):
        s.outlet_dir[c] = cfg.outlet_angle[c]
    return s


class Oracle:
    """The CPU restatement loaded with one case (fv3d.cases.Case)."""

    def __init__(self, case):
        self.case = case
        L = lib()
        self._cfg = config_struct(case.cfg)
        self.h = L.orc_create(C.byref(self._cfg))
        for b, xyz in enumerate(case.nodes):
            bi = case.blocks[b]
            a = np.ascontiguousarray(xyz, dtype=np.float64)
            L.orc_block_add(self.h, bi.im, bi.jm, bi.km, a.ctypes.data, bi.nblade, bi.rpm)
        for p in case.patches:
            if p.kind == "mixing":
                r1 = np.ascontiguousarray(p.rng, dtype=np.int32)
                r2 = np.ascontiguousarray(p.rng2, dtype=np.int32)
                if L.orc_mixing_plane_add(self.h, p.block, r1.ctypes.data, p.block2, r2.ctypes.data) != 0:
                    raise ValueError(f"oracle rejected {p}")
                continue
            if p.kind in ("interior", "periodic"):
                r1 = np.ascontiguousarray(p.rng, dtype=np.int32)
                r2 = np.ascontiguousarray(p.rng2, dtype=np.int32)
                if L.orc_connection_add(self.h, 0 if p.kind == "interior" else 1, p.orient, p.block, p.face,
                                        r1.ctypes.data, p.block2, p.face2, r2.ctypes.data) != 0:
                    raise ValueError(f"oracle rejected {p}")
                continue
            if p.kind not in _KIND:
                raise NotImplementedError(f"oracle patch kind {p.kind}")
            r = np.ascontiguousarray(p.rng, dtype=np.int32)
            L.orc_patch_add(self.h, _KIND[p.kind], p.block, p.face, r.ctypes.data, p.rotate)
        if L.orc_geometry(self.h) != 0:
            raise ValueError("oracle rejected the grid levels")
        if case.cfg.reuse_solution.upper() == "Y":
            if L.orc_levels(self.h) == 2:
                raise ValueError("two grid levels need the coarse start-up phase (reuse_solution = N)")
            for b, st in enumerate(case.states):
                pv = np.ascontiguousarray(np.stack([st["rho"], st["u"], st["v"], st["w"], st["p"], st["T"]], -1))
                e = np.ascontiguousarray(st["rhoE"])
                L.orc_set_state(self.h, b + 1, pv.ctypes.data, e.ctypes.data)
                if case.cfg.turbulence == 200:
                    ev = np.ascontiguousarray(st["nut"], dtype=np.float64)
                    L.orc_set_nut(self.h, b + 1, ev.ctypes.data)
        elif L.orc_initial(self.h) != 0:
            raise ValueError("start field needs an inlet and an outlet patch")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def call(self, stage: str) -> None:
        getattr(lib(), "orc_" + stage)(self.h)

    def iterate(self, n: int) -> np.ndarray:
        h = np.zeros((n, 7))
        lib().orc_iterate(self.h, n, h.ctypes.data)
        return h

    def res_calculate(self) -> np.ndarray:
        r = np.zeros(6)
        lib().orc_res_calculate(self.h, r.ctypes.data)
        return r

    def perf(self) -> np.ndarray:
        r = np.zeros(5)
        lib().orc_perf(self.h, r.ctypes.data)
        return r

    def viscous_on(self, block: int, pv_box: np.ndarray) -> None:
        """Curl, face gradients and viscous fluxes of a given primitive box
        (ghost cells included, shape (IM+1, JM+1, KM+1, 6)), no boundary conditions."""
        a = np.ascontiguousarray(pv_box, dtype=np.float64)
        lib().orc_test_viscous(self.h, block, a.ctypes.data)

    def sutherland(self, temp: float) -> float:
        return lib().orc_test_sutherland(self.h, temp)

    def levels(self) -> int:
        return lib().orc_levels(self.h)

    def get(self, block: int, name: str, level: int = 1) -> np.ndarray:
        """A field of one block; level 2 reads the coarse level of the two-grid cycle."""
        bi = self.case.blocks[block - 1]
        im, jm, km = bi.im, bi.jm, bi.km
        h = self.h
        if level == 2:
            h = lib().orc_coarse(self.h)
            if not h:
                raise ValueError("no coarse level")
            im, jm, km = (im + 1) // 2, (jm + 1) // 2, (km + 1) // 2
        ni, nj, nk = im - 1, jm - 1, km - 1
        if name in ("cv", "pv"):
            shape = (im + 1, jm + 1, km + 1, 5 if name == "cv" else 6)
        elif name in ("res", "dq", "fc", "fv"):
            shape = (ni, nj, nk, 5)
        elif name in ("fcI", "fcJ", "fcK", "fvI", "fvJ", "fvK"):
            shape = {"I": (im, nj, nk, 5), "J": (ni, jm, nk, 5), "K": (ni, nj, km, 5)}[name[2]]
        elif name in ("gradI", "gradJ", "gradK"):
            shape = {"I": (im, nj, nk, 12), "J": (ni, jm, nk, 12), "K": (ni, nj, km, 12)}[name[4]]
        elif name in ("metI", "metJ", "metK", "metC"):
            shape = {"I": (im, nj, nk, 9), "J": (ni, jm, nk, 9), "K": (ni, nj, km, 9), "C": (ni, nj, nk, 9)}[name[3]]
        elif name in ("mut", "ev"):
            shape = (im + 1, jm + 1, km + 1)
        elif name in ("resev", "dev", "dwall", "cfctur", "cfvtur"):
            shape = (ni, nj, nk)
        elif name in ("gevI", "gevJ", "gevK"):
            shape = {"I": (im, nj, nk, 3), "J": (ni, jm, nk, 3), "K": (ni, nj, km, 3)}[name[3]]
        elif name in ("SI", "SJ", "SK"):
            shape = {"SI": (im, nj, nk, 3), "SJ": (ni, jm, nk, 3), "SK": (ni, nj, km, 3)}[name]
        elif name in ("dt", "lamcI", "lamcJ", "lamcK", "lamvI", "lamvJ", "lamvK", "omega", "vol"):
            shape = (ni, nj, nk)
        elif name.startswith("d") and name[1] in "IJK":
            c = 6 if name.endswith("_pv") else 5
            shape = {"I": (2, nj, nk, c), "J": (ni, 2, nk, c), "K": (ni, nj, 2, c)}[name[1]]
        else:
            raise KeyError(name)
        out = np.zeros(shape)
        n = lib().orc_get(h, block, name.encode(), out.ctypes.data)
        if n != out.size:
            raise RuntimeError(f"orc_get({name}) returned {n}, expected {out.size}")
        return out


def set_threads(n: int = 0) -> int:
    """OpenMP threads of the restatement's cell and face loops (n > 0 sets
    them); returns the count in use.  Results are the same for any count."""
    return int(lib().orc_set_threads(n))


STAT_NAMES = ("recon_fallback", "limit_vars", "limit_vars_kn0", "limit_sa", "limit_sa_kn0", "low_dt")


def stats() -> Dict[str, int]:
    """Repair-branch counters of the restatement since the last call (then
    cleared): first-order reconstruction fallbacks, Limit_Vars replacements
    (and those with no admissible neighbour), Limit_SA replacements (and those
    with none), and blocks whose time-step damping Limit_Vars switched on."""
    out = (C.c_long * len(STAT_NAMES))()
    lib().orc_stats(out)
    return dict(zip(STAT_NAMES, list(out)))


def run_reference(case_dir: str, n_iter: int, out_dir: str, threads: int = 1, dump_first: bool = False,
                  dump: bool = True, timeout: Optional[float] = None, dump_coarse: bool = False) -> Dict:
    """Runs the compiled reference (oracle/_ref/fv3d_ref) on a case directory."""
    if not os.path.exists(REF_BIN):
        raise RuntimeError(f"reference harness not built: {REF_BIN}")
    os.makedirs(out_dir, exist_ok=True)
    args = [REF_BIN, case_dir, str(n_iter), out_dir, "--threads", str(threads), "--quiet"]
    if dump_first:
        args.append("--dump-first")
    if not dump:
        args.append("--no-dump")
    if dump_coarse:
        args.append("--dump-coarse")
    subprocess.run(args, check=True, timeout=timeout)
    hist = np.loadtxt(os.path.join(out_dir, "history.txt"), ndmin=2)
    with open(os.path.join(out_dir, "timing.json")) as f:
        timing = json.load(f)
    with open(os.path.join(out_dir, "levels.txt")) as f:
        levels = int(f.read())
    return {"history": hist, "timing": timing, "levels": levels}


def point_flux(scheme: int, gamma: float, ql, qr, que: float = 0.0) -> np.ndarray:
    """Face flux of the oracle (200 Steger-Warming, 300 Roe) on face-frame states."""
    a = np.ascontiguousarray(ql, dtype=np.float64)
    b = np.ascontiguousarray(qr, dtype=np.float64)
    out = np.zeros(5)
    if lib().orc_test_flux(scheme, gamma, a.ctypes.data, b.ctypes.data, que, out.ctypes.data) != 0:
        raise ValueError(f"flux scheme {scheme}")
    return out


def point_recon(var: int, scheme: int, gamma: float, ff, f, b, bb):
    """Left/right conservative states of one face from its stencil FF, F | B, BB."""
    arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (ff, f, b, bb)]
    l, r = np.zeros(5), np.zeros(5)
    if lib().orc_test_recon(var, scheme, gamma, *[x.ctypes.data for x in arrs], l.ctypes.data, r.ctypes.data) != 0:
        raise ValueError(f"reconstruction var {var}")
    return l, r
