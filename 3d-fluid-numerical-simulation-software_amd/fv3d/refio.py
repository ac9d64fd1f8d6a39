"""Reader/writer for the reference solver's case files.

The reference reads four text files from its working directory:

* ``config.dat`` -- positional, one value per line, ``#`` starts a comment
  (source/UserInput.cpp:17-239, include/fstringIO.h:15-42);
* the grid file named on config line 2 -- ``NB``, then per block its id,
  ``IM JM KM`` and ``KM*JM*IM`` node lines ``x y z`` with i fastest
  (source/UserInput.cpp:241-286);
* the boundary-definition file named on config line 5 -- patch tables for
  inlet, outlet, wall, periodic, interior and mixing-plane patches, then the
  block table ``BID IM JM KM NBlade rpm`` (source/UserInput.cpp:336-446);
* ``RestartSolution.dat`` -- one line per cell ``nb i j k p u v w T rho rhoE
  [nu~]`` with i fastest (source/UserInput.cpp:1118-1159).

Values are written with 17 significant digits so a reader that parses doubles
gets the exact bits back.  Config values are parsed by the reference with
``stof`` (float); :func:`cfg_float` reproduces that rounding.
"""
from __future__ import annotations

import dataclasses
import os
import struct
from typing import Dict, List, Sequence

import numpy as np


def cfg_float(x: float) -> float:
    """Round like ``std::stof`` followed by widening to double."""
    return struct.unpack("f", struct.pack("f", float(x)))[0]


def _g(x: float) -> str:
    return "%.17g" % float(x)


# --------------------------------------------------------------------------
# config.dat
# --------------------------------------------------------------------------

@dataclasses.dataclass
class Config:
    """Every field of ``config.dat`` in file order (source/UserInput.cpp:17-239)."""
    grid_file: str = "grid.dat"
    bc_file: str = "BoundaryDefinition.dat"
    turbo_type: str = "G"          # A axial / C centrifugal / G general outlet
    equation: str = "E"            # E Euler / N Navier-Stokes
    initial: str = "U"             # U uniform / I interpolate
    rgas: float = 287.06
    gamma: float = 1.4
    pr_lam: float = 0.72
    pr_tur: float = 0.9
    pt_in: float = 101325.0
    tt_in: float = 288.15
    p_out: float = 91192.5
    p21_ratio: float = 1.0
    ma0: float = 0.5
    len_ref: float = 1.0
    ev_ratio: float = 3.0
    inlet_angle: Sequence[float] = (0.0, 0.0, 1.0)
    outlet_angle: Sequence[float] = (0.0, 0.0, 0.0)
    mass_imposed: str = "N"
    mass_imposed_params: str = "33.8 1.0 100 0.0001"
    max_iter: int = 500
    coarse_iter: int = 0
    out_step: int = 100000000
    conv_tol: float = 1.0e-12
    n_threads: int = 1
    use_multigrid: str = "N"
    reuse_solution: str = "Y"
    reuse_wall_distance: str = "N"
    cfl: float = 5.0
    c_timestep: float = 1.0
    relax_factor: float = 1.0
    eps_entr: float = 0.05
    k2: float = 0.6
    k4: float = 100.0
    irs: str = "N"
    limiter: str = "Y"
    timestep: str = "L"
    recon_var: int = 200           # 100 conservative / 200 primitive / 300 characteristic
    recon_scheme: int = 500        # 0 / 100 NND / 200 MUSCL2u / 300 MUSCL2c / 400 MUSCL2o / 500 MUSCL3
    flux_scheme: int = 300         # 200 Steger-Warming / 300 Roe
    time_scheme: int = 500         # 500 LU-SGS
    turbulence: int = 0            # 0 none / 100 Baldwin-Lomax / 200 Spalart-Allmaras
    # not in config.dat: grid levels with use_multigrid (0/2 the reference's
    # two-grid cycle, 3-5 the N-level V-cycle of this library)
    mg_levels: int = 0

    def lines(self) -> List[str]:
        a = self.inlet_angle
        b = self.outlet_angle
        return [
            "# grid file", self.grid_file, "#", "# boundary condition file", self.bc_file,
            "#", "# Physics - general", "#--------------------------",
            f"{self.turbo_type}       # axial(A) / centrifugal(C) / general(G)",
            f"{self.equation}       # Euler(E) / Navier-Stokes(N)",
            f"{self.initial}       # uniform(U) / interpolate(I)",
            f"{self.rgas!r}    # Gas_Constant", f"{self.gamma!r}    # Specific_Heat_Ratio",
            f"{self.pr_lam!r}    # laminar Prandtl number", f"{self.pr_tur!r}    # turbulence Prandtl number",
            "#", "# Physics - internal flow", "#--------------------------",
            f"{self.pt_in!r}    # total pressure at inlet [Pa]",
            f"{self.tt_in!r}    # total temperature at inlet [K]",
            f"{self.p_out!r}    # static pressure at outlet [Pa]",
            f"{self.p21_ratio!r}    # approximate ratio of inlet to outlet static pressure",
            f"{self.ma0!r}    # Mach number for initial",
            f"{self.len_ref!r}    # reference length [m]",
            f"{self.ev_ratio!r}    # reference eddy viscous ratio",
            f"{a[0]!r} {a[1]!r} {a[2]!r}    # inlet flow direction",
            f"{b[0]!r} {b[1]!r} {b[2]!r}    # outlet flow direction",
            f"{self.mass_imposed}    # use flow mass outlet",
            f"{self.mass_imposed_params}    # mass imposed, relax, iter, res",
            "#", "# Iteration control", "# -------------------------",
            f"{self.max_iter}    # max. number of iterations",
            f"{self.coarse_iter}    # number of iterations on coarse grid",
            f"{self.out_step}    # number of iterations for save",
            f"{self.conv_tol!r}    # convergence tolerance",
            f"{self.n_threads}    # number of threads",
            f"{self.use_multigrid}    # use multigrid method",
            f"{self.reuse_solution}    # use previous solution for restart",
            f"{self.reuse_wall_distance}    # use previous wall distance",
            "#", "# Numerical parameters", "# -------------------------",
            f"{self.cfl!r}    # CFL-number", f"{self.c_timestep!r}    # time step coefficient",
            f"{self.relax_factor!r}    # relax factor in LU_SGS",
            f"{self.eps_entr!r}    # entropy correction coefficient",
            f"{self.k2!r}    # k2", f"{self.k4!r}    # k4",
            f"{self.irs}    # implicit residual smoothing",
            f"{self.limiter}    # limit flow variables",
            f"{self.timestep}    # local(L) / global(G) time-stepping",
            f"{self.recon_var}    # reconstruction variables",
            f"{self.recon_scheme}    # reconstruction scheme",
            f"{self.flux_scheme}    # space discrete scheme",
            f"{self.time_scheme}    # time marching scheme",
            f"{self.turbulence}    # turbulence model",
        ]


def _strip(line: str) -> str:
    """Comment/space stripping of include/fstringIO.h:15-42 (tabs are kept)."""
    ic = line.find("#")
    if ic >= 0:
        line = line[:ic]
    return line.strip(" ").rstrip("\r\n").strip(" ")


def read_config(path: str) -> Config:
    with open(path) as f:
        raw = [_strip(l.rstrip("\n")) for l in f]
    c = Config()
    v = lambda n: raw[n - 1].strip()
    c.grid_file, c.bc_file = v(2), v(5)
    c.turbo_type, c.equation, c.initial = v(9), v(10), v(11)
    c.rgas, c.gamma, c.pr_lam, c.pr_tur = (float(v(n)) for n in (12, 13, 14, 15))
    c.pt_in, c.tt_in, c.p_out, c.p21_ratio, c.ma0, c.len_ref, c.ev_ratio = (
        float(v(n)) for n in range(19, 26))
    c.inlet_angle = tuple(float(t) for t in v(26).split())
    c.outlet_angle = tuple(float(t) for t in v(27).split())
    c.mass_imposed, c.mass_imposed_params = v(28), v(29)
    c.max_iter, c.coarse_iter, c.out_step = int(v(33)), int(v(34)), int(v(35))
    c.conv_tol, c.n_threads = float(v(36)), int(v(37))
    c.use_multigrid, c.reuse_solution, c.reuse_wall_distance = v(38), v(39), v(40)
    c.cfl, c.c_timestep, c.relax_factor, c.eps_entr, c.k2, c.k4 = (
        float(v(n)) for n in range(44, 50))
    c.irs, c.limiter, c.timestep = v(50), v(51), v(52)
    c.recon_var, c.recon_scheme, c.flux_scheme, c.time_scheme, c.turbulence = (
        int(v(n).split()[0]) for n in range(53, 58))
    return c


def write_config(path: str, cfg: Config) -> None:
    with open(path, "w") as f:
        f.write("\n".join(cfg.lines()) + "\n")


# --------------------------------------------------------------------------
# boundary definition
# --------------------------------------------------------------------------

@dataclasses.dataclass
class Patch:
    """One row of a patch table.

    ``kind`` is one of inlet, outlet, wall, periodic, interior, mixing.
    ``face`` is the reference face type 1..6 = i-, i+, j-, j+, k-, k+.
    ``rng`` holds ``is ie js je ks ke``: the normal index is the node (face)
    index, the two tangential ranges are cell ranges (source/BndCondWall.cpp:59-78).
    Two-sided patches carry ``face2``/``block2``/``rng2`` and ``orient``.
    """
    kind: str
    block: int
    face: int
    rng: Sequence[int]
    rotate: int = 0
    orient: int = 0
    block2: int = 0
    face2: int = 0
    rng2: Sequence[int] = ()


@dataclasses.dataclass
class BlockInfo:
    bid: int
    im: int
    jm: int
    km: int
    nblade: int = 1
    rpm: float = 0.0


def write_bc(path: str, patches: List[Patch], blocks: List[BlockInfo]) -> None:
    by = {k: [p for p in patches if p.kind == k]
          for k in ("inlet", "outlet", "wall", "periodic", "interior", "mixing")}
    out = ["# Boundary definition", "#"]

    def section(title, rows, header):
        out.extend([f"# {title}", "#---------------------------"])
        out.append(f" {len(rows)}\t\t# number of {title}s")
        out.append(header)
        out.extend(rows)
        out.append("#")

    r = lambda q: " ".join(str(int(x)) for x in q)
    section("inlet", [f"{n+1} {p.block} {p.face} {r(p.rng)}" for n, p in enumerate(by["inlet"])],
            "#ID_BID FT_Ista_Iend_Jsta_Jend_Ksta_Kend")
    section("outlet", [f"{n+1} {p.block} {p.face} {r(p.rng)}" for n, p in enumerate(by["outlet"])],
            "#ID_BID_FT_Ista_Iend_Jsta_Jend_Ksta_Kend")
    section("wall", [f"{n+1} {p.block} {p.face} {p.rotate} {r(p.rng)}" for n, p in enumerate(by["wall"])],
            "#ID_BID_FT_wrot_Ista_Iend_Jsta_Jend_Ksta_Kend")
    for kind, title in (("periodic", "periodic"), ("interior", "interior"), ("mixing", "mixing plane")):
        section(title, [f"{n+1} {p.orient} {p.block} {p.face} {r(p.rng)} {p.block2} {p.face2} {r(p.rng2)}"
                        for n, p in enumerate(by[kind])],
                "#ID_ori_BID1_FT1_Ista1_Iend1_Jsta1_Jend1_Ksta1_Kend1_BID2_FT2_Ista2_Iend2_Jsta2_Jend2_Ksta2_Kend2")
    out = out[:-1]
    out.extend(["#", "# Blocks", "#--------------------------", f"{len(blocks)}\t\t# number of Blocks",
                "#BID_IM_JM_KM_Nblade_rpm"])
    out.extend(f"{b.bid} {b.im} {b.jm} {b.km} {b.nblade} {_g(b.rpm)}" for b in blocks)
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def read_bc(path: str):
    with open(path) as f:
        lines = [_strip(l.rstrip("\n")) for l in f]
    pos = [0]

    def nxt():
        s = lines[pos[0]]
        pos[0] += 1
        return s

    patches: List[Patch] = []
    for _ in range(4):
        nxt()
    for kind, nskip in (("inlet", 0), ("outlet", 3), ("wall", 3), ("periodic", 3),
                        ("interior", 3), ("mixing", 3)):
        for _ in range(nskip):
            nxt()
        n = int(nxt().split()[0])
        nxt()
        for _ in range(n):
            t = [int(float(x)) for x in nxt().split()]
            if kind in ("inlet", "outlet"):
                patches.append(Patch(kind, t[1], t[2], t[3:9]))
            elif kind == "wall":
                patches.append(Patch(kind, t[1], t[2], t[4:10], rotate=t[3]))
            else:
                patches.append(Patch(kind, t[2], t[3], t[4:10], orient=t[1], block2=t[10],
                                     face2=t[11], rng2=t[12:18]))
    for _ in range(3):
        nxt()
    nb = int(nxt().split()[0])
    nxt()
    blocks = []
    for _ in range(nb):
        t = nxt().split()
        blocks.append(BlockInfo(int(t[0]), int(t[1]), int(t[2]), int(t[3]), int(t[4]), float(t[5])))
    return patches, blocks


# --------------------------------------------------------------------------
# grid and restart files
# --------------------------------------------------------------------------

def write_grid(path: str, nodes: List[np.ndarray]) -> None:
    """``nodes[b]`` has shape (IM, JM, KM, 3): node (i,j,k) 1-based at [i-1,j-1,k-1]."""
    with open(path, "w") as f:
        f.write(f"{len(nodes)}\n")
        for b, xyz in enumerate(nodes):
            im, jm, km, _ = xyz.shape
            f.write(f"{b+1}\n{im} {jm} {km}\n")
            flat = np.ascontiguousarray(xyz.transpose(2, 1, 0, 3)).reshape(-1, 3)
            np.savetxt(f, flat, fmt="%.17g")


def read_grid(path: str) -> List[np.ndarray]:
    out = []
    with open(path) as f:
        nb = int(_strip(f.readline()).split()[0])
        for _ in range(nb):
            f.readline()
            im, jm, km = (int(x) for x in _strip(f.readline()).split()[:3])
            data = np.loadtxt(f, max_rows=im * jm * km).reshape(km, jm, im, 3)
            out.append(np.ascontiguousarray(data.transpose(2, 1, 0, 3)))
    return out


def write_restart(path: str, states: List[Dict[str, np.ndarray]]) -> None:
    """``states[b]`` maps p,u,v,w,T,rho,rhoE[,nut] to arrays of shape (ni,nj,nk)."""
    with open(path, "w") as f:
        for b, st in enumerate(states):
            ni, nj, nk = st["rho"].shape
            keys = ["p", "u", "v", "w", "T", "rho", "rhoE"] + (["nut"] if "nut" in st else [])
            ii, jj, kk = np.meshgrid(np.arange(1, ni + 1), np.arange(1, nj + 1), np.arange(1, nk + 1),
                                     indexing="ij")
            cols = [np.full(ni * nj * nk, b + 1)] + [a.transpose(2, 1, 0).ravel() for a in (ii, jj, kk)]
            vals = [st[k].transpose(2, 1, 0).ravel() for k in keys]
            for row in zip(*(cols + vals)):
                f.write(" ".join(str(int(x)) for x in row[:4]) + " " +
                        " ".join(_g(x) for x in row[4:]) + "\n")


def read_restart(path: str, dims: List[Sequence[int]], with_nut: bool = False):
    data = np.loadtxt(path)
    out, off = [], 0
    keys = ["p", "u", "v", "w", "T", "rho", "rhoE"] + (["nut"] if with_nut else [])
    for (ni, nj, nk) in dims:
        n = ni * nj * nk
        blk = data[off:off + n]
        off += n
        out.append({k: np.ascontiguousarray(blk[:, 4 + c].reshape(nk, nj, ni).transpose(2, 1, 0))
                    for c, k in enumerate(keys)})
    return out


def write_case(dirpath: str, cfg: Config, nodes: List[np.ndarray], patches: List[Patch],
               blocks: List[BlockInfo], states: List[Dict[str, np.ndarray]] | None) -> None:
    os.makedirs(dirpath, exist_ok=True)
    write_config(os.path.join(dirpath, "config.dat"), cfg)
    write_grid(os.path.join(dirpath, cfg.grid_file), nodes)
    write_bc(os.path.join(dirpath, cfg.bc_file), patches, blocks)
    if states is not None:
        write_restart(os.path.join(dirpath, "RestartSolution.dat"), states)
