"""Synthetic cases written in the reference solver's own file formats.

No grid ships with the reference (its Rotor 67 grid is missing), so every case
is generated here: node coordinates, patch tables, ``config.dat`` and a
restart file that carries the seeded initial field.

Index conventions follow the reference: inlet and outlet must be K-faces
(source/BndCondInlet.cpp:34-63, source/BndCondOutlet.cpp:233-250), so the
streamwise direction is index k and physical z (the rotation axis, see
source/Flux_Split_Scheme.cpp:737-738).
"""
from __future__ import annotations

import dataclasses
from typing import Dict, List

import numpy as np

from .refio import BlockInfo, Config, Patch, cfg_float, write_case


@dataclasses.dataclass
class Case:
    cfg: Config
    nodes: List[np.ndarray]                 # per block (IM, JM, KM, 3)
    patches: List[Patch]
    blocks: List[BlockInfo]
    states: List[Dict[str, np.ndarray]]      # per block restart fields (ni, nj, nk)

    def write(self, dirpath: str) -> None:
        write_case(dirpath, self.cfg, self.nodes, self.patches, self.blocks, self.states)

    @property
    def n_cells(self) -> int:
        return sum((b.im - 1) * (b.jm - 1) * (b.km - 1) for b in self.blocks)


def nondim_constants(cfg: Config) -> Dict[str, float]:
    """Non-dimensional gas constants (source/UserInput.cpp:1070-1092)."""
    gamma = cfg_float(cfg.gamma)
    return {
        "gamma": gamma,
        "rgas": 1.0 / (1.0 * 1.0 * gamma),
        "pt": 1.0 / gamma,
        "tt": 1.0,
    }


def uniform_state(cfg: Config, shape, inlet_dir=None, seed: int | None = None,
                  amp: float = 1.0e-3) -> Dict[str, np.ndarray]:
    """Uniform field at ``Ma0`` (source/Initialization.cpp:17-76), then a seeded
    ``rho *= 1 + U(-amp, amp)`` with p and velocity held and T, rhoE made
    consistent with the perturbed density."""
    c = nondim_constants(cfg)
    g, r = c["gamma"], c["rgas"]
    ma0 = cfg_float(cfg.ma0)
    a = np.array(inlet_dir if inlet_dir is not None else cfg.inlet_angle, dtype=np.float64)
    press = c["pt"] * (1.0 + (g - 1.0) / 2.0 * ma0 * ma0) ** (0.0 - g / (g - 1.0))
    temp = c["tt"] * (1.0 + (g - 1.0) / 2.0 * ma0 * ma0) ** (0.0 - 1.0)
    dens = press / (r * temp)
    vel = (g * r * temp) ** 0.5 * ma0
    n_m = float(np.sqrt(a @ a))
    rho = np.full(shape, dens)
    if seed is not None:
        rng = np.random.default_rng(seed)
        rho = rho * (1.0 + rng.uniform(-amp, amp, size=shape))
    u = np.full(shape, vel * a[0] / n_m)
    v = np.full(shape, vel * a[1] / n_m)
    w = np.full(shape, vel * a[2] / n_m)
    p = np.full(shape, press)
    t = p / (rho * r)
    e = p / (g - 1.0) + rho * (u * u + v * v + w * w) / 2.0
    return {"p": p, "u": u, "v": v, "w": w, "T": t, "rho": rho, "rhoE": e}


def _tanh_cluster(n: int, beta: float) -> np.ndarray:
    """n+1 points on [0,1] clustered at both ends."""
    xi = np.linspace(0.0, 1.0, n + 1)
    if beta <= 0.0:
        return xi
    return 0.5 * (1.0 + np.tanh(beta * (2.0 * xi - 1.0)) / np.tanh(beta))


def bump_channel(ni: int = 32, nj: int = 32, nk: int = 96, seed: int = 1, beta: float = 2.0,
                 bump_height: float = 0.1, cfl: float = 5.0, max_iter: int = 500,
                 flux: int = 300, recon_scheme: int = 500, recon_var: int = 200,
                 p_ratio: float = 0.9, n_threads: int = 1, viscous: bool = False,
                 len_ref: float = 1.0) -> Case:
    """Channel with a circular-arc bump on the lower wall, spanwise uniform.

    Streamwise z in [-1, 2] chords (index k), wall-normal y in [y_bump(z), 1]
    (index j, tanh-clustered at both walls), span x in [0, 0.5] (index i).
    Inlet k=1 (total conditions), outlet k=KM at ``p_ratio * Pt``, slip/no-slip
    walls on the other four faces.  Inviscid Euler by default (``viscous``
    selects laminar Navier-Stokes; the grid is in units of ``len_ref`` metres,
    which sets the Reynolds number), Roe + MUSCL3 on primitive variables,
    LU-SGS with local time step.
    """
    im, jm, km = ni + 1, nj + 1, nk + 1
    z = np.linspace(-1.0, 2.0, km)
    h = bump_height
    rad = (0.25 + h * h) / (2.0 * h)
    ylow = np.where((z >= 0.0) & (z <= 1.0), (h - rad) + np.sqrt(np.maximum(rad * rad - (z - 0.5) ** 2, 0.0)), 0.0)
    ylow = np.maximum(ylow, 0.0)
    eta = _tanh_cluster(nj, beta)
    x = np.linspace(0.0, 0.5, im)
    xyz = np.empty((im, jm, km, 3))
    xyz[..., 0] = x[:, None, None]
    xyz[..., 1] = ylow[None, None, :] + (1.0 - ylow[None, None, :]) * eta[None, :, None]
    xyz[..., 2] = z[None, None, :]

    cfg = Config(grid_file="grid.dat", turbo_type="G", equation="N" if viscous else "E", len_ref=len_ref,
                 cfl=cfl, max_iter=max_iter,
                 flux_scheme=flux, recon_scheme=recon_scheme, recon_var=recon_var,
                 p_out=p_ratio * 101325.0, n_threads=n_threads)
    patches = [
        Patch("inlet", 1, 5, (1, ni, 1, nj, 1, 1)),
        Patch("outlet", 1, 6, (1, ni, 1, nj, km, km)),
        Patch("wall", 1, 3, (1, ni, 1, 1, 1, nk)),
        Patch("wall", 1, 4, (1, ni, jm, jm, 1, nk)),
        Patch("wall", 1, 1, (1, 1, 1, nj, 1, nk)),
        Patch("wall", 1, 2, (im, im, 1, nj, 1, nk)),
    ]
    blocks = [BlockInfo(1, im, jm, km, 1, 0.0)]
    states = [uniform_state(cfg, (ni, nj, nk), seed=seed)]
    return Case(cfg, [xyz], patches, blocks, states)


def with_wall_shear(case: Case, delta: float = 0.3) -> Case:
    """Scales the velocity of every block's restart state by a wall profile
    tanh(d_i/delta) * tanh(d_j/delta), with d the normalised index distance to
    the nearer I and J faces, so a viscous or turbulence-model case starts
    with shear layers at the walls.  Pressure and density are kept; T and rhoE
    follow."""
    c = nondim_constants(case.cfg)
    g, r = c["gamma"], c["rgas"]
    for st in case.states:
        ni, nj, nk = st["rho"].shape
        di = np.minimum(np.arange(ni) + 0.5, ni - np.arange(ni) - 0.5) / ni
        dj = np.minimum(np.arange(nj) + 0.5, nj - np.arange(nj) - 0.5) / nj
        f = (np.tanh(di / delta)[:, None] * np.tanh(dj / delta)[None, :])[..., None]
        for q in ("u", "v", "w"):
            st[q] = st[q] * f
        st["T"] = st["p"] / (st["rho"] * r)
        st["rhoE"] = st["p"] / (g - 1.0) + st["rho"] * (st["u"] ** 2 + st["v"] ** 2 + st["w"] ** 2) / 2.0
    return case


def from_meta(meta: Dict) -> Case:
    """The case a fixture's metadata describes: ``kwargs`` for bump_channel and
    the iteration count, optionally ``wall_shear`` (with_wall_shear) and
    ``turbulence`` (the config's turbulence model code)."""
    case = bump_channel(max_iter=meta["iterations"], **meta["kwargs"])
    if meta.get("wall_shear"):
        case = with_wall_shear(case)
    if meta.get("turbulence"):
        case.cfg.turbulence = meta["turbulence"]
    return case
