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
from typing import Dict, List, Sequence

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


def with_sa(case: Case, amp: float = 0.5) -> Case:
    """Spalart-Allmaras on a case: turbulence model 200 and a restart nu~ of
    ev_ratio * (1 + amp * sin(2 pi i/ni) cos(2 pi k/nk)) per block, so the
    first iteration sees non-uniform gradients of nu~."""
    case.cfg.turbulence = 200
    for st in case.states:
        ni, nj, nk = st["rho"].shape
        a = np.sin(2.0 * np.pi * (np.arange(ni) + 0.5) / ni)[:, None, None]
        b = np.cos(2.0 * np.pi * (np.arange(nk) + 0.5) / nk)[None, None, :]
        st["nut"] = cfg_float(case.cfg.ev_ratio) * (1.0 + amp * a * b) * np.ones((ni, nj, nk))
    return case


def with_cell_values(case: Case, block: int, index, rho=None, p=None, u=None, nut=None) -> Case:
    """Overwrites restart values of block ``block`` (1-based) at ``index`` (a
    numpy index into the (ni, nj, nk) cell arrays, 0-based cells): density,
    pressure, x velocity and nu~ where given, with T and rhoE made consistent.
    Used to start a run from states outside the limiter's admissible range
    (source/Limit_FlowVars.cpp:16-20, 124) so that its repair branches run."""
    c = nondim_constants(case.cfg)
    g, r = c["gamma"], c["rgas"]
    st = case.states[block - 1]
    if rho is not None:
        st["rho"][index] = rho
    if p is not None:
        st["p"][index] = p
    if u is not None:
        st["u"][index] = u
    if nut is not None:
        st["nut"][index] = nut
    st["T"] = st["p"] / (st["rho"] * r)
    st["rhoE"] = st["p"] / (g - 1.0) + st["rho"] * (st["u"] ** 2 + st["v"] ** 2 + st["w"] ** 2) / 2.0
    return case


def passage_batch(n: int = 64, seed: int = 3, min_cells: float = 1e5, max_cells: float = 2e6,
                  turbulence: int = 200, design_rpm: float = 16000.0, design_p_ratio: float = 0.76,
                  max_iter: int = 500) -> List[Case]:
    """Independent passages of uneven size for a batched run: cell counts drawn
    log-uniform in [min_cells, max_cells], each passage the five-block
    cambered-blade family (64x64 cross-section and k-blocks 12/40/128/60/24 at
    1.08 M cells, every direction scaled by the cube root of the size ratio,
    even cell counts), outlet pressure design_p_ratio * Pt times a
    back-pressure factor drawn uniform in [0.95, 1.25], rotation design_rpm
    times a factor drawn uniform in [0.7, 1.05].  Navier-Stokes with
    Spalart-Allmaras (turbulence 200, nu~ = ev_ratio) or Baldwin-Lomax (100)."""
    rng = np.random.default_rng(seed)
    out = []
    base_k = (12, 40, 128, 60, 24)
    for q in range(n):
        cells = float(np.exp(rng.uniform(np.log(min_cells), np.log(max_cells))))
        bp = float(rng.uniform(0.95, 1.25))
        sp = float(rng.uniform(0.7, 1.05))
        f = (cells / (64.0 * 64.0 * sum(base_k))) ** (1.0 / 3.0)
        even = lambda x: max(4, 2 * int(round(x / 2.0)))
        s = even(64 * f)
        nk = tuple(even(k * f) for k in base_k)
        c = annular_passage(ni=s, nj=s, nk_blocks=nk, seed=seed * 1000 + q, rpm=design_rpm * sp, blade_blocks=(2,),
                            camber=0.1, viscous=True, len_ref=1.0, max_iter=max_iter,
                            z_extent=(-0.06, -0.035, -0.015, 0.035, 0.055, 0.08))
        c.cfg.p_out = design_p_ratio * 101325.0 * bp
        if turbulence == 200:
            c = with_sa(c, amp=0.0)
        else:
            c.cfg.turbulence = turbulence
        out.append(c)
    return out


def vane_passage(ni: int = 88, nk_blocks: Sequence[int] = (24, 48, 144, 72, 40), seed: int = 1,
                 turbulence: int = 200, levels: int = 0, coarse_iter: int = 2, p_ratio: float = 0.95,
                 camber: float = 0.1, max_iter: int = 500) -> Case:
    """The learned-closure passage: the five-block annular passage (hub 0.2 m,
    tip 0.25 m, 22 blades) with a turning vane in the middle block (the
    parabola theta = camber * pitch * t^2, which meets the axial inflow at
    zero incidence), ni x ni cells in cross-section and nk_blocks along the
    axis (88 x 88 x (24, 48, 144, 72, 40) = 2,540,032 cells, every size a
    multiple of 8 so the grid halves three times), Navier-Stokes with
    Spalart-Allmaras (nu~ = ev_ratio at the start) or Baldwin-Lomax, Roe +
    MUSCL3 on primitive variables, outlet p_ratio * Pt, stationary; multigrid
    from the uniform field on the coarse level with ``levels`` grid levels
    (0: the reference's two-grid cycle)."""
    c = annular_passage(ni=ni, nj=ni, nk_blocks=nk_blocks, seed=seed, rpm=0.0, blade_blocks=(2,), camber=camber,
                        viscous=True, len_ref=1.0, max_iter=max_iter, profile="turn", p_ratio=p_ratio,
                        z_extent=(-0.06, -0.035, -0.015, 0.035, 0.055, 0.08))
    if turbulence == 200:
        c = with_sa(c, amp=0.0)
    else:
        c.cfg.turbulence = turbulence
    c.cfg.use_multigrid, c.cfg.reuse_solution, c.cfg.initial = "Y", "N", "U"
    c.cfg.coarse_iter, c.cfg.mg_levels = coarse_iter, levels
    return c


def rotor_stator_stage(ni: int = 128, nj_rotor: int = 128, nj_stator: int = 96,
                       rotor_nk: Sequence[int] = (24, 40, 64, 120, 64, 40, 24),
                       stator_nk: Sequence[int] = (24, 48, 72, 144, 96, 64, 32),
                       nblade_rotor: int = 22, nblade_stator: int = 30, rpm: float = 3000.0,
                       camber_rotor: float = 0.04, camber_stator: float = -0.03, p_ratio: float = 0.95,
                       seed: int = 1, turbulence: int = 200, levels: int = 0, coarse_iter: int = 2,
                       max_iter: int = 500) -> Case:
    """A rotor row and a stator row joined by a mixing plane: two annular
    passages (hub 0.2 m, tip 0.25 m, the same radial cells) of nblade_rotor and
    nblade_stator blades, nj_rotor and nj_stator cells pitchwise, cut along k
    into the blocks of rotor_nk and stator_nk cells (128 x (128, 96) x
    (376, 480) = 12,058,624 cells in 14 blocks by default; every size a
    multiple of 8, so the grid halves three times).  The rotor spins at rpm
    with its hub; its blades (the middle block) follow the relative flow of
    the axial inflow at zero incidence on every radius (a helix at -omega/V_x,
    V_x from the isentropic outlet Mach number) with a turning vane of
    camber_rotor pitches on top; the stator's vane (its middle block) turns by
    camber_stator pitches.  (At 0.08 / -0.06 pitches the full-size stage's
    residual stops falling near 1.3e-4 of its first value, in the FP64 and the
    FP32-operator forms alike, while the uncambered stage and the 5M-cell one
    converge: the more heavily turned blades do not settle to a steady flow at
    that resolution.  0.04 / -0.03 reaches a 1e-6 drop.)  Blocks 1..len(rotor_nk) are the rotor row, the
    rest the stator row; the rotor's last block (K high) meets the stator's
    first (K low) at the mixing plane (source/BndCondMixingPlane.cpp:287-522).
    Navier-Stokes with Spalart-Allmaras, Roe + MUSCL3 on primitive variables,
    LU-SGS with local dt at CFL 5, inlet total conditions, outlet at
    p_ratio * Pt, multigrid from the uniform field (``levels`` grid levels, 0:
    the reference's two-grid cycle)."""
    nr, ns = len(rotor_nk), len(stator_nk)
    zr = tuple(np.linspace(-0.2, 0.0, nr + 1))
    zs = tuple(np.linspace(0.0, 0.26, ns + 1))
    # axial velocity of the isentropic flow at the outlet pressure ratio
    g = 1.4
    mach = np.sqrt(2.0 / (g - 1.0) * (p_ratio ** (-(g - 1.0) / g) - 1.0))
    temp = 288.15 / (1.0 + 0.5 * (g - 1.0) * mach * mach)
    vx = mach * np.sqrt(g * 287.0 * temp)
    omega = rpm * 2.0 * np.pi / 60.0
    kw = dict(ni=ni, seed=seed, viscous=True, len_ref=1.0, max_iter=max_iter, profile="turn", p_ratio=p_ratio)
    R = annular_passage(nj=nj_rotor, nk_blocks=rotor_nk, nblade=nblade_rotor, rpm=rpm, z_extent=zr,
                        blade_blocks=(nr // 2,), camber=camber_rotor, helix=-omega / vx, **kw)
    S = annular_passage(nj=nj_stator, nk_blocks=stator_nk, nblade=nblade_stator, rpm=0.0, z_extent=zs,
                        blade_blocks=(ns // 2,), camber=camber_stator, **kw)
    nodes = R.nodes + S.nodes
    blocks = R.blocks + [BlockInfo(bi.bid + nr, bi.im, bi.jm, bi.km, bi.nblade, bi.rpm) for bi in S.blocks]
    states = R.states + S.states
    patches = [p for p in R.patches if p.kind != "outlet"]
    for p in S.patches:
        if p.kind != "inlet":
            patches.append(dataclasses.replace(p, block=p.block + nr, block2=p.block2 + nr if p.block2 else 0))
    km1 = R.nodes[nr - 1].shape[2]
    patches.append(Patch("mixing", nr, 6, (1, ni, 1, nj_rotor, km1, km1), orient=1, block2=nr + 1, face2=5,
                         rng2=(1, ni, 1, nj_stator, 1, 1)))
    kind_order = {"inlet": 0, "outlet": 1, "wall": 2, "periodic": 3, "interior": 4, "mixing": 5}
    patches.sort(key=lambda p: kind_order[p.kind])
    c = Case(R.cfg, nodes, patches, blocks, states)
    if turbulence == 200:
        c = with_sa(c, amp=0.0)
    else:
        c.cfg.turbulence = turbulence
    c.cfg.use_multigrid, c.cfg.reuse_solution, c.cfg.initial = "Y", "N", "U"
    c.cfg.coarse_iter, c.cfg.mg_levels = coarse_iter, levels
    return c


def from_meta(meta: Dict) -> Case:
    """The case a fixture's metadata describes: ``kwargs`` for bump_channel (or
    for annular_passage when ``generator`` names it, with ``turn`` a list of
    (block, quarter) re-indexings) and the iteration count, optionally
    ``wall_shear`` (with_wall_shear), ``turbulence`` (the config's
    turbulence model code) and ``cfg`` (other config fields by name, e.g. the
    two-grid settings)."""
    if meta.get("generator") == "annular_passage":
        case = annular_passage(max_iter=meta["iterations"], **meta["kwargs"])
        for blk, quarter in meta.get("turn", []):
            turn_block(case, blk, quarter)
    else:
        case = bump_channel(max_iter=meta["iterations"], **meta["kwargs"])
    if meta.get("wall_shear"):
        case = with_wall_shear(case)
    if meta.get("turbulence") == 200:
        case = with_sa(case)
    elif meta.get("turbulence"):
        case.cfg.turbulence = meta["turbulence"]
    for k, v in meta.get("cfg", {}).items():
        if not hasattr(case.cfg, k):
            raise KeyError(k)
        setattr(case.cfg, k, tuple(v) if isinstance(v, list) else v)
    return case


def annular_passage(ni: int = 8, nj: int = 8, nk_blocks: Sequence[int] = (16,), seed: int = 1,
                    hub: float = 0.2, tip: float = 0.25, nblade: int = 22, rpm: float = 0.0,
                    z_extent: Sequence[float] = (), camber: float = 0.0, blade_blocks: Sequence[int] = (),
                    beta: float = 1.5, cfl: float = 5.0, max_iter: int = 500, flux: int = 300,
                    recon_scheme: int = 500, recon_var: int = 200, p_ratio: float = 0.95,
                    viscous: bool = False, len_ref: float = 1.0, rotate_walls: bool = True,
                    periodic_from_high: bool = False, stagger: float = 0.0, profile: str = "arc",
                    helix: float = 0.0) -> Case:
    """Annular-sector blade passage: i radial (hub I- to tip I+, tanh-clustered),
    j pitchwise over one blade pitch 2*pi/nblade, k axial.  The passage is cut
    along k into blocks of ``nk_blocks`` cells joined by interior patches
    (K+ of block b to K- of block b+1, orientation 2).  In the blocks listed in
    ``blade_blocks`` (0-based) the J faces are the two sides of a zero-thickness
    circular-arc blade of pitchwise ``camber`` (as a fraction of the pitch) and
    are walls; elsewhere J- and J+ of a block are periodic (orientation 2).
    Hub and tip are walls (rotating with the block when ``rotate_walls``).
    Inlet K- of the first block, outlet K+ of the last.  All blocks share one
    rotation speed ``rpm`` and blade count.  ``stagger`` (degrees) shears the
    whole passage helically: at radius r the pitchwise position advances by
    tan(stagger) * (z - z_mid) / r, so the blade and the periodic surfaces meet
    a relative inflow at that angle.  ``periodic_from_high`` lists the
    periodic pair from J+ to J- (its second side then lies at the smaller
    angle, which makes the reference flip the block's pitch angle every time
    it applies the condition, source/BndCondInteriorPeriodic.cpp:54-59).
    ``profile`` "arc" makes the blade a circular arc through both ends (equal
    and opposite edge angles); "turn" makes it a turning vane, the parabola
    theta = camber * pitch * t^2, which meets an axial inflow at zero
    incidence and turns it; the blocks downstream of it are shifted by the
    vane's pitchwise offset so the interfaces stay conforming.  ``helix``
    (radians per metre) turns every surface of constant j by helix * z about
    the axis, the same angle at every radius: a rotor row whose blade meets the
    relative flow of an axial inflow V_x at rotation omega at zero incidence
    on every radius has helix = -omega / V_x."""
    nb = len(nk_blocks)
    if not z_extent:
        z_extent = [float(b) for b in range(nb + 1)]
        z_extent = [0.1 * (z - nb / 2.0) for z in z_extent]
    pitch = 2.0 * np.pi / nblade
    r = hub + (tip - hub) * _tanh_cluster(ni, beta)
    cfg = Config(grid_file="grid.dat", turbo_type="G", equation="N" if viscous else "E", len_ref=len_ref,
                 cfl=cfl, max_iter=max_iter, flux_scheme=flux, recon_scheme=recon_scheme, recon_var=recon_var,
                 p_out=p_ratio * 101325.0, n_threads=1)
    nodes, patches, blocks, states = [], [], [], []
    for b, nk in enumerate(nk_blocks):
        im, jm, km = ni + 1, nj + 1, nk + 1
        z = np.linspace(z_extent[b], z_extent[b + 1], km)
        t = (z - z_extent[b]) / (z_extent[b + 1] - z_extent[b])
        if b in blade_blocks and camber != 0.0 and profile == "turn":
            th0 = pitch * camber * t * t
        elif b in blade_blocks and camber != 0.0:
            # circular arc through (0, 0) and (1, 0) with mid-height camber
            h = camber
            rad = (0.25 + h * h) / (2.0 * h)
            th0 = pitch * ((h - rad) + np.sqrt(np.maximum(rad * rad - (t - 0.5) ** 2, 0.0)))
        elif profile == "turn" and any(q < b for q in blade_blocks):
            th0 = np.full(km, pitch * camber)
        else:
            th0 = np.zeros(km)
        zmid = 0.5 * (z_extent[0] + z_extent[-1])
        shear = np.tan(np.radians(stagger)) * (z[None, None, :] - zmid) / r[:, None, None]
        th = th0[None, None, :] + pitch * np.linspace(0.0, 1.0, jm)[None, :, None] + shear + helix * z[None, None, :]
        xyz = np.empty((im, jm, km, 3))
        rr = r[:, None, None]
        xyz[..., 0] = rr * np.cos(th)
        xyz[..., 1] = rr * np.sin(th)
        xyz[..., 2] = z[None, None, :]
        nodes.append(xyz)
        bid = b + 1
        blocks.append(BlockInfo(bid, im, jm, km, nblade, rpm))
        rw = 1 if (rotate_walls and rpm != 0.0) else 0
        patches.append(Patch("wall", bid, 1, (1, 1, 1, nj, 1, nk), rotate=rw))
        patches.append(Patch("wall", bid, 2, (im, im, 1, nj, 1, nk), rotate=0))
        if b in blade_blocks:
            patches.append(Patch("wall", bid, 3, (1, ni, 1, 1, 1, nk), rotate=rw))
            patches.append(Patch("wall", bid, 4, (1, ni, jm, jm, 1, nk), rotate=rw))
        else:
            lo, hi_ = (3, (1, ni, 1, 1, 1, nk)), (4, (1, ni, jm, jm, 1, nk))
            a, c = (hi_, lo) if periodic_from_high else (lo, hi_)
            patches.append(Patch("periodic", bid, a[0], a[1], orient=2, block2=bid, face2=c[0], rng2=c[1]))
        if b == 0:
            patches.append(Patch("inlet", bid, 5, (1, ni, 1, nj, 1, 1)))
        else:
            patches.append(Patch("interior", bid - 1, 6, (1, ni, 1, nj, nodes[b - 1].shape[2], nodes[b - 1].shape[2]),
                                 orient=2, block2=bid, face2=5, rng2=(1, ni, 1, nj, 1, 1)))
        if b == nb - 1:
            patches.append(Patch("outlet", bid, 6, (1, ni, 1, nj, km, km)))
        states.append(uniform_state(cfg, (ni, nj, nk), seed=seed + b))
    # table order: inlets, outlets, walls, periodic, interior
    order = {"inlet": 0, "outlet": 1, "wall": 2, "periodic": 3, "interior": 4}
    patches.sort(key=lambda p: order[p.kind])
    return Case(cfg, nodes, patches, blocks, states)


def with_mixing_plane(case: Case, block1: int, block2: int) -> Case:
    """The interior patch between block1's K-high face and block2's K-low face
    becomes a mixing plane (source/BndCondMixingPlane.cpp:287-522): each side
    then sees the pitch-wise flux average of the other."""
    for q, p in enumerate(case.patches):
        if p.kind != "interior":
            continue
        if (p.block, p.face, p.block2, p.face2) == (block2, 5, block1, 6):
            p = dataclasses.replace(p, block=block1, face=6, rng=p.rng2, block2=block2, face2=5, rng2=p.rng)
        if (p.block, p.face, p.block2, p.face2) == (block1, 6, block2, 5):
            case.patches[q] = dataclasses.replace(p, kind="mixing", orient=1)
            return case
    raise ValueError(f"no interior patch from block {block1} (K high) to block {block2} (K low)")


def turn_block(case: Case, block: int, quarter: int) -> Case:
    """Re-indexes block ``block`` (1-based) by a rotation of its (i, j) index
    plane by ``quarter`` x 90 degrees, keeping the grid, the state and every
    patch on it the same physical objects; handedness is preserved.  With
    quarter 1 the new (i, j) = (j, NI-1-i), with 2 (NI-1-i, NJ-1-j), with 3
    (NJ-1-j, i) (0-based cell indices).  Patch orientations on interfaces with
    other blocks are recomputed for the reference's CMesh::Trans convention
    (source/ShadowOutIn.cpp:166-325)."""
    b = block - 1
    q = quarter % 4
    if q == 0:
        return case
    xyz = case.nodes[b]
    im, jm = xyz.shape[:2]
    ni, nj = im - 1, jm - 1

    def rot_arr(a):
        # a[i, j, ...] -> new array indexed by the new (i, j)
        if q == 1:
            return np.ascontiguousarray(np.transpose(a, (1, 0) + tuple(range(2, a.ndim)))[:, ::-1])
        if q == 2:
            return np.ascontiguousarray(a[::-1, ::-1])
        return np.ascontiguousarray(np.transpose(a, (1, 0) + tuple(range(2, a.ndim)))[::-1, :])

    case.nodes[b] = rot_arr(xyz)
    case.states[b] = {k: rot_arr(v) for k, v in case.states[b].items()}
    bi = case.blocks[b]
    if q in (1, 3):
        bi.im, bi.jm = bi.jm, bi.im
    # old cell (i, j) (1-based) -> new; node index n (1..N) likewise as a face index
    def cell(i, j):
        if q == 1:
            return j, ni + 1 - i
        if q == 2:
            return ni + 1 - i, nj + 1 - j
        return nj + 1 - j, i

    def face_map(face, rng):
        d, hi = (face - 1) // 2, (face - 1) % 2
        i0, i1, j0, j1, k0, k1 = rng
        if d == 2:
            a, c = cell(i0, j0), cell(i1, j1)
            ri, rj = sorted((a[0], c[0])), sorted((a[1], c[1]))
            return face, (ri[0], ri[1], rj[0], rj[1], k0, k1)
        # the cells next to an I or J face, mapped; their new normal index is
        # 1 (low side) or the new cell count (high side)
        if d == 0:
            ci = 1 if hi == 0 else ni
            a, c = cell(ci, j0), cell(ci, j1)
        else:
            cj = 1 if hi == 0 else nj
            a, c = cell(i0, cj), cell(i1, cj)
        nim, njm = (nj, ni) if q in (1, 3) else (ni, nj)
        nd = 1 - d if q in (1, 3) else d
        if nd == 0:
            newhi = 0 if a[0] == 1 else 1
            rj = sorted((a[1], c[1]))
            return 1 + newhi, (1 if newhi == 0 else nim + 1,) * 2 + (rj[0], rj[1], k0, k1)
        newhi = 0 if a[1] == 1 else 1
        ri = sorted((a[0], c[0]))
        return 3 + newhi, (ri[0], ri[1]) + (1 if newhi == 0 else njm + 1,) * 2 + (k0, k1)

    for p in case.patches:
        if p.block == block:
            p.face, p.rng = face_map(p.face, p.rng)
        if p.kind in ("periodic", "interior", "mixing") and p.block2 == block:
            p.face2, p.rng2 = face_map(p.face2, p.rng2)
    for p in case.patches:
        if p.kind in ("periodic", "interior") and block in (p.block, p.block2):
            p.orient = _orientation(case, p)
    return case


def _shadow_axes(face: int):
    """(first, second) tangential axes of the reference's shadow array of a face
    (source/ShadowOutIn.cpp:12-164)."""
    return {1: (2, 1), 2: (1, 2), 3: (0, 2), 4: (2, 0), 5: (1, 0), 6: (0, 1)}[face]


def _orientation(case: Case, p: Patch) -> int:
    """The CMesh::Trans code that maps side 1's shadow array onto side 2's for
    a conformal interface, found by matching face-centre coordinates."""
    def face_centres(block, face, rng):
        xyz = case.nodes[block - 1]
        d = (face - 1) // 2
        al, am = _shadow_axes(face)
        L = rng[2 * al + 1] - rng[2 * al] + 1
        M = rng[2 * am + 1] - rng[2 * am] + 1
        out = np.zeros((L, M, 3))
        for l in range(L):
            for m in range(M):
                c = [0, 0, 0]
                c[al] = rng[2 * al] + l
                c[am] = rng[2 * am] + m
                c[d] = rng[2 * d]
                # face centre: mean of its four nodes (node indices, 1-based)
                pts = []
                for s in (0, 1):
                    for t in (0, 1):
                        n = list(c)
                        n[al] += s
                        n[am] += t
                        pts.append(xyz[n[0] - 1, n[1] - 1, n[2] - 1])
                out[l, m] = np.mean(pts, axis=0)
        return out

    a = face_centres(p.block, p.face, p.rng)
    b = face_centres(p.block2, p.face2, p.rng2)
    if p.kind == "periodic":
        # compare in (r, z) and pitchwise order only
        def rz(x):
            return np.stack([np.hypot(x[..., 0], x[..., 1]), x[..., 2]], -1)
        a, b = rz(a), rz(b)
    L1, M1 = a.shape[:2]
    cands = {1: lambda A: A[:, ::-1], 2: lambda A: np.transpose(A, (1, 0, 2)), 3: lambda A: A[::-1, :],
             4: lambda A: np.transpose(A, (1, 0, 2))[::-1, ::-1]}
    for o, f in cands.items():
        t = f(a)
        if t.shape == b.shape and np.allclose(t, b, rtol=0, atol=1e-9):
            return o
    raise ValueError(f"no orientation maps {p}")
