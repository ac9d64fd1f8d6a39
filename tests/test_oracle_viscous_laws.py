"""The laminar Navier-Stokes part of the C restatement against closed forms.

On a uniform Cartesian grid (i along x, j along y, k along z) with fields that
are linear in the cell centroids (ghost cells included), the index-space
differences of the gradient stencils are known exactly, so every face gradient,
viscous flux and vorticity has a closed form.  Two properties of the
reference's coordinate coefficients enter those forms and are pinned here:

* gradz_kesi repeats the gradx_kesi expression (source/GeometryVariables.cpp:
  200, 259, 318, 376), so d/dz picks up + d/dx on every face and cell;
* the I-face zita vector averages Node[i][j][k] with itself (line 247), so on
  an I face e_zita = (0, dy/2, dz) and d/dz picks up - (dy / 2dz) d/dy.

The three-cell corner averages and the J-face corner j == JM, k == 1 (whose
three-cell sum the reference also scales by 0.25, source/Spectral_curl_grad.cpp:
331) are checked through constant and linear fields as well.
"""
import numpy as np
import pytest

from fv3d import cases
import oracle as O

NI, NJ, NK = 6, 5, 7
DX, DY, DZ = 0.1, 0.07, 0.13
X0 = np.array([0.3, -0.2, 0.5])


def cartesian_case(len_ref=1e-4):
    case = cases.bump_channel(ni=NI, nj=NJ, nk=NK, seed=1, viscous=True, len_ref=len_ref)
    im, jm, km = NI + 1, NJ + 1, NK + 1
    i, j, k = np.meshgrid(np.arange(im), np.arange(jm), np.arange(km), indexing="ij")
    case.nodes[0] = np.stack([X0[0] + DX * i, X0[1] + DY * j, X0[2] + DZ * k], -1).astype(np.float64)
    return case


def centroids():
    """Cell centroids over the whole (IM+1, JM+1, KM+1) box, ghost cells by
    linear extension (cell i spans nodes i..i+1, 1-based)."""
    im, jm, km = NI + 1, NJ + 1, NK + 1
    i, j, k = np.meshgrid(np.arange(im + 1), np.arange(jm + 1), np.arange(km + 1), indexing="ij")
    return np.stack([X0[0] + DX * (i - 0.5), X0[1] + DY * (j - 0.5), X0[2] + DZ * (k - 0.5)], -1)


def box(u=None, v=None, w=None, T=None):
    c = centroids()
    pv = np.zeros(c.shape[:3] + (6,))
    pv[..., 0] = 1.3
    pv[..., 4] = 0.7
    for n, f in ((1, u), (2, v), (3, w), (5, T)):
        pv[..., n] = f(c) if f is not None else (1.0 if n == 5 else 0.0)
    return pv


def run(pv, **kw):
    o = O.Oracle(cartesian_case(**kw))
    o.viscous_on(1, pv)
    return o


def interior_faces(d):
    """Face index ranges (0-based into the oracle's face arrays) whose
    tangential stencils use no corner-modified averages."""
    n = [NI, NJ, NK]
    sl = []
    for a in range(3):
        if a == d:
            sl.append(slice(1, n[a]))          # faces 2..N-1 (1-based), not the block ends
        else:
            sl.append(slice(1, n[a] - 1))      # tangential cells 2..N-2
    return tuple(sl)


@pytest.mark.parametrize("q", range(4))
def test_linear_field_gradients(q):
    cvec = np.array([0.8, -1.7, 0.45])
    f = lambda c: 2.0 + c @ cvec
    kw = {"u": f} if q == 0 else {"v": f} if q == 1 else {"w": f} if q == 2 else {"T": f}
    o = run(box(**kw))
    cx, cy, cz = cvec
    want = {0: (cx, cy, cz + cx - cy * DY / (2 * DZ)), 1: (cx, cy, cz + cx), 2: (cx, cy, cz + cx)}
    for d, name in enumerate(("gradI", "gradJ", "gradK")):
        g = o.get(1, name)[interior_faces(d)]
        for comp in range(3):
            np.testing.assert_allclose(g[..., comp * 4 + q], want[d][comp], rtol=1e-12, atol=1e-12)
        # the other three variables are constant: zero gradient
        others = [comp * 4 + r for comp in range(3) for r in range(4) if r != q]
        assert np.max(np.abs(g[..., others])) == 0.0


def test_constant_field_zero_gradient_except_quirk_corner():
    o = run(box(u=lambda c: np.full(c.shape[:3], 1.0)))
    for d, name in enumerate(("gradI", "gradJ", "gradK")):
        g = o.get(1, name)
        if d == 1:
            # J face j == JM (index NJ), k == 1 (index 0): M = 0.25 * (3-cell sum) / 3
            q = g[:, NJ, 0, :]
            # d_zita = 1 - 0.25 = 0.75 and only gradz_zita = 1/dz is non-zero
            np.testing.assert_allclose(q[:, 8], 0.75 / DZ, rtol=1e-13)
            assert np.max(np.abs(q[:, [0, 4]])) == 0.0
            g = g.copy()
            g[:, NJ, 0, :] = 0.0
        assert np.max(np.abs(g[..., [0, 4, 8]])) < 1e-13, name


def test_corner_three_cell_averages():
    # u = y: at the I face i=1, j=1 the lower yita average leaves out the corner
    # ghost (0, 0) and averages (i,j), (i-1,j), (i,j-1): mean y offset -1/3 of a
    # cell instead of -1/2, so d_yita = (1/2 + 1/3) dy; the i offsets of the
    # two averages (-1/2 and -1/3) do not matter for u = y.
    o = run(box(u=lambda c: c[..., 1]))
    g = o.get(1, "gradI")
    corner = g[0, 0, 1:-1]                  # i = 1, j = 1, interior k
    np.testing.assert_allclose(corner[:, 4], 5.0 / 6.0, rtol=1e-12)                    # d/dy
    np.testing.assert_allclose(corner[:, 8], -5.0 / 6.0 * DY / (2 * DZ), rtol=1e-12)   # d/dz
    # i = 1, j = JM-1: the upper average leaves out (0, JM)
    np.testing.assert_allclose(g[0, NJ - 1, 1:-1, 4], 5.0 / 6.0, rtol=1e-12)
    # u = x at the J face j = 1, i = 1: lower kesi average leaves out (0, 0);
    # its cells (i,j), (i-1,j), (i,j-1) have mean x offset -1/3
    o = run(box(u=lambda c: c[..., 0]))
    gj = o.get(1, "gradJ")
    np.testing.assert_allclose(gj[0, 0, 1:-1, 0], 5.0 / 6.0, rtol=1e-12)


@pytest.mark.parametrize("var", range(3))
@pytest.mark.parametrize("axis", range(3))
def test_uniform_strain_dissipation(var, axis):
    """Velocity component `var` = gam * coordinate `axis`, T uniform: the stress
    is uniform, so the momentum fluxes telescope to zero, and the energy flux
    difference is the dissipation mu gam^2 Vol for a shear and 4/3 mu gam^2 Vol
    for an extension (the quirks add uniform stresses that do no net work)."""
    gam = 0.9
    f = lambda c: gam * c[..., axis]
    o = run(box(**{"uvw"[var]: f}))
    mu = o.sutherland(1.0)
    fv = o.get(1, "fv")[1:-1, 1:-1, 1:-1]
    vol = DX * DY * DZ
    # the reference writes 4/3 as the float literal 4.0f/3.0f
    factor = float(np.float32(4.0) / np.float32(3.0)) if var == axis else 1.0
    np.testing.assert_allclose(fv[..., 4], factor * mu * gam * gam * vol, rtol=1e-10)
    scale = mu * gam * DX * DZ
    assert np.max(np.abs(fv[..., 0])) == 0.0
    assert np.max(np.abs(fv[..., 1:4])) < 1e-10 * scale


def test_rigid_rotation_vorticity():
    om = 1.7
    # u = -om*y, v = om*x: curl = 2 om e_z; with gradz_kesi = gradx_kesi the
    # cell picks up v_z = v_kesi/dx = om, so |Omega| = sqrt(0.5 (om^2 + 4 om^2))
    o = run(box(u=lambda c: -om * c[..., 1], v=lambda c: om * c[..., 0]))
    np.testing.assert_allclose(o.get(1, "omega"), om * np.sqrt(2.5), rtol=1e-12)
    # w = om*x only: w_x = om, nothing else (u_z = u_kesi/dx = 0)
    o = run(box(w=lambda c: om * c[..., 0]))
    np.testing.assert_allclose(o.get(1, "omega"), om / np.sqrt(2.0), rtol=1e-12)


def test_sutherland_reference_point():
    o = O.Oracle(cartesian_case(len_ref=1e-4))
    # at the reference temperature the non-dimensional viscosity is 1/Re_ref;
    # Re_ref = rho_ref V_ref L / mu_ref with the inlet total state (288.15 K,
    # 101325 Pa, R = 287.06, gamma = 1.4) and mu_ref = 1.716e-5
    rho, V = 101325.0 / 287.06 / 288.15, np.sqrt(1.4 * 287.06 * 288.15)
    re = rho * V * 1e-4 / 1.716e-5
    np.testing.assert_allclose(o.sutherland(1.0), 1.0 / re, rtol=1e-6)   # float-parsed constants
    # power law: mu(4)/mu(1) = 8 (1 + s)/(4 + s), s = 110.4/288.15
    s = 110.4 / 288.15
    np.testing.assert_allclose(o.sutherland(4.0) / o.sutherland(1.0), 8 * (1 + s) / (4 + s), rtol=1e-6)


def test_viscous_spectral_radius_cartesian():
    case = cartesian_case()
    st = case.states[0]
    rho, T = 1.3, 0.9
    for key in st:
        st[key] = np.full_like(st[key], 0.0)
    st["rho"][:] = rho
    st["T"][:] = T
    st["p"][:] = rho * T / 1.4
    st["w"][:] = 0.2
    st["rhoE"][:] = st["p"] / 0.4 + 0.5 * rho * 0.04
    o = O.Oracle(case)
    o.call("spectral")
    o.call("timestep")
    mu = o.sutherland(T)
    pr = np.float32(0.72)
    coef = max(4.0 / 3.0, 1.4) / rho * mu / pr
    areas = {"lamvI": DY * DZ, "lamvJ": DX * DZ, "lamvK": DX * DY}
    for name, a in areas.items():
        np.testing.assert_allclose(o.get(1, name), coef * a * a / (DX * DY * DZ), rtol=1e-6)


@pytest.mark.parametrize("axis", range(3))
def test_linear_temperature_heat_flux(axis):
    """T linear in one coordinate, fluid at rest: the conduction flux k dT/dn
    through each face is uniform per direction, so the cell sums vanish, and the
    face energy flux equals cond * (grad T . S) with cond from the face T."""
    beta = 0.35
    o = run(box(T=lambda c: 1.0 + beta * c[..., axis]))
    fv = o.get(1, "fv")[1:-1, 1:-1, 1:-1]
    assert np.max(np.abs(fv[..., :4])) == 0.0
    # face of the gradient's direction: S is along that axis with area A
    name = "fv" + "IJK"[axis]
    area = {0: DY * DZ, 1: DX * DZ, 2: DX * DY}[axis]
    ff = o.get(1, name)
    # face temperatures along the axis vary; cond = mu(T) gamma R/(gamma-1)/Pr
    c = centroids()
    sl = [slice(1, -1)] * 3
    t_face = 1.0 + beta * (c[..., axis] - 0.5 * (DX, DY, DZ)[axis])
    g = float(np.float32(1.4))
    rgas = 1.0 / g
    pr = float(np.float32(0.72))
    mu = np.vectorize(o.sutherland)
    # interior faces away from the corner-stencil rows (gradients exact there)
    idx = [slice(1, n - 1) for n in (NI, NJ, NK)]
    idx[axis] = slice(1, (NI, NJ, NK)[axis])
    want_t = t_face[1:NI + 2, 1:NJ + 2, 1:NK + 2][tuple(idx)]
    want = mu(want_t) * g * rgas / (g - 1.0) / pr * beta * area
    np.testing.assert_allclose(ff[tuple(idx)][..., 4], want, rtol=1e-10)
