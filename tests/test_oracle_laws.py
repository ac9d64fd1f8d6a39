"""The C restatement against what the mathematics fixes.

These checks hold whatever the reference does, so a restatement error that the
reference fixtures happen not to exercise (a dropped term, a wrong sign, a
transposed index) still fails one of them:

* consistency of both face fluxes, F(Q, Q) = the physical flux through a moving
  face, computed here from the primitive variables;
* upwinding: with every characteristic speed above the entropy-fix width the
  Roe flux is the flux of the upwind state (the Roe-matrix property);
* Steger-Warming splitting: F+(Q) + F-(Q) = F(Q);
* every reconstruction reproduces constant data, and linear data gives the face
  midpoint;
* minmod-based reconstructions fall to first order at an extremum;
* face vectors close every cell and Cartesian boxes get exact volumes;
* the residual telescopes: the sum over cells equals the outward boundary flux.
"""
import numpy as np
import pytest

from fv3d import cases
import oracle as O

GAMMA = 1.4


def physical_flux(q, que):
    """Flux through a face moving with normal speed que, face frame."""
    r, ru, rv, rw, re = q
    u, v, w = ru / r, rv / r, rw / r
    p = (GAMMA - 1.0) * (re - 0.5 * r * (u * u + v * v + w * w))
    un = u - que
    return np.array([r * un, ru * un + p, rv * un, rw * un, (re + p) * un + que * p])


def cons(rho, u, v, w, p):
    return np.array([rho, rho * u, rho * v, rho * w, p / (GAMMA - 1.0) + 0.5 * rho * (u * u + v * v + w * w)])


def random_states(n, seed, speed=1.0):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        rho = rng.uniform(0.5, 2.0)
        u, v, w = rng.uniform(-speed, speed, 3)
        p = rng.uniform(0.3, 1.5)
        out.append(cons(rho, u, v, w, p))
    return out


@pytest.mark.parametrize("scheme", [200, 300])
@pytest.mark.parametrize("que", [0.0, 0.37, -0.81])
def test_flux_consistency(scheme, que):
    for q in random_states(20, seed=11):
        f = O.point_flux(scheme, GAMMA, q, q, que)
        ref = physical_flux(q, que)
        np.testing.assert_allclose(f, ref, rtol=1e-12, atol=1e-12 * np.max(np.abs(ref)))


@pytest.mark.parametrize("direction", [1, -1])
def test_roe_supersonic_upwind(direction):
    # c = 1 (p = rho / gamma), normal speed 3 in the flow direction: every
    # |lambda| >= 2 > 0.5, so the entropy fix is inactive and the Roe flux is the
    # upwind flux.  The shear-wave strengths divide by (v + 1e-8), hence the 1e-7.
    rng = np.random.default_rng(5)
    for _ in range(20):
        rho_l, rho_r = rng.uniform(0.8, 1.2, 2)
        ql = cons(rho_l, 3.0 * direction + rng.uniform(-0.2, 0.2), 1.0 + rng.uniform(0, 0.3),
                  0.8 + rng.uniform(0, 0.3), rho_l / GAMMA)
        qr = cons(rho_r, 3.0 * direction + rng.uniform(-0.2, 0.2), 1.0 + rng.uniform(0, 0.3),
                  0.8 + rng.uniform(0, 0.3), rho_r / GAMMA)
        f = O.point_flux(300, GAMMA, ql, qr, 0.0)
        up = physical_flux(ql if direction > 0 else qr, 0.0)
        np.testing.assert_allclose(f, up, rtol=1e-7, atol=1e-7 * np.max(np.abs(up)))


def test_roe_differs_from_central_when_subsonic():
    # guards against a dissipation term that vanishes identically
    ql, qr = cons(1.0, 0.3, 0.2, 0.1, 0.7), cons(1.3, 0.1, 0.25, 0.05, 0.9)
    f = O.point_flux(300, GAMMA, ql, qr, 0.0)
    central = 0.5 * (physical_flux(ql, 0.0) + physical_flux(qr, 0.0))
    assert np.max(np.abs(f - central)) > 1e-3


def test_steger_warming_upwind_when_supersonic():
    # all eigenvalues of the left state positive, all of the right state's
    # split part negative-free: F = F+(QL) + F-(QR) = F(QL) + O(eps^2)
    ql = cons(1.0, 4.0, 0.3, -0.2, 1.0 / GAMMA)
    qr = cons(1.2, 4.5, 0.1, 0.2, 1.1 / GAMMA)
    f = O.point_flux(200, GAMMA, ql, qr, 0.0)
    np.testing.assert_allclose(f, physical_flux(ql, 0.0), rtol=1e-10)


def prim(rho, u, v, w, p):
    return np.array([rho, u, v, w, p, p / rho])


@pytest.mark.parametrize("var", [100, 200])
@pytest.mark.parametrize("scheme", [100, 200, 300, 400, 500])
def test_reconstruction_constant_data(var, scheme):
    w = prim(1.1, 0.3, -0.2, 0.4, 0.8)
    x = w if var == 200 else cons(*w[:5])
    l, r = O.point_recon(var, scheme, GAMMA, x, x, x, x)
    q = cons(*w[:5])
    np.testing.assert_allclose(l, q, rtol=1e-15)
    np.testing.assert_allclose(r, q, rtol=1e-15)


@pytest.mark.parametrize("var", [100, 200])
@pytest.mark.parametrize("scheme", [100, 200, 300, 400, 500])
def test_reconstruction_linear_data(var, scheme):
    # every record linear in the stencil position: both sides give the midpoint
    base = prim(1.0, 0.2, -0.1, 0.3, 0.9) if var == 200 else cons(1.0, 0.2, -0.1, 0.3, 0.9)
    slope = np.array([0.05, 0.02, 0.03, -0.01, 0.04, 0.01])[: base.size]
    if var == 100:
        slope = np.array([0.05, 0.06, -0.02, 0.04, 0.1])
    st = [base + slope * t for t in (-1.0, 0.0, 1.0, 2.0)]
    l, r = O.point_recon(var, scheme, GAMMA, *st)
    mid = 0.5 * (st[1] + st[2])
    want = cons(*mid[:5]) if var == 200 else mid
    np.testing.assert_allclose(l, want, rtol=2e-12)
    np.testing.assert_allclose(r, want, rtol=2e-12)


@pytest.mark.parametrize("var", [100, 200])
@pytest.mark.parametrize("scheme", [100, 200, 300])
def test_minmod_first_order_at_extremum(var, scheme):
    # F is a local maximum in every component: minmod slopes vanish, L = F
    F = prim(1.2, 0.5, 0.4, 0.3, 1.0) if var == 200 else cons(1.2, 0.5, 0.4, 0.3, 1.0)
    lo = F * 0.9
    l, r = O.point_recon(var, scheme, GAMMA, lo, F, lo, lo * 0.95)
    want_l = cons(*F[:5]) if var == 200 else F
    np.testing.assert_allclose(l, want_l, rtol=1e-15)


def _face_sums(o, case):
    SI, SJ, SK = o.get(1, "SI"), o.get(1, "SJ"), o.get(1, "SK")
    return SI[1:] - SI[:-1] + SJ[:, 1:] - SJ[:, :-1] + SK[:, :, 1:] - SK[:, :, :-1], max(
        np.abs(SI).max(), np.abs(SJ).max(), np.abs(SK).max())


def test_face_vectors_close_every_cell():
    case = cases.bump_channel(ni=6, nj=8, nk=14, seed=1)
    o = O.Oracle(case)
    net, scale = _face_sums(o, case)
    assert np.abs(net).max() <= 1e-14 * scale


def test_box_grid_volumes_and_areas():
    rng = np.random.default_rng(3)
    case = cases.bump_channel(ni=5, nj=4, nk=7, seed=1, bump_height=0.1)
    xs = np.cumsum(np.r_[0.0, rng.uniform(0.5, 1.5, 5)])
    ys = np.cumsum(np.r_[0.0, rng.uniform(0.5, 1.5, 4)])
    zs = np.cumsum(np.r_[0.0, rng.uniform(0.5, 1.5, 7)])
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    case.nodes[0] = np.stack([X, Y, Z], -1)
    o = O.Oracle(case)
    vol = o.get(1, "vol")
    want = np.diff(xs)[:, None, None] * np.diff(ys)[None, :, None] * np.diff(zs)[None, None, :]
    np.testing.assert_allclose(vol, want, rtol=1e-13)
    SI = o.get(1, "SI")
    np.testing.assert_allclose(np.abs(SI[..., 0]), np.broadcast_to(
        np.diff(ys)[None, :, None] * np.diff(zs)[None, None, :], SI.shape[:3]), rtol=1e-13)
    assert np.abs(SI[..., 1:]).max() <= 1e-14


@pytest.mark.parametrize("flux,recon", [(300, 500), (200, 100)])
def test_residual_telescopes(flux, recon):
    case = cases.bump_channel(ni=6, nj=5, nk=12, seed=2, flux=flux, recon_scheme=recon)
    o = O.Oracle(case)
    o.call("resflux")
    res = o.get(1, "res")
    fI, fJ, fK = o.get(1, "fcI"), o.get(1, "fcJ"), o.get(1, "fcK")
    boundary = (fI[-1].sum((0, 1)) - fI[0].sum((0, 1)) + fJ[:, -1].sum((0, 1)) - fJ[:, 0].sum((0, 1)) +
                fK[:, :, -1].sum((0, 1)) - fK[:, :, 0].sum((0, 1)))
    scale = np.abs(fI).sum((0, 1, 2)) + np.abs(fJ).sum((0, 1, 2)) + np.abs(fK).sum((0, 1, 2))
    assert np.all(np.abs(res.sum((0, 1, 2)) - boundary) <= 1e-13 * scale)


@pytest.mark.parametrize("var", [100, 200])
@pytest.mark.parametrize("scheme", [100, 200, 300, 500])
def test_reconstruction_mirror_symmetry(var, scheme):
    # Reversing the stencil (BB, B | F, FF) must swap the two sides exactly: the
    # right state is the left state of the mirrored stencil.  (MUSCL2o adds eps to
    # signed differences and is only approximately symmetric, so it is left out.)
    rng = np.random.default_rng(17 + scheme + var)
    for _ in range(20):
        st = []
        for _ in range(4):
            w = prim(rng.uniform(0.8, 1.2), *rng.uniform(-0.5, 0.5, 3), rng.uniform(0.7, 1.1))
            st.append(w if var == 200 else cons(*w[:5]))
        l, r = O.point_recon(var, scheme, GAMMA, *st)
        lm, rm = O.point_recon(var, scheme, GAMMA, *st[::-1])
        assert np.array_equal(r, lm) and np.array_equal(l, rm)


# Characteristic reconstruction (source/Reconstruction.cpp:363-468): S takes
# conservative states to characteristic variables of the face state and S1
# takes them back.  S1 S is the identity up to the reference's float
# quotients (-1/3, 2/3, 1/6 in S1), so a mistyped entry of either matrix
# breaks constant and linear data; the limiter is the conservative one.
@pytest.mark.parametrize("scheme", [100, 200, 300, 400, 500])
def test_characteristic_reconstruction_constant_data(scheme):
    w = prim(1.1, 0.3, -0.2, 0.4, 0.8)
    l, r = O.point_recon(300, scheme, GAMMA, w, w, w, w)
    q = cons(*w[:5])
    np.testing.assert_allclose(l, q, rtol=1e-7, atol=1e-8)
    np.testing.assert_allclose(r, q, rtol=1e-7, atol=1e-8)


@pytest.mark.parametrize("scheme", [100, 200, 300, 500])
def test_characteristic_reconstruction_linear_data(scheme):
    # conservative states linear in the stencil position stay linear in the
    # characteristic variables of one frozen face state: the midpoint again
    base = cons(1.0, 0.2, -0.1, 0.3, 0.9)
    slope = np.array([0.05, 0.06, -0.02, 0.04, 0.1])
    qs = [base + slope * t for t in (-1.0, 0.0, 1.0, 2.0)]
    st = []
    for q in qs:
        rho, u, v, w = q[0], q[1] / q[0], q[2] / q[0], q[3] / q[0]
        p = (GAMMA - 1.0) * (q[4] - 0.5 * rho * (u * u + v * v + w * w))
        st.append(prim(rho, u, v, w, p))
    l, r = O.point_recon(300, scheme, GAMMA, *st)
    mid = 0.5 * (qs[1] + qs[2])
    np.testing.assert_allclose(l, mid, rtol=1e-7)
    np.testing.assert_allclose(r, mid, rtol=1e-7)


@pytest.mark.parametrize("scheme", [100, 200, 300, 500])
def test_characteristic_reconstruction_mirror_symmetry(scheme):
    # the face state is the mean of F and B, the same for the mirrored stencil
    rng = np.random.default_rng(29 + scheme)
    for _ in range(20):
        st = [prim(rng.uniform(0.8, 1.2), *rng.uniform(-0.5, 0.5, 3), rng.uniform(0.7, 1.1)) for _ in range(4)]
        l, r = O.point_recon(300, scheme, GAMMA, *st)
        lm, rm = O.point_recon(300, scheme, GAMMA, *st[::-1])
        assert np.array_equal(r, lm) and np.array_equal(l, rm)


def test_characteristic_differs_from_conservative():
    # an entropy wave (density jump at constant p and u) next to a velocity
    # kink: limiting characteristic fields separately gives other states
    st = [prim(1.0, 0.3, 0.0, 0.0, 0.8), prim(1.0, 0.35, 0.0, 0.0, 0.8), prim(1.4, 0.2, 0.0, 0.0, 0.8),
          prim(1.45, 0.3, 0.0, 0.0, 0.8)]
    l3, _ = O.point_recon(300, 500, GAMMA, *st)
    l1, _ = O.point_recon(100, 500, GAMMA, *[cons(*w[:5]) for w in st])
    assert np.max(np.abs(l3 - l1)) > 1e-4
