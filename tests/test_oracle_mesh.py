"""Pins of the oracle's stretched meshes (oracle/mesh.c and the per-cell widths of fv.c,
multiphase.c and riemann.c) against what the mathematics fixes, and of the library's mesh generator
(m2c_mesh_1d, host code) against the oracle's bit for bit.

Mesh generator (MeshGenerator.cpp:141-259): a uniform axis has centres x0 + (i + 1/2) h; between two
control points the width is linear in x, h(x) = a x + b, the cell count is the integral of dx / h
rounded down and every cell covers the same share of that integral, so the widths form a geometric
progression of ratio exp(a / kappa) whose sum is the axis length.

Finite volumes on a stretched mesh: the generalised van Albada and minmod limiters with the
per-cell coefficients A = (w[i-1] + w[i]) / (w[i] + w[i+1]), B = 2 w[i] / (w[i] + w[i+1])
(Reconstructor.cpp:53-121, Reconstructor.h:95-110) reproduce a linear field exactly, so the face
states of a linear density meet and the mass flux of the local Lax-Friedrichs flux vanishes at
interior faces; the face areas and cell volumes of SpaceOperator.cpp:160-177 keep the totals of mass,
momentum and energy between walls; the time step of a uniform state is cfl times the narrowest
width over the sound speed. Level set: third-order differences on the stretched coordinates are
exact on cubics, a plane signed distance is a reinitialisation fixed point, and the level-set
gradient at a face (SpaceOperator.cpp:3048-3110) interpolates the tangential derivatives to the face
position x[i] - w[i]/2."""
import math

import numpy as np
import pytest

import m2c_amd as M
from oracle import pyoracle as po


def geometric(h1, h2, L):
    a = (h2 - h1) / L
    nreal = math.log(h2 / h1) / a
    n = int(nreal + 1e-12)
    return a, nreal, n


def test_uniform_axis_closed_form():
    x, w = po.mesh_1d(-1.0, 2.0, 12)
    assert len(x) == 12
    assert np.all(w == 0.25)
    assert np.max(np.abs(x - (-1.0 + (np.arange(12) + 0.5) * 0.25))) <= 1e-15


def test_constant_width_control_point_is_uniform():
    x, w = po.mesh_1d(0.0, 2.0, control_points=[(0.7, 0.125)])
    assert len(x) == 16
    assert np.max(np.abs(w - 0.125)) <= 1e-15
    assert np.max(np.abs(x - (np.arange(16) + 0.5) * 0.125)) <= 1e-14


@pytest.mark.parametrize("h1,h2", [(0.01, 0.05), (0.08, 0.02), (0.02, 0.021)])
def test_linear_width_is_a_geometric_progression(h1, h2):
    L = 1.5
    x, w = po.mesh_1d(0.0, L, control_points=[(0.0, h1), (L, h2)])
    a, nreal, n = geometric(h1, h2, L)
    assert len(w) == n
    r = math.exp(a * nreal / n)
    assert np.max(np.abs(w[1:] / w[:-1] - r)) <= 1e-10
    # the first cell solves the integral of dx / h over it = Nreal / N
    assert abs(w[0] - h1 / a * (r - 1.0)) <= 1e-12 * h1
    assert abs(w.sum() - L) <= 1e-13
    assert abs(x[0] - 0.5 * w[0]) <= 1e-16
    assert np.max(np.abs(np.diff(x) - 0.5 * (w[1:] + w[:-1]))) <= 1e-14


def test_piecewise_control_points():
    pts = [(0.2, 0.02), (0.5, 0.01), (0.9, 0.04)]
    x, w = po.mesh_1d(0.0, 1.2, control_points=pts)
    nreal = 0.2 / 0.02 + geometric(0.02, 0.01, 0.3)[1] + geometric(0.01, 0.04, 0.4)[1] + 0.3 / 0.04
    assert len(w) == int(nreal + 1e-12)
    assert abs(w.sum() - 1.2) <= 1e-13
    # widths follow the control points' profile to within the rounding of the count
    h = np.interp(x, [0.0, 0.2, 0.5, 0.9, 1.2], [0.02, 0.02, 0.01, 0.04, 0.04])
    assert np.max(np.abs(w / h - 1.0)) < 0.05
    assert np.all(np.diff(w[(x > 0.21) & (x < 0.49)]) < 0) and np.all(np.diff(w[(x > 0.51) & (x < 0.89)]) > 0)


def test_control_points_outside_the_axis_are_errors():
    with pytest.raises(ValueError):
        po.mesh_1d(0.0, 1.0, control_points=[(-0.1, 0.05)])
    with pytest.raises(ValueError):
        po.mesh_1d(0.0, 1.0, control_points=[(0.5, 0.05), (1.1, 0.05)])


MESHES = [(0.0, 1.0, [(0.0, 0.01), (1.0, 0.05)]), (0.0, 1.2, [(0.2, 0.02), (0.5, 0.01), (0.9, 0.04)]),
          (-3.0, 4.0, [(0.0, 0.05), (-3.0, 0.3), (4.0, 0.2)]), (0.0, 2.0, [(0.7, 0.125)]),
          (1.0, 1.7, [(1.3, 0.001), (1.0, 0.01), (1.7, 0.01)])]


@pytest.mark.parametrize("x0,x1,pts", MESHES)
def test_library_generator_matches_oracle_bitwise(x0, x1, pts):
    xo, wo = po.mesh_1d(x0, x1, control_points=pts)
    xl, wl = M.mesh_1d(x0, x1, points=pts)
    assert np.array_equal(xo, xl) and np.array_equal(wo, wl)
    xu, wu = M.mesh_1d(x0, x1, n=37)
    xv, wv = po.mesh_1d(x0, x1, 37)
    assert np.array_equal(xu, xv) and np.array_equal(wu, wv)


def test_library_generator_rejects_what_the_oracle_rejects():
    with pytest.raises(Exception):
        M.mesh_1d(0.0, 1.0, points=[(-0.1, 0.05)])


# ------------------------------------------------------------------ finite volumes ----------

XPTS = [(0.0, 0.03), (0.5, 0.06), (1.0, 0.02)]
YPTS = [(0.0, 0.05), (0.6, 0.1)]
ZPTS = [(0.0, 0.12), (0.8, 0.06)]


def stretched(mats, bc=None, ls=False, limiter=po.LIM_VANALBADA, flux=po.FLUX_HLLC, uniform=False, **kw):
    """The oracle on a mesh stretched along all three axes (uniform=True: the uniform mesh of the
    same cell counts), with the stretched axes' (centres, widths)."""
    box = (0.0, 1.0, 0.0, 0.6, 0.0, 0.8)
    pts = (XPTS, YPTS, ZPTS)
    axes = [po.mesh_1d(box[2 * d], box[2 * d + 1], control_points=pts[d]) for d in range(3)]
    n = [len(a[0]) for a in axes]
    cfg = po.make_config(*n, box, bc or [po.BC_SYMMETRY] * 6, mats, flux=flux, limiter=limiter,
                         control_points=None if uniform else pts, ls_matid=1 if ls else None, **kw)
    return po.Oracle(cfg), axes


@pytest.mark.parametrize("limiter", [po.LIM_VANALBADA, po.LIM_GENMINMOD])
def test_linear_density_gives_no_interior_mass_flux(limiter):
    def interior_mass_residual(uniform):
        o, axes = stretched([po.sg(1.4, 0.0)], flux=po.FLUX_LLF, limiter=limiter, vanalbada_pow=False,
                            uniform=uniform)
        (x, wx), (y, wy), (z, wz) = axes
        Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
        V0 = np.zeros(X.shape + (5,))
        V0[..., 0] = 1.0 + 0.6 * X + 0.3 * Y + 0.2 * Z
        V0[..., 4] = 1.0
        V, ID = o.ghosted(V0)
        st, R = o.residual(V, ID)
        assert st == 0
        # cells whose neighbours' slopes are unlimited: two away from every wall
        return R[3:-3, 3:-3, 3:-3, 0]

    # the residual is per unit volume: a face's dissipation a/2 * area * jump over the volume is
    # a/2 * (jump / width), measured against the density gradient's jump per width
    scale = 0.5 * math.sqrt(1.4) * 0.6
    R = interior_mass_residual(False)
    assert np.max(np.abs(R)) <= 1e-12 * scale
    # the same states with the uniform coefficients A = B = 1 leave jumps at the faces
    R = interior_mass_residual(True)
    assert np.max(np.abs(R)) > 1e-4 * scale


def test_conservation_between_walls():
    o, axes = stretched([po.sg(1.4, 0.0)], ts=po.TS_RK3, cfl=0.4)
    (x, wx), (y, wy), (z, wz) = axes
    vol = wz[:, None, None] * wy[None, :, None] * wx[None, None, :]
    rng = np.random.default_rng(5)
    shp = vol.shape
    V0 = np.zeros(shp + (5,))
    V0[..., 0] = 1.0 + 0.2 * rng.uniform(size=shp)
    V0[..., 1:4] = 0.3 * rng.uniform(-1, 1, shp + (3,))
    V0[..., 4] = 1.0 + 0.2 * rng.uniform(size=shp)
    V, ID = o.ghosted(V0)

    def totals(Vg):
        U = np.array([o.p2c(v) for v in Vg[1:-1, 1:-1, 1:-1].reshape(-1, 5)]).reshape(shp + (5,))
        return (U * vol[..., None]).reshape(-1, 5).sum(axis=0)

    t0 = totals(V)
    st, t, ns = o.run(V, ID, 1e300, 5)
    assert st == 0 and ns == 5
    t1 = totals(V)
    # mass and energy are conserved (the walls' mass and energy fluxes vanish); momentum changes by
    # the wall pressure, so it is not pinned
    assert abs(t1[0] - t0[0]) <= 1e-13 * t0[0]
    assert abs(t1[4] - t0[4]) <= 1e-13 * t0[4]


def test_time_step_of_a_uniform_state_is_the_narrowest_width():
    o, axes = stretched([po.sg(1.4, 0.0)], cfl=0.5)
    shp = tuple(len(a[0]) for a in axes[::-1])
    V0 = np.zeros(shp + (5,))
    V0[..., 0] = 1.0
    V0[..., 4] = 1.0
    V, ID = o.ghosted(V0)
    st, dt = o.time_step(V, ID)
    assert st == 0
    wmin = min(a[1].min() for a in axes)
    assert abs(dt - 0.5 * wmin / math.sqrt(1.4)) <= 1e-15 * dt


# ------------------------------------------------------------------ level set ---------------

def cubic(t):
    return 0.3 - 0.7 * t + 0.25 * t * t - 0.04 * t ** 3


def dcubic(t):
    return -0.7 + 0.5 * t - 0.12 * t * t


@pytest.mark.parametrize("comp", [0, 1, 2])
def test_fd3_exact_on_cubics_on_stretched_coordinates(comp):
    o, axes = stretched([po.sg(1.4, 0.0), po.sg(1.4, 0.0)], ls=True, ls_bc=(2, 2, 2, 2, 2, 2))
    Z, Y, X = np.meshgrid(axes[2][0], axes[1][0], axes[0][0], indexing="ij")
    t = (X, Y, Z)[comp]
    phi = o.ghosted_phi(cubic(t))
    V = np.zeros(o.shape + (5,))
    V[..., 0] = V[..., 4] = 1.0
    V[..., 1 + comp] = 1.0
    R = o.ls_residual(V, phi)[1:-1, 1:-1, 1:-1]
    sl = [slice(None)] * 3
    sl[2 - comp] = slice(2, -2)
    sl = tuple(sl)
    assert np.max(np.abs(R[sl] + dcubic(t)[sl])) <= 1e-11


def test_plane_is_a_reinitialisation_fixed_point_on_stretched_mesh():
    o, axes = stretched([po.sg(1.4, 0.0), po.sg(1.4, 0.0)], ls=True, ls_bc=(2, 2, 2, 2, 2, 2), reinit_freq=1)
    Z, Y, X = np.meshgrid(axes[2][0], axes[1][0], axes[0][0], indexing="ij")
    f = (0.6 * (X - 0.47) + 0.8 * (Y - 0.31))
    phi = o.ghosted_phi(f)
    before = phi.copy()
    o.reinitialize(phi)
    assert np.max(np.abs(phi[1:-1, 1:-1, 1:-1] - before[1:-1, 1:-1, 1:-1])) <= 1e-12


def test_face_gradient_uses_the_face_position():
    o, axes = stretched([po.sg(1.4, 0.0), po.sg(1.4, 0.0)], ls=True, ls_bc=(2, 2, 2, 2, 2, 2), mp_normal=0)
    (x, wx), (y, wy), _ = axes
    Z, Y, X = np.meshgrid(axes[2][0], y, x, indexing="ij")
    phi = o.ghosted_phi(X * Y + 0.1)
    for (i, j, k) in [(5, 4, 3), (17, 2, 1), (22, 6, 2)]:
        # x-face between cells i-1 and i: d(xy)/dx = y, d(xy)/dy = x at the face x[i] - w[i]/2
        g = np.array([y[j], x[i] - 0.5 * wx[i], 0.0])
        d = o.interface_normal(phi, 0, i, j, k, 1, 1)
        assert np.max(np.abs(d - g / np.linalg.norm(g))) <= 1e-12
        g = np.array([y[j] - 0.5 * wy[j], x[i], 0.0])
        d = o.interface_normal(phi, 1, i, j, k, 1, 1)
        assert np.max(np.abs(d - g / np.linalg.norm(g))) <= 1e-12
