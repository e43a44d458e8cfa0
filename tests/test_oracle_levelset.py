"""Pins of the oracle's level-set restatement (oracle/multiphase.c) against what the mathematics
fixes: third-order one-sided differences are exact on cubics (GradientCalculatorFD3.cpp:54-171,
300-469), including the reference's right-biased y stencil that is third order only in columns
i < NY-1 (GradientCalculatorFD3.cpp:415); the signed distance of a plane is a fixed point of the
reinitialisation and a steeper plane is brought back to unit slope with its first layer fixed
(LevelSetReinitializer.cpp:74-172); the material-ID rule of MultiPhaseOperator.cpp:127-330 on
hand-worked cases; the linear-extrapolation ghost rule reproduces linear fields."""
import numpy as np
import pytest

from oracle import pyoracle as po

NX, NY, NZ = 16, 12, 10
H = 0.5


def make(ls_bc=(1, 1, 1, 1, 1, 1), nx=NX, ny=NY, nz=NZ, mats=2):
    cfg = po.make_config(nx, ny, nz, (0.0, nx * H, 0.0, ny * H, 0.0, nz * H), [po.BC_SYMMETRY] * 6,
                         [po.sg(1.4, 0.0) for _ in range(mats)], ls_matid=1, ls_bc=ls_bc, reinit_freq=1)
    return po.Oracle(cfg)


def centres(n):
    return (np.arange(n) + 0.5) * H


def cubic(t):
    return 0.3 - 0.7 * t + 0.25 * t * t - 0.04 * t ** 3


def dcubic(t):
    return -0.7 + 0.5 * t - 0.12 * t * t


def velocity_field(o, comp, sign):
    V = np.zeros(o.shape + (5,))
    V[..., 0] = 1.0
    V[..., 4] = 1.0
    V[..., 1 + comp] = sign
    return V


@pytest.mark.parametrize("comp", [0, 1, 2])
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_fd3_exact_on_cubics(comp, sign):
    o = make(ls_bc=(2, 2, 2, 2, 2, 2))
    z, y, x = np.meshgrid(centres(NZ), centres(NY), centres(NX), indexing="ij")
    t = (x, y, z)[comp]
    phi = o.ghosted_phi(cubic(t))
    R = o.ls_residual(velocity_field(o, comp, sign), phi)[1:-1, 1:-1, 1:-1]
    expect = -sign * dcubic(t)
    n = (NX, NY, NZ)[comp]
    # interior nodes whose 4-point stencil stays inside the box
    lo, hi = (2, n - 2) if sign > 0 else (1, n - 3)
    sl = [slice(None)] * 3
    sl[2 - comp] = slice(lo, hi + 1)
    got, exp = R[tuple(sl)], expect[tuple(sl)]
    if comp == 1 and sign < 0:
        # third order only in columns i < NY-1
        good = np.arange(NX) < NY - 1
        assert np.max(np.abs(got[..., good] - exp[..., good])) <= 1e-12
        # first-order difference (phi_{j+1} - phi_j)/dy elsewhere
        yy = centres(NY)
        fo = (cubic(yy[1:]) - cubic(yy[:-1])) / H
        fo_exp = np.broadcast_to(-sign * fo[lo:hi + 1][None, :, None], got[..., ~good].shape)
        assert np.max(np.abs(got[..., ~good] - fo_exp)) <= 1e-12
        assert np.max(np.abs(got[..., ~good] - exp[..., ~good])) > 1e-3
    else:
        assert np.max(np.abs(got - exp)) <= 1e-12


def test_linear_extrapolation_ghosts_reproduce_linear_field():
    o = make(ls_bc=(2, 2, 2, 2, 2, 2))
    z, y, x = np.meshgrid(centres(NZ), centres(NY), centres(NX), indexing="ij")
    f = 0.5 + 0.25 * x - 0.125 * y + 0.375 * z
    phi = o.ghosted_phi(f)
    assert abs(phi[1, 1, 0] - (0.5 + 0.25 * (-0.5 * H) - 0.125 * y[0, 0, 0] + 0.375 * z[0, 0, 0])) <= 1e-13
    assert abs(phi[1, 1, -1] - (0.5 + 0.25 * (NX + 0.5) * H - 0.125 * y[0, 0, 0] + 0.375 * z[0, 0, 0])) <= 1e-13
    assert abs(phi[0, 3, 4] - (0.5 + 0.25 * x[0, 0, 3] - 0.125 * y[0, 2, 0] + 0.375 * (-0.5 * H))) <= 1e-13


def test_zero_neumann_and_nonnegative_ghosts():
    o = make(ls_bc=(1, 3, 1, 1, 1, 1))
    f = np.full((NZ, NY, NX), -5.0)
    phi = o.ghosted_phi(f)
    assert phi[3, 3, 0] == -5.0                       # zero Neumann: image value
    assert phi[3, 3, -1] == -0.5 * H                  # non-negative: no deeper than the wall distance


def test_plane_is_a_reinitialisation_fixed_point():
    o = make(ls_bc=(2, 2, 2, 2, 2, 2))
    z, y, x = np.meshgrid(centres(NZ), centres(NY), centres(NX), indexing="ij")
    phi = o.ghosted_phi(x - 3.3)
    before = phi.copy()
    its = o.reinitialize(phi)
    assert its == 0
    assert np.array_equal(phi[1:-1, 1:-1, 1:-1], before[1:-1, 1:-1, 1:-1])


def test_steep_plane_reinitialised_to_unit_slope_first_layer_fixed():
    o = make(ls_bc=(2, 2, 2, 2, 2, 2))
    z, y, x = np.meshgrid(centres(NZ), centres(NY), centres(NX), indexing="ij")
    f = 2.0 * (x - 3.3)
    phi = o.ghosted_phi(f)
    its = o.reinitialize(phi)
    assert its > 0
    new = phi[1:-1, 1:-1, 1:-1]
    # first layer (the two columns around the zero crossing at x = 3.3) keeps its values
    first = (np.abs(x - 3.3) < H)
    assert np.array_equal(new[first], f[first])
    # signs are preserved and away from the interface the slope is 1
    assert np.all(np.sign(new) == np.sign(f))
    gx = np.diff(new, axis=2) / H
    far = (np.abs(x[..., 1:] - 3.3) > 2 * H) & (np.abs(x[..., :-1] - 3.3) > 2 * H)
    assert np.max(np.abs(gx[far] - 1.0)) < 2e-2


def test_sphere_distance_nearly_fixed():
    o = make(ls_bc=(2, 2, 2, 2, 2, 2))
    z, y, x = np.meshgrid(centres(NZ), centres(NY), centres(NX), indexing="ij")
    f = np.sqrt((x - 4.0) ** 2 + (y - 3.0) ** 2 + (z - 2.5) ** 2) - 1.6
    phi = o.ghosted_phi(f)
    o.reinitialize(phi)
    new = phi[1:-1, 1:-1, 1:-1]
    assert np.all(np.sign(new) == np.sign(f))
    # first-order upwind Hamiltonian: exact distances change little within a cell of the interface
    # and by a first-order amount (a fraction of the distance) further out
    near = np.abs(f) < H
    assert np.max(np.abs(new - f)[near]) <= 0.2 * H
    assert np.all(np.abs(new - f) <= 0.1 * np.abs(f) + 0.2 * H)


def test_material_id_rule_hand_worked():
    o = make(nx=8, ny=1, nz=1, mats=3)
    ID = np.zeros(o.shape, dtype=np.int32)
    ID[1, 1, 1:-1] = [0, 0, 1, 1, 1, 1, 2, 2]
    phi0 = np.zeros(o.shape)
    phi0[1, 1, 1:-1] = [1, 1, -1, -1, -1, -1, 1, 1]
    # cell 1 enters the level-set subdomain; cells 3, 4, 5 leave it: in pass 1 (i order) only 5
    # has an untracked neighbour (material 2), pass 2 then resolves 4 from 5 and pass 3 resolves 3
    phi = np.zeros(o.shape)
    phi[1, 1, 1:-1] = [1, -1, -1, 1, 1, 1, 1, 1]
    o.update_id(phi0, phi, ID)
    assert list(ID[1, 1, 1:-1]) == [0, 1, 1, 2, 2, 2, 2, 2]
    # a node leaving between two tracked nodes has no candidate: orphan, ID 0
    ID[1, 1, 1:-1] = [2, 1, 1, 1, 1, 1, 1, 2]
    phi0[1, 1, 1:-1] = [1, -1, -1, -1, -1, -1, -1, 1]
    phi[1, 1, 1:-1] = [1, -1, -1, 1, -1, -1, -1, 1]
    o.update_id(phi0, phi, ID)
    assert list(ID[1, 1, 1:-1]) == [2, 1, 1, 0, 1, 1, 1, 2]
    # a node leaving next to both 0 and 2 takes the smallest (std::set order)
    ID[1, 1, 1:-1] = [0, 0, 1, 2, 2, 2, 2, 2]
    phi0[1, 1, 1:-1] = [1, 1, -1, 1, 1, 1, 1, 1]
    phi[1, 1, 1:-1] = [1, 1, 1, 1, 1, 1, 1, 1]
    o.update_id(phi0, phi, ID)
    assert ID[1, 1, 3] == 0
