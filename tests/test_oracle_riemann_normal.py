"""Pins of the oracle's interface Riemann direction for MultiPhase.RiemannNormal = LevelSet / Average
(SpaceOperator::GetNormalForBimaterialRiemann, CalculateGradPhiAtCellInterface and
CalculateGradientAtCellInterface, SpaceOperator.cpp:2966-3190; oracle/riemann.c):

* a linear level set: the face difference, the central differences and their interpolation to the
  face are exact, so the direction is the plane's unit normal, turned when the minus cell is
  material 0, with its axis component made non-negative; Average is the normalised sum with the
  axis;
* a sphere: the direction approaches the radial unit vector at the face centre at second order;
* a planar interface normal to x: the level-set direction is the mesh normal to rounding (the
  central-difference weights of SpaceOperator.h:218-223 sum to zero only to rounding on the
  computed coordinates, leaving transverse components of ~1e-16), so a run with RiemannNormal = LevelSet follows the Mesh run to
  the level at which the solver's tolerance-driven iteration passes on last-bit changes (~1e-6
  after four steps), while an oblique plane changes the solution at the 1e-2 level.
"""
import numpy as np
import pytest

from oracle import pyoracle as po

N = 10


def oracle(mode, n=N, h=1.0, bc=po.BC_SYMMETRY):
    cfg = po.make_config(n, n, n, (0, n * h, 0, n * h, 0, n * h), [bc] * 6, [po.sg(1.4, 0.0), po.sg(1.67, 0.0)],
                         ls_matid=1, mp_normal=mode, ls_bc=(1,) * 6)
    return po.Oracle(cfg)


def centres(n=N, h=1.0):
    c = (np.arange(n) + 0.5) * h
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    return x, y, z


@pytest.mark.parametrize("d", [0, 1, 2])
def test_linear_level_set_gives_the_plane_normal(d):
    a = np.array([0.3, -0.5, 0.8])
    x, y, z = centres()
    o = oracle(0)
    phi = o.ghosted_phi(a[0] * x + a[1] * y + a[2] * z - 2.0)
    g = a / np.linalg.norm(a)
    q = [4, 5, 6]
    for myid, nbid in ((1, 0), (0, 1)):
        n = o.interface_normal(phi, d, *q, myid, nbid)
        e = g * (-1.0 if nbid == 0 else 1.0)
        e[d] = abs(e[d])                       # only the axis component is turned
        assert np.allclose(n, e, rtol=0, atol=1e-14), (myid, nbid, n, e)
    oa = oracle(2)
    n = oa.interface_normal(phi, d, *q, 1, 0)
    e = -g
    e[d] = abs(e[d]) + 1.0
    assert np.allclose(n, e / np.linalg.norm(e), rtol=0, atol=1e-14)
    assert np.array_equal(oracle(1).interface_normal(phi, d, *q, 1, 0), np.eye(3)[d])


def test_sphere_normal_converges_at_second_order():
    errs = []
    for n in (16, 32):
        h = 1.0 / n
        x, y, z = centres(n, h)
        c = np.array([0.5, 0.5, 0.5])
        o = oracle(0, n, h)
        phi = o.ghosted_phi(np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) - 0.3)
        # the x-face of cell (i, j, k) near the sphere, at x = i h, y = (j + 1/2) h, z = (k + 1/2) h
        i, j, k = int(0.8 * n), int(0.6 * n), int(0.55 * n)
        p = np.array([i * h, (j + 0.5) * h, (k + 0.5) * h])
        r = (p - c) / np.linalg.norm(p - c)
        got = o.interface_normal(phi, 0, i, j, k, 0, 1)    # minus cell material 1: no turn
        errs.append(np.max(np.abs(got - r)))
    assert errs[1] < errs[0] / 3.0 and errs[1] < 5e-3, errs


def planar_case(tilt):
    n = 12
    h = 1.0 / n
    x, y, z = centres(n, h)
    phi_int = (x - 0.5 - tilt * (y - 0.5)) / np.sqrt(1 + tilt * tilt)    # air (material 1) where phi < 0
    V = np.zeros((n, n, n, 5))
    V[..., 0] = np.where(phi_int < 0, 0.5, 1.0)
    V[..., 1] = 0.1
    V[..., 2] = 0.05
    V[..., 4] = np.where(phi_int < 0, 2.0, 1.0)
    ID = (phi_int < 0).astype(np.int32)
    return n, h, V, ID, phi_int


@pytest.mark.parametrize("tilt", [0.0, 0.4])
def test_planar_interface_level_set_normal_against_the_mesh_normal(tilt):
    n, h, V0, ID0, phi0 = planar_case(tilt)
    out = {}
    for mode in (1, 0):
        cfg = po.make_config(n, n, n, (0, 1, 0, 1, 0, 1), [po.BC_SYMMETRY] * 6, [po.sg(1.4, 0.0), po.sg(1.67, 0.0)],
                             ls_matid=1, mp_normal=mode, ts=po.TS_RK2, reinit_freq=-1)
        o = po.Oracle(cfg)
        V, ID = o.ghosted(V0, ID0)
        phi = o.ghosted_phi(phi0)
        st, t, k = o.run(V, ID, 1e300, 4, phi=phi)
        assert st == 0 and k == 4
        out[mode] = V
    diff = float(np.max(np.abs(out[0] - out[1])[1:-1, 1:-1, 1:-1]))
    if tilt == 0.0:
        assert diff <= 1e-5, diff
        o = oracle(0, n, h)
        ph = o.ghosted_phi(phi0)
        for j in range(n):
            nrm = o.interface_normal(ph, 0, n // 2, j, 3, 0, 1)
            assert abs(nrm[0] - 1.0) <= 2.3e-16 and abs(nrm[1]) <= 1e-15 and abs(nrm[2]) <= 1e-15
    else:
        assert diff >= 1e-3, diff
