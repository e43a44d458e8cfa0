"""Mesh faces cut by embedded walls inside the sweep (m2c_set_embedded_faces) against the CPU oracle,
bit for bit: the residual of a sphere obstacle in a supersonic stream over the fluxes, the three
one-sided Riemann normals, both wall reconstructions and slope zeroing, then RK2/RK3 runs with a
moving wall; the mixed form within the 1e-5 relative L1 bar. The reference path is
SpaceOperator::ComputeAdvectionFluxes with EmbeddedBoundaryDataSets (SpaceOperator.cpp:2438-2484,
2569-2611, 2696-2737) and Reconstructor.cpp:656-664; the oracle's cut-face code is pinned in
test_oracle_embedded.py."""
import numpy as np
import pytest

import m2c_amd as M
from m2c_amd import workloads as W
from oracle import pyoracle as po
from tests.cases import oracle_config

pytestmark = pytest.mark.gpu


def oracle_for(settings, faces, walls, mode, recon, V0, ID0, **over):
    o = po.Oracle(oracle_config(settings, vanalbada_pow=False, **over))
    V, ID = o.ghosted(V0, ID0)
    assert o.set_embedded_faces(faces, walls, normal_mode=mode, recon=recon) == 0
    return o, V, ID


def solver_for(settings, faces, walls, mode, recon, V0, ID0, **over):
    s = M.Solver(W.config_from(settings, **over))
    s.upload(V0, ID0)
    s.set_embedded_faces(faces, walls, normal_mode=mode, recon=recon)
    return s


@pytest.mark.parametrize("flux", [M.FLUX_HLLC, M.FLUX_ROE, M.FLUX_LLF, M.FLUX_GODUNOV])
@pytest.mark.parametrize("mode,recon,slope_zero", [(1, 0, False), (0, 1, False), (2, 1, True), (0, 0, True)])
def test_residual_bitwise(flux, mode, recon, slope_zero):
    settings, V0, ID0, faces, walls = W.sphere_obstacle(24, 18, 16, radius=4.5, vwall=(0.05, -0.02, 0.01))
    print(f"case: sphere 24x18x16 flux {flux} mode {mode} recon {recon} slope_zero {slope_zero}, "
          f"{len(faces)} cut faces", flush=True)
    s = solver_for(settings, faces, walls, mode, recon, V0, ID0, flux=flux, slope_zero=slope_zero)
    R = s.residual()
    st = s.status()
    s.close()
    o, V, ID = oracle_for(settings, faces, walls, mode, recon, V0, ID0, flux=flux, slope_zero=slope_zero)
    ost, Ro = o.residual(V, ID)
    assert ost >= 0
    Ro = Ro[1:-1, 1:-1, 1:-1]
    assert np.array_equal(R, Ro), f"max diff {np.max(np.abs(R - Ro))}"
    # the walls matter: without them the residual next to the sphere differs
    o2 = po.Oracle(oracle_config(settings, vanalbada_pow=False, flux=flux, slope_zero=slope_zero))
    _, R2 = o2.residual(V, ID)
    assert not np.array_equal(R2[1:-1, 1:-1, 1:-1], Ro)
    assert st.error == 0


@pytest.mark.parametrize("ts", [M.TS_RK2, M.TS_RK3])
def test_run_bitwise(ts):
    settings, V0, ID0, faces, walls = W.sphere_obstacle(32, 20, 20, radius=5.0, vwall=(0.1, 0.0, 0.0))
    steps = 15
    print(f"case: sphere 32x20x20 ts {ts} {steps} steps, {len(faces)} cut faces", flush=True)
    s = solver_for(settings, faces, walls, 2, 1, V0, ID0, ts=ts, slope_zero=True)
    st = s.run(1e300, steps, sync_every=0)
    V, ID = s.download()
    s.close()
    assert st.error == 0 and st.steps == steps
    o, Vo, IDo = oracle_for(settings, faces, walls, 2, 1, V0, ID0, ts=ts, slope_zero=True)
    ost, t, n = o.run(Vo, IDo, 1e300, steps)
    assert ost == 0 and n == steps
    assert st.t == t
    Vo = Vo[1:-1, 1:-1, 1:-1]
    assert np.array_equal(V, Vo), f"max diff {np.max(np.abs(V - Vo))}"
    assert np.array_equal(ID, ID0)
    # the flow has moved: a bow shock ahead of the sphere
    assert np.abs(V - V0).max() > 0.1


def test_mixed_form_with_walls():
    settings, V0, ID0, faces, walls = W.sphere_obstacle(32, 20, 20, radius=5.0)
    steps = 15
    print(f"case: sphere 32x20x20 mixed {steps} steps", flush=True)
    s = solver_for(settings, faces, walls, 1, 0, V0, ID0, precision=M.PREC_MIXED)
    st = s.run(1e300, steps, sync_every=0)
    V, _ = s.download()
    s.close()
    assert st.error == 0
    o, Vo, IDo = oracle_for(settings, faces, walls, 1, 0, V0, ID0)
    o.run(Vo, IDo, 1e300, steps)
    Vo = Vo[1:-1, 1:-1, 1:-1]
    act = ID0 == 0
    for f in (0, 4):
        l1 = np.abs(V[..., f] - Vo[..., f])[act].sum() / np.abs(Vo[..., f])[act].sum()
        print(f"  field {f}: L1 {l1:.2e}", flush=True)
        assert l1 <= 1e-5


def test_walls_removed_and_rejected():
    settings, V0, ID0, faces, walls = W.sphere_obstacle(16, 12, 12, radius=3.0)
    s = solver_for(settings, faces, walls, 1, 0, V0, ID0)
    R1 = s.residual()
    s.set_embedded_faces([], [])
    R0 = s.residual()
    o = po.Oracle(oracle_config(settings, vanalbada_pow=False))
    V, ID = o.ghosted(V0, ID0)
    _, Ro = o.residual(V, ID)
    assert np.array_equal(R0, Ro[1:-1, 1:-1, 1:-1])
    assert not np.array_equal(R0, R1)
    bad = faces.copy()
    bad[0, 1 + bad[0, 0]] = 0                     # a domain-boundary face
    with pytest.raises(M.M2CError):
        s.set_embedded_faces(bad, walls)
    with pytest.raises(M.M2CError):
        s.set_embedded_faces(np.concatenate([faces, faces[:1]]), np.concatenate([walls, walls[:1]]))
    s.close()
    # a level set (interfaces) and periodic axes are not combined with walls
    sb, Vb, IDb, phib = W.shock_bubble(scale=0.125)
    s2 = M.Solver(W.config_from(sb))
    with pytest.raises(M.M2CError):
        s2.set_embedded_faces([[0, 4, 4, 4]], [np.r_[1.0, 0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 0]])
    s2.close()


def piston_case(nx=40, ny=6, nz=5, xw=25.3, uw=-0.2):
    """Gas left of a plane wall x = xw moving at uw along x (into the gas for uw < 0), the cells
    beyond it inactive; the wall given as a triangulated surface (two triangles) and its crossings
    with the x-edges between the last gas cell and the first solid cell of every row, as an
    embedded-boundary operator hands them over (triangle, barycentric coordinates, distance from the
    edge's first node)."""
    gas = M.material(M.EOS_SG, [1.4, 0.0])
    settings = dict(nx=nx, ny=ny, nz=nz, box=(0.0, float(nx), 0.0, float(ny), 0.0, float(nz)),
                    bc=[M.BC_INLET, M.BC_OUTLET, M.BC_SLIPWALL, M.BC_SLIPWALL, M.BC_SYMMETRY, M.BC_SYMMETRY],
                    mats=[gas, gas], inactive_id=1, inlet=(1.0, 0.0, 0.0, 0.0, 1.0), outlet=(1.0, 0.0, 0.0, 0.0, 1.0),
                    flux=M.FLUX_HLLC, limiter=M.LIM_VANALBADA, ts=M.TS_RK2, cfl=0.5)
    x = np.arange(nx) + 0.5
    ID = np.broadcast_to((x > xw).astype(np.int32), (nz, ny, nx)).copy()
    rng = np.random.default_rng(4)
    V = np.zeros((nz, ny, nx, 5))
    V[..., 0] = 1.0 + 0.02 * rng.uniform(-1, 1, (nz, ny, nx))
    V[..., 4] = 1.0 + 0.02 * rng.uniform(-1, 1, (nz, ny, nx))
    V[ID == 1, 1:4] = 0.0
    nodes = np.array([[xw, 0.0, 0.0], [xw, ny, 0.0], [xw, ny, nz], [xw, 0.0, nz]])
    elems = np.array([[0, 1, 2], [0, 2, 3]])
    surf = dict(udot=np.tile([uw, 0.0, 0.0], (4, 1)), elems=elems, elem_norm=[[-1.0, 0.0, 0.0]] * 2,
                tid=[], xi=[], dist=[])
    i0 = int(np.argmax(x > xw))
    faces, tid, xi, dist = [], [], [], []
    for k in range(nz):
        for j in range(ny):
            p = np.array([j + 0.5, k + 0.5])
            for t, e in enumerate(elems):
                a, b, c = (nodes[e][:, 1:] - 0.0)
                T = np.array([b - a, c - a]).T
                l1, l2 = np.linalg.solve(T, p - a)
                if l1 >= 0 and l2 >= 0 and l1 + l2 <= 1:
                    tid.append(t)
                    xi.append([1.0 - l1 - l2, l1, l2])
                    break
            dist.append(xw - x[i0 - 1])
            faces.append([0, i0, j, k])
    surf.update(tid=tid, xi=xi, dist=dist)
    idx = np.arange(len(faces)).reshape(-1, 1)
    return settings, V, ID, np.array(faces, np.int32), [surf], idx, idx


def test_walls_from_surface_crossings_bitwise():
    settings, V0, ID0, faces, surfaces, xf, xb = piston_case()
    cut, walls = M.edge_surface_walls(surfaces, xf, xb)
    co, wo = po.edge_surface_walls(surfaces, xf, xb)
    assert cut.all() and np.array_equal(walls, wo)
    assert np.allclose(walls[:, 3:6], [-0.2, 0.0, 0.0], rtol=0, atol=1e-15)
    assert np.array_equal(walls[:, 0:3], np.tile([-1.0, 0.0, 0.0], (len(faces), 1)))
    print(f"case: piston 40x6x5, {len(faces)} cut faces from surface crossings, 12 steps", flush=True)
    s = solver_for(settings, faces, walls, 0, 1, V0, ID0)
    st = s.run(1e300, 12, sync_every=0)
    V, _ = s.download()
    s.close()
    o, Vo, IDo = oracle_for(settings, faces, walls, 0, 1, V0, ID0)
    ost, t, n = o.run(Vo, IDo, 1e300, 12)
    assert st.error == 0 and ost >= 0 and n == 12 and st.t == t
    Vo = Vo[1:-1, 1:-1, 1:-1]
    assert np.array_equal(V, Vo), f"max diff {np.max(np.abs(V - Vo))}"
    # the moving wall compresses the gas next to it
    assert V[:, :, 24, 0].mean() > 1.02
