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
