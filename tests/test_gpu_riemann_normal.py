"""GPU parity of the interface Riemann direction MultiPhase.RiemannNormal = LevelSet / Average
(SpaceOperator::GetNormalForBimaterialRiemann and CalculateGradPhiAtCellInterface,
SpaceOperator.cpp:2966-3190): the interface faces of the shock-bubble and of an oblique planar
interface, solved along the level-set gradient of Phi(n) at the face (or its average with the
axis), bit-identical to the oracle over several RK2/RK3 steps with reinitialisation."""
import numpy as np
import pytest

import m2c_amd as M
from m2c_amd import workloads as W
from tests.cases import run_oracle

pytestmark = pytest.mark.gpu


def gpu_run(settings, V0, ID0, phi0, steps):
    s = M.Solver(W.config_from(settings))
    s.upload(V0, ID0)
    s.upload_levelset(phi0)
    st = s.run(1e300, steps, sync_every=0)
    V, ID = s.download()
    phi = s.download_levelset()
    s.close()
    return st, V, ID, phi


@pytest.mark.parametrize("mode,ts", [(0, M.TS_RK2), (2, M.TS_RK2), (0, M.TS_RK3)])
def test_shock_bubble_bitwise(mode, ts):
    settings, V0, ID0, phi0 = W.shock_bubble(scale=0.0625, reinit_freq=3)
    settings = dict(settings, mp_normal=mode, ts=ts)
    print(f"case: shock-bubble 24x12x12 RiemannNormal {mode} ts {ts}, 8 steps", flush=True)
    st, V, ID, phi = gpu_run(settings, V0, ID0, phi0, 8)
    ost, t, n, Vo, _, IDo, phio = run_oracle(settings, V0, ID0, maxsteps=8, phi0=phi0, vanalbada_pow=False)
    assert st.error == 0 and ost == 0 and n == st.steps
    assert np.array_equal(ID, IDo) and np.array_equal(phi, phio)
    assert np.array_equal(V, Vo), f"max diff {np.max(np.abs(V - Vo))}"


def test_oblique_plane_bitwise_and_not_the_mesh_run():
    n = 16
    c = (np.arange(n) + 0.5) / n
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    phi0 = (x - 0.5 - 0.4 * (y - 0.5) + 0.2 * (z - 0.5)) / np.sqrt(1.2)
    V0 = np.zeros((n, n, n, 5))
    V0[..., 0] = np.where(phi0 < 0, 0.5, 1.0)
    V0[..., 1] = 0.1
    V0[..., 4] = np.where(phi0 < 0, 2.0, 1.0)
    ID0 = (phi0 < 0).astype(np.int32)
    settings = dict(nx=n, ny=n, nz=n, box=(0, 1, 0, 1, 0, 1), bc=[M.BC_SYMMETRY] * 6,
                    mats=[M.material(M.EOS_SG, [1.4, 0.0]), M.material(M.EOS_SG, [1.67, 0.0])], ls_matid=1,
                    ts=M.TS_RK2, reinit_freq=2, mp_normal=0)
    print("case: oblique plane 16^3 RiemannNormal LevelSet, 6 steps", flush=True)
    st, V, ID, phi = gpu_run(settings, V0, ID0, phi0, 6)
    ost, t, k, Vo, _, IDo, phio = run_oracle(settings, V0, ID0, maxsteps=6, phi0=phi0, vanalbada_pow=False)
    assert st.error == 0 and ost == 0 and k == st.steps
    assert np.array_equal(ID, IDo) and np.array_equal(phi, phio) and np.array_equal(V, Vo)
    st, Vm, _, _ = gpu_run(dict(settings, mp_normal=1), V0, ID0, phi0, 6)
    assert np.max(np.abs(Vm - V)) > 1e-3


def test_rejections():
    settings, V0, ID0, phi0 = W.shock_bubble(scale=0.0625, reinit_freq=3)
    # no level set, an out-of-range code, and the learned solver (it solves along the axes)
    with pytest.raises(M.M2CError):
        M.Solver(W.config_from(dict(settings, mp_normal=0, ls_matid=None)))
    with pytest.raises(M.M2CError):
        M.Solver(W.config_from(dict(settings, mp_normal=3)))
    s = M.Solver(W.config_from(dict(settings, mp_normal=2)))
    net = M.load_surrogate()
    with pytest.raises(M.M2CError):
        s.set_surrogate(net)
    net.close()
    s.close()
