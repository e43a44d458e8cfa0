"""GPU parity of linear reconstruction at material interfaces (MultiPhase.ReconstructionAtInterface =
Linear, SpaceOperator.cpp:2504-2507, 2529-2535): the interface Riemann pass solves between the
reconstructed face states (formed as the stage kernel forms them) and takes its one-sided fluxes
from them; bit-identical to the oracle with the numerical and the exact interface flux, the
level-set interface normal, and without a level set."""
import numpy as np
import pytest

import m2c_amd as M
from m2c_amd import workloads as W
from tests.cases import run_oracle, smooth_state

pytestmark = pytest.mark.gpu


def gpu_run(settings, V0, ID0, phi0, steps):
    s = M.Solver(W.config_from(settings))
    s.upload(V0, ID0)
    if phi0 is not None:
        s.upload_levelset(phi0)
    st = s.run(1e300, steps, sync_every=0)
    V, ID = s.download()
    phi = s.download_levelset() if phi0 is not None else None
    s.close()
    return st, V, ID, phi


@pytest.mark.parametrize("mp_flux,normal,ts", [(0, 1, M.TS_RK2), (1, 1, M.TS_RK2), (0, 0, M.TS_RK3)])
def test_shock_bubble_bitwise(mp_flux, normal, ts):
    settings, V0, ID0, phi0 = W.shock_bubble(scale=0.0625, reinit_freq=3)
    settings = dict(settings, mp_recon=1, mp_flux=mp_flux, mp_normal=normal, ts=ts)
    print(f"case: shock-bubble linear interface recon, flux {mp_flux}, normal {normal}, ts {ts}", flush=True)
    st, V, ID, phi = gpu_run(settings, V0, ID0, phi0, 8)
    ost, t, n, Vo, _, IDo, phio = run_oracle(settings, V0, ID0, maxsteps=8, phi0=phi0, vanalbada_pow=False)
    assert st.error == 0 and ost == 0 and n == st.steps
    assert np.array_equal(ID, IDo) and np.array_equal(phi, phio)
    assert np.array_equal(V, Vo), f"max diff {np.max(np.abs(V - Vo))}"
    st, V0c, _, _ = gpu_run(dict(settings, mp_recon=0), V0, ID0, phi0, 8)
    assert np.max(np.abs(V0c - V)) > 0.0


def test_two_gases_without_level_set_bitwise():
    nx, ny, nz = 20, 10, 6
    V0 = smooth_state(nx, ny, nz, seed=4, amp=0.1)
    c = (np.arange(nx) + 0.5) / nx
    ID0 = np.broadcast_to((c > 0.5).astype(np.int32), (nz, ny, nx)).copy()
    settings = dict(nx=nx, ny=ny, nz=nz, box=(0.0, 1.0, 0.0, 0.5, 0.0, 0.3),
                    bc=[M.BC_SYMMETRY, M.BC_OUTLET, M.BC_SLIPWALL, M.BC_SLIPWALL, M.BC_PERIODIC, M.BC_PERIODIC],
                    mats=[M.material(M.EOS_SG, [1.4, 0.0]), M.material(M.EOS_SG, [1.67, 0.0])],
                    outlet=(1.0, 0.0, 0.0, 0.0, 1.0), mp_recon=1, ts=M.TS_RK2)
    print("case: two gases, linear interface recon, 6 steps", flush=True)
    st, V, _, _ = gpu_run(settings, V0, ID0, None, 6)
    ost, t, n, Vo, _ = run_oracle(settings, V0, ID0, maxsteps=6, vanalbada_pow=False)
    assert st.error == 0 and ost == 0 and n == st.steps
    assert np.array_equal(V, Vo), f"max diff {np.max(np.abs(V - Vo))}"


def test_learned_solver_rejected():
    settings, V0, ID0, phi0 = W.shock_bubble(scale=0.0625, reinit_freq=3)
    s = M.Solver(W.config_from(dict(settings, mp_recon=1)))
    net = M.load_surrogate()
    with pytest.raises(M.M2CError):
        s.set_surrogate(net)
    net.close()
    s.close()


@pytest.mark.parametrize("mp_recon,flux", [(0, M.FLUX_HLLC), (1, M.FLUX_HLLC), (0, M.FLUX_GODUNOV)])
def test_constant_reconstruction_depth_bitwise(mp_recon, flux):
    # MultiPhase.ConstantReconstructionDepth: cells within 2.5 cells of the interface (|phi(t_n)|)
    # reconstruct constantly, in the stage kernel, the interface pass and the Godunov pass alike
    settings, V0, ID0, phi0 = W.shock_bubble(scale=0.0625, reinit_freq=3)
    V0 = V0.copy()
    nz, ny, nx = ID0.shape
    x = (np.arange(nx) + 0.5) / nx
    V0[..., 0] *= 1.0 + 0.05 * np.sin(2 * np.pi * x)
    V0[..., 1] += 0.5 * np.cos(2 * np.pi * x)
    h = settings["box"][1] / settings["nx"]
    settings = dict(settings, conrec_depth=2.5 * h, mp_recon=mp_recon, flux=flux)
    print(f"case: shock-bubble constant-reconstruction depth, interface recon {mp_recon}, flux {flux}", flush=True)
    st, V, ID, phi = gpu_run(settings, V0, ID0, phi0, 6)
    ost, t, n, Vo, _, IDo, phio = run_oracle(settings, V0, ID0, maxsteps=6, phi0=phi0, vanalbada_pow=False)
    assert st.error == 0 and ost == 0 and n == st.steps
    assert np.array_equal(ID, IDo) and np.array_equal(phi, phio)
    assert np.array_equal(V, Vo), f"max diff {np.max(np.abs(V - Vo))}"
