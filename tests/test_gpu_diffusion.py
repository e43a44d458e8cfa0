"""GPU parity of the viscous and heat-diffusion residual terms against the oracle.

Bar: bit-identical FP64 for the Constant viscosity model and heat diffusion (residual, RK2/RK3 runs,
two materials with a level set); Sutherland's law calls pow(), which the device library and glibc
round differently in the last place, so there the bar is 1e-12 relative.
"""
import numpy as np
import pytest

import m2c_amd as M
from m2c_amd import workloads as W
from oracle import pyoracle as po
from tests.cases import oracle_config, rel_linf, run_oracle, smooth_state

pytestmark = pytest.mark.gpu


def visc_gas(mu=0.004, bulk=0.001, kappa=0.006, cv=2.5):
    # explicit steps stay inside the diffusive limit (dt < h^2 rho / (6 mu)) at CFL 0.4 on h ~ 0.05
    return M.material(M.EOS_SG, [1.4, 0.0], tlaw=(cv, 0.0, 0.0, 0.0, 0.0, 0.0), viscosity=M.VISC_CONSTANT, mu=mu,
                      bulk=bulk, kappa=kappa)


def walls_case(mats, nx=21, ny=13, nz=9, ts=M.TS_RK2, seed=3):
    V0 = smooth_state(nx, ny, nz, seed=seed)
    settings = dict(nx=nx, ny=ny, nz=nz, box=(0.0, 1.0, 0.0, 0.6, 0.0, 0.4),
                    bc=[M.BC_INLET, M.BC_OUTLET, M.BC_SLIPWALL, M.BC_STICKWALL, M.BC_SYMMETRY, M.BC_SLIPWALL],
                    mats=mats, inlet=(1.2, 0.3, 0.0, 0.0, 1.1), outlet=(0.9, 0.1, 0.0, 0.0, 0.95),
                    flux=M.FLUX_HLLC, limiter=M.LIM_VANALBADA, ts=ts, cfl=0.4)
    return settings, V0, np.zeros((nz, ny, nx), dtype=np.int32)


def test_residual_bitwise():
    settings, V0, ID0 = walls_case([visc_gas(mu=0.02, bulk=0.005, kappa=0.03)])
    print("case: viscous + heat residual 21x13x9", flush=True)
    s = M.Solver(W.config_from(settings))
    s.upload(V0, ID0)
    R = s.residual()
    s.close()
    o = po.Oracle(oracle_config(settings, vanalbada_pow=False))
    Vg, IDg = o.ghosted(V0, ID0)
    st, Ro = o.residual(Vg, IDg)
    assert st == 0
    Ro = Ro[1:-1, 1:-1, 1:-1]
    assert np.array_equal(R, Ro), f"max diff {np.max(np.abs(R - Ro))}"
    # and the terms matter: the inviscid residual differs
    o0 = po.Oracle(oracle_config(dict(settings, mats=[M.material(M.EOS_SG, [1.4, 0.0])]), vanalbada_pow=False))
    st, R0 = o0.residual(*o0.ghosted(V0, ID0))
    assert np.max(np.abs(R0[1:-1, 1:-1, 1:-1] - Ro)) > 1e-3


@pytest.mark.parametrize("ts", [M.TS_RK2, M.TS_RK3])
def test_runs_bitwise(ts):
    settings, V0, ID0 = walls_case([visc_gas(mu=0.005, bulk=0.0, kappa=0.008)], ts=ts, seed=5)
    print(f"case: viscous + heat run ts={ts} 12 steps", flush=True)
    s = M.Solver(W.config_from(settings))
    s.upload(V0, ID0)
    st = s.run(1e300, 12, sync_every=4)
    V, _ = s.download()
    s.close()
    ost, t, n, Vo, _ = run_oracle(settings, V0, ID0, maxsteps=12, vanalbada_pow=False)
    assert st.error == 0 and ost == 0 and n == st.steps == 12 and t == st.t
    assert np.array_equal(V, Vo), f"max diff {np.max(np.abs(V - Vo))}"


def test_two_materials_with_a_level_set_bitwise():
    settings, V0, ID0, phi0 = W.shock_bubble(scale=0.0625, reinit_freq=3)
    water = M.material(M.EOS_SG, list(W.WATER), tlaw=(4180.0, 0.0, 0.0, 0.0, 0.0, 0.0), viscosity=M.VISC_CONSTANT,
                       mu=1.0e-3, bulk=0.0, kappa=0.6)
    air = M.material(M.EOS_SG, list(W.AIR), tlaw=(717.5, 0.0, 0.0, 0.0, 0.0, 0.0), viscosity=M.VISC_CONSTANT,
                     mu=1.8e-5, bulk=1.0e-5, kappa=0.026)
    settings = dict(settings, mats=[water, air])
    print("case: viscous conducting shock-bubble 24x12x12, 8 steps", flush=True)
    s = M.Solver(W.config_from(settings))
    s.upload(V0, ID0)
    s.upload_levelset(phi0)
    st = s.run(1e300, 8, sync_every=0)
    V, ID = s.download()
    phi = s.download_levelset()
    s.close()
    ost, t, n, Vo, _, IDo, phio = run_oracle(settings, V0, ID0, maxsteps=8, phi0=phi0, vanalbada_pow=False)
    assert st.error == 0 and ost == 0 and n == st.steps
    assert np.array_equal(ID, IDo) and np.array_equal(phi, phio)
    assert np.array_equal(V, Vo), f"max diff {np.max(np.abs(V - Vo))}"


def test_sutherland_viscosity():
    m = M.material(M.EOS_SG, [1.4, 0.0], tlaw=(2.5, 0.0, 0.0, 0.0, 0.0, 0.0), viscosity=M.VISC_SUTHERLAND,
                   mu=0.004, su_T0=1.0, su_S=0.4, bulk=0.001, kappa=0.005)
    settings, V0, ID0 = walls_case([m], seed=9)
    print("case: Sutherland viscosity run 10 steps", flush=True)
    s = M.Solver(W.config_from(settings))
    s.upload(V0, ID0)
    st = s.run(1e300, 10, sync_every=0)
    V, _ = s.download()
    s.close()
    ost, t, n, Vo, _ = run_oracle(settings, V0, ID0, maxsteps=10, vanalbada_pow=False)
    assert st.error == 0 and ost == 0 and n == st.steps
    assert rel_linf(V, Vo) <= 1e-12


def test_mixed_precision_with_diffusion():
    settings, V0, ID0 = walls_case([visc_gas()], seed=11)
    out = {}
    for prec in (M.PREC_FP64, M.PREC_MIXED):
        s = M.Solver(W.config_from(settings, precision=prec))
        s.upload(V0, ID0)
        st = s.run(1e300, 10, sync_every=0)
        assert st.error == 0
        out[prec], _ = s.download()
        s.close()
    a, b = out[M.PREC_MIXED], out[M.PREC_FP64]
    assert np.sum(np.abs(a[..., 0] - b[..., 0])) / np.sum(np.abs(b[..., 0])) <= 1e-5


def test_rejections():
    # one cell along an axis, a batch, and a temperature law missing for heat diffusion
    settings, V0, ID0 = walls_case([visc_gas()], nz=1)
    with pytest.raises(M.M2CError):
        M.Solver(W.config_from(settings))
    jwl = M.material(M.EOS_JWL, [0.28, 3.712e11, 3.23e9, 4.15, 0.95, 1.63e-3], kappa=1.0)
    settings, V0, ID0 = walls_case([jwl])
    with pytest.raises(M.M2CError):
        M.Solver(W.config_from(settings))
    settings, V0, ID0 = walls_case([visc_gas()])
    with pytest.raises(M.M2CError):
        M.Batch([W.config_from(settings)])
