"""GPU parity of the two-material path: exact Riemann solves at interface faces, level-set
advection and reinitialisation, material-ID update and the state update of swept cells, against
the CPU oracle on the same inputs.

Two oracle runs per case: with the van Albada powers evaluated as dq*dq (what the kernels compute)
the GPU result must be bit-identical (state, IDs, level set, time, Riemann-error count). With the
reference's pow(dq, 2) (Reconstructor.h:107-108; glibc's pow differs from dq*dq in the last bit for
a small fraction of inputs) the IDs must agree exactly and the state to 1e-10 relative L-infinity
where the reference form itself is that stable: the exact Riemann solver's tolerance-driven
iteration counts (tol_main = 1e-4) can turn a last-bit change of a face state into a change of
order 1e-8 in its star state, so for some cases the oracle's two forms already differ by more
than 1e-10 and the GPU is held to that difference instead."""
import numpy as np
import pytest

import m2c_amd as M
from m2c_amd import workloads as W
from tests import cases

pytestmark = pytest.mark.gpu


def gpu_run_mp(settings, V0, ID0, phi0, maxsteps, tmax=1e300, **over):
    s = M.Solver(W.config_from(settings, **over))
    s.upload(V0, ID0)
    if phi0 is not None:
        s.upload_levelset(phi0)
    st = s.run(tmax, maxsteps, sync_every=0)
    V, ID = s.download()
    phi = s.download_levelset() if phi0 is not None else None
    return st, V, ID, phi


def compare(settings, V0, ID0, phi0, nsteps, bar=False, **over):
    print(f"case nx={settings['nx']} ny={settings['ny']} nz={settings['nz']} ts={over.get('ts', settings.get('ts'))} "
          f"steps={nsteps}", flush=True)
    sets = dict(settings)
    sets.update(over)
    st, Vg, IDg, phig = gpu_run_mp(settings, V0, ID0, phi0, nsteps, **over)
    assert st.error == 0, st.error
    ost, ot, on, Vo, o, IDo, phio = cases.run_oracle(sets, V0, ID0, maxsteps=nsteps, phi0=phi0, vanalbada_pow=False)
    assert ost == 0
    rev, clip, rerr = o.counters()
    print(f"  ID changed {int((IDo != ID0).sum())} differ {int((IDg != IDo).sum())}, "
          f"phi max diff {float(np.max(np.abs(phig - phio))):.3e}, V rel Linf {cases.rel_linf(Vg, Vo):.3e}, "
          f"riemann errors oracle {rerr} gpu {st.riemann_errors}, reinit its {st.reinit_iterations}, "
          f"unresolved {st.unresolved_cells}, t {st.t:.6e}", flush=True)
    assert st.steps == on
    assert st.t == ot
    assert np.array_equal(IDg, IDo)
    assert np.array_equal(Vg, Vo)
    assert np.array_equal(phig, phio)
    assert st.riemann_errors == rerr
    pst, pt, pn, Vp, po_, IDp, phip = cases.run_oracle(sets, V0, ID0, maxsteps=nsteps, phi0=phi0, vanalbada_pow=True)
    assert pst == 0
    err_gpu = cases.rel_linf(Vg, Vp)
    err_mode = cases.rel_linf(Vo, Vp)
    print(f"  vs pow(dq,2) oracle: GPU {err_gpu:.3e}, oracle dq*dq vs pow {err_mode:.3e}", flush=True)
    assert np.array_equal(IDg, IDp)
    assert err_gpu <= max(1e-10, err_mode)
    if bar:
        assert err_gpu <= 1e-10
    return Vg, IDg, phig


def test_shock_bubble_rk2():
    settings, V0, ID0, phi0 = W.shock_bubble(scale=0.125)
    compare(settings, V0, ID0, phi0, 40, bar=True)


def test_shock_bubble_rk3():
    settings, V0, ID0, phi0 = W.shock_bubble(scale=0.125)
    compare(settings, V0, ID0, phi0, 16, ts=M.TS_RK3, reinit_freq=4)


def test_shock_bubble_forward_euler():
    settings, V0, ID0, phi0 = W.shock_bubble(scale=0.125)
    compare(settings, V0, ID0, phi0, 10, ts=M.TS_FE, cfl=0.3, reinit_freq=3)


def test_shock_bubble_linear_extrapolation_bc():
    settings, V0, ID0, phi0 = W.shock_bubble(scale=0.125)
    compare(settings, V0, ID0, phi0, 12, ls_bc=(2, 2, 3, 3, 2, 2))


def test_shock_bubble_non_power_of_two_spacing():
    # h = 0.7: level-set differences divide by the coordinate spacing, FD3 weights come per index
    settings, V0, ID0, phi0 = W.shock_bubble(scale=0.125, h=0.7)
    compare(settings, V0, ID0, phi0, 12, reinit_freq=3)


def test_shock_bubble_power_of_two_spacing():
    # h = 0.5: divisions by the spacing become exact multiplications by 2
    settings, V0, ID0, phi0 = W.shock_bubble(scale=0.125, h=0.5)
    compare(settings, V0, ID0, phi0, 12, reinit_freq=3)


def test_two_gas_interface_without_level_set():
    # a static two-material Riemann problem: the interface faces use the exact solver, IDs fixed
    n = 64
    settings = dict(nx=n, ny=2, nz=2, box=(0.0, 1.0, 0.0, 2.0 / n, 0.0, 2.0 / n),
                    bc=[M.BC_INLET, M.BC_OUTLET, M.BC_SYMMETRY, M.BC_SYMMETRY, M.BC_SYMMETRY, M.BC_SYMMETRY],
                    mats=[M.material(M.EOS_SG, [1.4, 0.0]), M.material(M.EOS_SG, [1.67, 0.0])],
                    inlet=(1.0, 0.0, 0.0, 0.0, 1.0), outlet=(0.125, 0.0, 0.0, 0.0, 0.1), ts=M.TS_RK2, cfl=0.4)
    x = (np.arange(n) + 0.5) / n
    V0 = np.zeros((2, 2, n, 5))
    V0[..., 0] = np.where(x < 0.5, 1.0, 0.125)
    V0[..., 4] = np.where(x < 0.5, 1.0, 0.1)
    ID0 = np.broadcast_to(np.where(x < 0.5, 0, 1), (2, 2, n)).astype(np.int32).copy()
    ost, ot, on, Vo, o = cases.run_oracle(settings, V0, ID0, maxsteps=20, vanalbada_pow=False)
    assert ost == 0
    s = M.Solver(W.config_from(settings))
    s.upload(V0, ID0)
    st = s.run(1e300, 20, sync_every=0)
    Vg, IDg = s.download()
    assert st.error == 0
    assert np.array_equal(IDg, ID0)
    assert np.array_equal(Vg, Vo), cases.rel_linf(Vg, Vo)


def test_water_air_riemann_residual():
    # one residual evaluation across a water/air interface in all three directions
    settings, V0, ID0, phi0 = W.shock_bubble(scale=0.125)
    o_cfg = cases.oracle_config(settings, vanalbada_pow=False)
    from oracle import pyoracle as po
    o = po.Oracle(o_cfg)
    V, ID = o.ghosted(V0, ID0)
    st, Ro = o.residual(V, ID)
    s = M.Solver(W.config_from(settings))
    s.upload(V0, ID0)
    Rg = s.residual()
    assert np.array_equal(Rg, Ro[1:-1, 1:-1, 1:-1])
