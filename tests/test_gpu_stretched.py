"""Stretched meshes (Mesh.ControlPointX/Y/Z) on the GPU against the oracle, bit for bit: the stage
kernels' per-cell limiter coefficients, face areas, volumes and CFL widths over every flux and
limiter and the boundary conditions; the two-material shock-bubble on a mesh refined around the
bubble (interface Riemann solves, level-set advection on the stretched coordinates, reinitialisation
with per-cell steps); the mixed-precision form within its tolerance of the FP64 run; and the
configurations a stretched mesh does not run, refused with their error codes."""
import numpy as np
import pytest

import m2c_amd as M
from m2c_amd import workloads as W
from tests.cases import rel_linf, run_oracle, smooth_state

pytestmark = pytest.mark.gpu

PTS = ([(0.0, 0.03), (0.45, 0.015), (1.0, 0.04)], [(0.0, 0.05), (0.5, 0.02)], [(0.1, 0.04), (0.3, 0.02)])
BOX = (0.0, 1.0, 0.0, 0.5, 0.0, 0.3)


def stretched_settings(**kw):
    n = [len(M.mesh_1d(BOX[2 * d], BOX[2 * d + 1], points=PTS[d])[0]) for d in range(3)]
    s = dict(nx=n[0], ny=n[1], nz=n[2], box=BOX,
             bc=[M.BC_INLET, M.BC_OUTLET, M.BC_SLIPWALL, M.BC_STICKWALL, M.BC_SYMMETRY, M.BC_SLIPWALL],
             mats=[M.material(M.EOS_SG, [1.4, 0.0])], inlet=(1.2, 0.3, 0.0, 0.0, 1.1),
             outlet=(0.9, 0.1, 0.0, 0.0, 0.95), flux=M.FLUX_HLLC, limiter=M.LIM_VANALBADA, ts=M.TS_RK2, cfl=0.5,
             control_points=PTS)
    s.update(kw)
    return s


def gpu_run(settings, V0, ID0, maxsteps, phi0=None, **over):
    s = M.Solver(W.config_from(settings, **over))
    s.upload(V0, ID0)
    if phi0 is not None:
        s.upload_levelset(phi0)
    st = s.run(1e300, maxsteps, sync_every=0)
    V, ID = s.download()
    phi = s.download_levelset() if phi0 is not None else None
    s.close()
    return st, V, ID, phi


@pytest.mark.parametrize("flux", [M.FLUX_HLLC, M.FLUX_ROE, M.FLUX_LLF, M.FLUX_GODUNOV])
@pytest.mark.parametrize("limiter", [M.LIM_NONE, M.LIM_GENMINMOD, M.LIM_VANALBADA])
def test_stretched_single_material_bitwise(flux, limiter):
    settings = stretched_settings(flux=flux, limiter=limiter, ts=M.TS_RK3 if flux == M.FLUX_ROE else M.TS_RK2)
    nx, ny, nz = settings["nx"], settings["ny"], settings["nz"]
    print(f"case: stretched {nx}x{ny}x{nz} flux={flux} limiter={limiter}", flush=True)
    V0 = smooth_state(nx, ny, nz, seed=3 * flux + limiter)
    ID0 = np.zeros((nz, ny, nx), dtype=np.int32)
    st, V, _, _ = gpu_run(settings, V0, ID0, 6)
    assert st.error == 0 and st.steps == 6
    ost, t, n, Vo, _ = run_oracle(settings, V0, ID0, maxsteps=6, vanalbada_pow=False)
    assert ost == 0 and t == st.t
    assert np.array_equal(V, Vo), f"max diff {np.max(np.abs(V - Vo))}"


def test_stretched_mixed_precision_within_tolerance():
    settings = stretched_settings()
    nx, ny, nz = settings["nx"], settings["ny"], settings["nz"]
    V0 = smooth_state(nx, ny, nz, seed=9)
    ID0 = np.zeros((nz, ny, nx), dtype=np.int32)
    print(f"case: stretched {nx}x{ny}x{nz} mixed precision", flush=True)
    st64, V64, _, _ = gpu_run(settings, V0, ID0, 6)
    st32, V32, _, _ = gpu_run(settings, V0, ID0, 6, precision=M.PREC_MIXED)
    assert st64.error == 0 and st32.error == 0
    # FP32 reconstruction and fluxes, FP64 state: the mixed form's bar on smooth flows
    assert rel_linf(V32, V64) <= 1e-5


def test_stretched_shock_bubble_bitwise():
    settings, V0, ID0, phi0 = W.shock_bubble_stretched(scale=0.125, reinit_freq=3)
    print(f"case: stretched shock-bubble {settings['nx']}x{settings['ny']}x{settings['nz']}, 7 steps", flush=True)
    st, V, ID, phi = gpu_run(settings, V0, ID0, 7, phi0=phi0)
    assert st.error == 0 and st.steps == 7
    assert st.reinit_iterations > 0
    ost, t, n, Vo, _, IDo, phio = run_oracle(settings, V0, ID0, maxsteps=7, phi0=phi0, vanalbada_pow=False)
    assert ost == 0 and t == st.t
    assert np.array_equal(ID, IDo)
    assert np.array_equal(phi, phio), f"max diff {np.max(np.abs(phi - phio))}"
    assert np.array_equal(V, Vo), f"max diff {np.max(np.abs(V - Vo))}"


def test_stretched_shock_bubble_levelset_normals_bitwise():
    # the exact Riemann problems along the level-set normal, whose face gradient interpolates to the
    # face position of the stretched axis
    settings, V0, ID0, phi0 = W.shock_bubble_stretched(scale=0.125, reinit_freq=3)
    settings["mp_normal"] = 0
    print(f"case: stretched shock-bubble {settings['nx']}x{settings['ny']}x{settings['nz']}, level-set normals",
          flush=True)
    st, V, ID, phi = gpu_run(settings, V0, ID0, 4, phi0=phi0)
    assert st.error == 0
    ost, t, n, Vo, _, IDo, phio = run_oracle(settings, V0, ID0, maxsteps=4, phi0=phi0, vanalbada_pow=False)
    assert ost == 0 and t == st.t
    assert np.array_equal(ID, IDo) and np.array_equal(phi, phio)
    assert np.array_equal(V, Vo), f"max diff {np.max(np.abs(V - Vo))}"


def test_stretched_unsupported_configurations():
    s = stretched_settings()
    # a cell count that is not the generator's
    with pytest.raises(M.M2CError) as e:
        M.Solver(W.config_from(s, nx=s["nx"] + 1))
    assert e.value.code == M.ERR_ARG
    # periodic axes
    with pytest.raises(M.M2CError) as e:
        M.Solver(W.config_from(s, bc=[M.BC_PERIODIC] * 6))
    assert e.value.code == M.ERR_UNSUPPORTED
    # viscous terms
    with pytest.raises(M.M2CError) as e:
        M.Solver(W.config_from(s, mats=[M.material(M.EOS_SG, [1.4, 0.0], viscosity=M.VISC_CONSTANT, mu=0.01)]))
    assert e.value.code == M.ERR_UNSUPPORTED
    # batches
    with pytest.raises(M.M2CError) as e:
        M.Batch([W.config_from(s, bc=[M.BC_SLIPWALL] * 6)])
    assert e.value.code == M.ERR_UNSUPPORTED
