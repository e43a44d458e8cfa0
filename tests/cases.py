"""Shared test cases: one settings dict drives both the HIP library and the CPU oracle."""
import numpy as np

import m2c_amd as M
from m2c_amd import workloads as W
from oracle import pyoracle as po


def oracle_eos(m):
    e = po.Eos()
    e.type = m.eos
    e.rhomin, e.pmin, e.rhomax, e.pmax, e.failsafe_density = m.rhomin, m.pmin, m.rhomax, m.pmax, m.failsafe_density
    p = list(m.p)
    if m.eos == M.EOS_SG:
        e.gam, e.pstiff = p[0], p[1]
    elif m.eos == M.EOS_NASG:
        e.gam, e.pc, e.b, e.q = p[0], p[1], p[2], p[3]
    elif m.eos == M.EOS_MG:
        e.rho0, e.c0, e.Gamma0, e.s, e.e0 = p[0], p[1], p[2], p[3], p[4]
    elif m.eos == M.EOS_JWL:
        e.omega, e.A1, e.A2, e.R1, e.R2, e.rho0 = p[0], p[1], p[2], p[3], p[4], p[5]
    elif m.eos == M.EOS_DUMMY:
        e.d_p0, e.d_rho0, e.d_e0 = p[0], p[1], p[2]
    e.t_cv, e.t_T0, e.t_e0, e.t_rho0, e.t_cp, e.t_h0 = list(m.tlaw)
    e.visc, e.mu, e.bulk, e.su_T0, e.su_S, e.kappa = m.viscosity, m.mu, m.bulk, m.su_T0, m.su_S, m.kappa
    return e


def oracle_config(settings, vanalbada_pow=True, nthreads=0, **over):
    s = dict(settings)
    s.update(over)
    mats = [oracle_eos(m) for m in s["mats"]]
    return po.make_config(s["nx"], s["ny"], s["nz"], s["box"], s["bc"], mats,
                          inlet=s.get("inlet", (1, 0, 0, 0, 1)), outlet=s.get("outlet", (1, 0, 0, 0, 1)),
                          flux=s.get("flux", M.FLUX_HLLC), limiter=s.get("limiter", M.LIM_VANALBADA),
                          ts=s.get("ts", M.TS_RK2), cfl=s.get("cfl", 0.5), rec_linear=s.get("rec_linear", True),
                          vanalbada_pow=vanalbada_pow, inactive_id=s.get("inactive_id", -1), nthreads=nthreads,
                          roe_delta=s.get("roe_delta", 0.2), gminmod_alpha=s.get("gminmod_alpha", 1.2),
                          slope_zero=s.get("slope_zero", False), mp_flux=s.get("mp_flux", 0),
                          mp_recon=s.get("mp_recon", 0), ls_matid=s.get("ls_matid"),
                          ls_bc=s.get("ls_bc", (1, 1, 1, 1, 1, 1)), reinit_freq=s.get("reinit_freq", -1),
                          reinit_maxits=s.get("reinit_maxits", 30), reinit_cfl=s.get("reinit_cfl", 0.8),
                          reinit_tol=s.get("reinit_tol", 2.0e-4), mp_normal=s.get("mp_normal", 1),
                          conrec_depth=s.get("conrec_depth", 0.0), control_points=s.get("control_points"))


def run_oracle(settings, V0, ID0, tmax=1e300, maxsteps=10, vanalbada_pow=True, phi0=None, **over):
    """Run the oracle; with phi0 the level set is advanced too and (.., ID, phi) are appended."""
    o = po.Oracle(oracle_config(settings, vanalbada_pow=vanalbada_pow, **over))
    V, ID = o.ghosted(V0, ID0)
    if phi0 is None:
        st, t, n = o.run(V, ID, tmax, maxsteps)
        return st, t, n, V[1:-1, 1:-1, 1:-1].copy(), o
    phi = o.ghosted_phi(phi0)
    st, t, n = o.run(V, ID, tmax, maxsteps, phi=phi)
    return (st, t, n, V[1:-1, 1:-1, 1:-1].copy(), o, ID[1:-1, 1:-1, 1:-1].copy(), phi[1:-1, 1:-1, 1:-1].copy())


def smooth_state(nx, ny, nz, seed=1, amp=0.2):
    """A smooth random state with rho, p > 0 and velocities of order 0.3."""
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(nz) / max(nz, 1), np.arange(ny) / max(ny, 1), np.arange(nx) / max(nx, 1),
                          indexing="ij")
    V = np.zeros((nz, ny, nx, 5))
    base = [1.0, 0.0, 0.0, 0.0, 1.0]
    for c in range(5):
        f = np.full_like(x, base[c])
        for _ in range(3):
            k = rng.integers(1, 4, size=3)
            ph = rng.uniform(0, 2 * np.pi)
            f += amp * np.sin(2 * np.pi * (k[0] * x + k[1] * y + k[2] * z) + ph)
        V[..., c] = f
    return V


def rel_linf(a, b):
    """max |a-b| / max|b| per conserved/primitive field, returned as the worst field."""
    out = 0.0
    for c in range(a.shape[-1]):
        den = max(np.max(np.abs(b[..., c])), 1e-300)
        out = max(out, np.max(np.abs(a[..., c] - b[..., c])) / den)
    return out
