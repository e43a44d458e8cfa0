"""Pins of the oracle's finite-volume restatement (oracle/eos.h, flux.h, fv.c) against closed forms
and invariants: equation-of-state round trips and the stiffened-gas sound speed
(VarFcnSG.h:64-68, VarFcnBase.h:169-219); flux consistency F(V, V) = physical flux and mirror
symmetry for HLLC / generalised Roe / LLF (FluxFcnHLLC.h, FluxFcnGenRoe.h, FluxFcnLLF.h,
FluxFcnBase.h:83-123); the MUSCL slope of van Albada with k = 2 on a uniform mesh
(Reconstructor.h:95-110, Reconstructor.cpp:103-121); discrete conservation of mass, momentum and
energy with periodic boundaries; Sod's problem against the exact solution (Toro, table 4.3)."""
import math

import numpy as np
import pytest

from oracle import pyoracle as po

FLUXES = [po.FLUX_HLLC, po.FLUX_ROE, po.FLUX_LLF]


def oracle(mats, flux=po.FLUX_HLLC, **kw):
    cfg = po.make_config(4, 4, 4, (0, 1, 0, 1, 0, 1), [po.BC_SYMMETRY] * 6, mats, flux=flux, **kw)
    return po.Oracle(cfg)


def materials():
    return {
        "sg": po.sg(1.4, 0.0),
        "water": po.sg(6.12, 3.43e8),
        "nasg": po.eos(po.EOS_NASG, gam=1.19, pc=6.217e8, b=6.7212e-4, q=-1177788.0),
        "mg": po.eos(po.EOS_MG, rho0=2785.0, c0=5328.0, Gamma0=2.0, s=1.338, e0=0.0),
        "jwl": po.eos(po.EOS_JWL, omega=0.28, A1=3.712e11, A2=3.23e9, R1=4.15, R2=0.95, rho0=1630.0),
    }


STATES = {
    "sg": np.array([1.3, 0.4, -0.2, 0.1, 2.5]),
    "water": np.array([1012.0, 12.0, -3.0, 1.0, 2e6]),
    "nasg": np.array([1005.0, 2.0, 0.5, -1.0, 5e5]),
    "mg": np.array([2900.0, 50.0, -20.0, 10.0, 2e9]),
    "jwl": np.array([1700.0, 100.0, 5.0, -5.0, 5e9]),
}


@pytest.mark.parametrize("name", list(STATES))
def test_eos_round_trip(name):
    o = oracle([materials()[name]])
    V = STATES[name]
    U = o.p2c(V)
    assert U[0] == V[0] and U[1] == V[0] * V[1]
    V2 = o.c2p(U)
    assert np.allclose(V2, V, rtol=1e-12, atol=0)
    assert o.c2(V[0], V[4]) > 0


def test_stiffened_gas_sound_speed_closed_form():
    o = oracle([po.sg(6.12, 3.43e8)])
    for rho, p in [(1000.0, 1e5), (1200.0, 1e9), (990.0, -1e8)]:
        assert abs(o.c2(rho, p) - 6.12 * (p + 3.43e8) / rho) <= 1e-12 * 6.12 * (p + 3.43e8) / rho


def physical_flux(V, gam, pinf, d):
    rho, u, p = V[0], V[1:4], V[4]
    e = (p + gam * pinf) / ((gam - 1.0) * rho)
    E = rho * (e + 0.5 * u @ u)
    un = u[d]
    F = np.zeros(5)
    F[0] = rho * un
    F[1:4] = rho * un * u
    F[1 + d] += p
    F[4] = (E + p) * un
    return F


@pytest.mark.parametrize("flux", FLUXES)
@pytest.mark.parametrize("d", [0, 1, 2])
def test_flux_consistency(flux, d):
    for mat, V in [((1.4, 0.0), STATES["sg"]), ((6.12, 3.43e8), STATES["water"])]:
        o = oracle([po.sg(*mat)], flux=flux)
        F = o.flux(d, V, V)
        Fx = physical_flux(V, mat[0], mat[1], d)
        assert np.allclose(F, Fx, rtol=1e-12, atol=1e-12 * np.abs(Fx).max())


@pytest.mark.parametrize("flux", FLUXES)
def test_flux_mirror_symmetry(flux):
    # the generalised Roe flux starts its entropy-fix tolerance from the signed lam[0]
    # (FluxFcnGenRoe.h:150-156), which breaks mirror symmetry; without the fix it is symmetric
    o = oracle([po.sg(1.4, 0.0)], flux=flux, roe_delta=0.0)
    rng = np.random.default_rng(5)
    for _ in range(50):
        Vm = np.array([rng.uniform(0.2, 2), *rng.uniform(-1, 1, 3), rng.uniform(0.2, 3)])
        Vp = np.array([rng.uniform(0.2, 2), *rng.uniform(-1, 1, 3), rng.uniform(0.2, 3)])
        M = np.array([1.0, -1.0, 1.0, 1.0, 1.0])
        F = o.flux(0, Vm, Vp)
        G = o.flux(0, Vp * M, Vm * M)
        # mirrored problem: mass, tangential momentum and energy fluxes change sign, normal
        # momentum flux does not
        expect = -F * np.array([1.0, -1.0, 1.0, 1.0, 1.0])
        assert np.allclose(G, expect, rtol=1e-10, atol=1e-12)


def test_roe_entropy_fix_tolerance_quirk():
    # a left-running sonic state: |lam[0]| is the largest wave speed; the reference's tolerance
    # ignores it, so the mirrored problem (where it becomes lam[4]) is fixed differently
    o = oracle([po.sg(1.4, 0.0)], flux=po.FLUX_ROE, roe_delta=0.2)
    M = np.array([1.0, -1.0, 1.0, 1.0, 1.0])
    Vm = np.array([1.0, -0.9, 0.0, 0.0, 1.0])
    Vp = np.array([0.9, -0.8, 0.0, 0.0, 0.9])
    F = o.flux(0, Vm, Vp)
    G = o.flux(0, Vp * M, Vm * M)
    assert not np.allclose(G, -F * np.array([1.0, -1.0, 1.0, 1.0, 1.0]), rtol=1e-12, atol=0)


def test_van_albada_uniform_slope_closed_form():
    # uniform mesh: A = B = 1, k = 2, slope = dq0 dq1 (dq0 + dq1) / (dq0^2 + dq1^2). With u = 0 and
    # constant p, the LLF mass flux of a face is -a/2 (rho_R - rho_L), a = max sound speed of the
    # two face states, so the mass residual of each cell is fixed by the reconstructed densities.
    cfg = po.make_config(3, 1, 1, (0, 3, 0, 1, 0, 1), [po.BC_SYMMETRY] * 6, [po.sg(1.4, 0.0)],
                         flux=po.FLUX_LLF, limiter=po.LIM_VANALBADA, vanalbada_pow=False)
    o = po.Oracle(cfg)
    rho = np.array([1.0, 1.5, 3.0])
    V0 = np.zeros((1, 1, 3, 5))
    V0[..., 0] = rho
    V0[..., 4] = 1.0
    V, ID = o.ghosted(V0)
    st, R = o.residual(V, ID)
    assert st == 0
    dq0, dq1 = 0.5, 1.5
    sig1 = dq0 * dq1 * (dq0 + dq1) / (dq0 * dq0 + dq1 * dq1)
    # symmetry ghosts mirror the end cells, so their outer one-sided differences vanish: slope 0
    faces = [(1.0, 1.0), (1.0, 1.5 - 0.5 * sig1), (1.5 + 0.5 * sig1, 3.0), (3.0, 3.0)]

    def mass_flux(rl, rr):
        a = max(math.sqrt(1.4 / rl), math.sqrt(1.4 / rr))
        return -0.5 * a * (rr - rl)

    F = [mass_flux(*f) for f in faces]
    for i in range(3):
        expect = -(F[i + 1] - F[i])
        assert abs(R[1, 1, 1 + i, 0] - expect) <= 1e-14 * max(1.0, abs(expect))


def test_conservation_periodic():
    n = (12, 10, 8)
    cfg = po.make_config(*n, (0, 1.2, 0, 1.0, 0, 0.8), [po.BC_PERIODIC] * 6, [po.sg(1.4, 0.0)], ts=po.TS_RK3,
                         cfl=0.4)
    o = po.Oracle(cfg)
    rng = np.random.default_rng(11)
    V0 = np.zeros(n[::-1] + (5,))
    V0[..., 0] = 1.0 + 0.2 * rng.uniform(size=n[::-1])
    V0[..., 1:4] = 0.3 * rng.uniform(-1, 1, n[::-1] + (3,))
    V0[..., 4] = 1.0 + 0.2 * rng.uniform(size=n[::-1])
    V, ID = o.ghosted(V0)

    def totals(Vg):
        tot = np.zeros(5)
        for U in (o.p2c(v) for v in Vg[1:-1, 1:-1, 1:-1].reshape(-1, 5)):
            tot += U
        return tot

    t0 = totals(V)
    st, t, ns = o.run(V, ID, 1e300, 6)
    assert st == 0 and ns == 6
    t1 = totals(V)
    scale = np.array([t0[0], np.abs(t0[1:4]).max() + t0[0], np.abs(t0[1:4]).max() + t0[0], np.abs(t0[1:4]).max() + t0[0], t0[4]])
    assert np.all(np.abs(t1 - t0) <= 1e-12 * scale)


def sod_exact(x, t, x0=0.5, g=1.4):
    rl, ul, pl, rr, ur, pr = 1.0, 0.0, 1.0, 0.125, 0.0, 0.1
    ps, us = 0.30313017805064697, 0.92745262004895056
    cl, cr = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    rsl = rl * (ps / pl) ** (1 / g)
    rsr = rr * ((ps / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pr + 1))
    csl = cl * (ps / pl) ** ((g - 1) / (2 * g))
    sh = ur + cr * math.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
    out = np.empty_like(x)
    for i, xi in enumerate(x):
        s = (xi - x0) / t
        if s < ul - cl:
            out[i] = rl
        elif s < us - csl:
            u = 2 / (g + 1) * (cl + (g - 1) / 2 * ul + s)
            c = 2 / (g + 1) * (cl + (g - 1) / 2 * (ul - s))
            out[i] = rl * (c / cl) ** (2 / (g - 1))
        elif s < us:
            out[i] = rsl
        elif s < sh:
            out[i] = rsr
        else:
            out[i] = rr
    return out


def test_sod_converges_to_exact_solution():
    n = 400
    cfg = po.make_config(n, 1, 1, (0, 1, 0, 1 / n, 0, 1 / n),
                         [po.BC_INLET, po.BC_OUTLET] + [po.BC_SYMMETRY] * 4, [po.sg(1.4, 0.0)],
                         inlet=(1, 0, 0, 0, 1), outlet=(0.125, 0, 0, 0, 0.1), ts=po.TS_RK2, cfl=0.5)
    o = po.Oracle(cfg)
    x = (np.arange(n) + 0.5) / n
    V0 = np.zeros((1, 1, n, 5))
    V0[..., 0] = np.where(x < 0.5, 1.0, 0.125)
    V0[..., 4] = np.where(x < 0.5, 1.0, 0.1)
    V, ID = o.ghosted(V0)
    st, t, ns = o.run(V, ID, 0.2, 100000)
    assert st == 0 and abs(t - 0.2) < 1e-15
    rho = V[1, 1, 1:-1, 0]
    ex = sod_exact(x, t)
    assert np.mean(np.abs(rho - ex)) < 4e-3
    # plateau between contact and shock
    m = (x > 0.72) & (x < 0.82)
    assert np.max(np.abs(rho[m] - 0.26557371170530714)) < 2e-3
