"""Pins of the oracle's HLLC and generalised Roe fluxes (oracle/flux.h) by properties the
mathematics fixes:
* HLLC resolves isolated contact and shear discontinuities exactly (Toro, Riemann Solvers and
  Numerical Methods for Fluid Dynamics, section 10.4): with equal normal velocity and pressure on
  both sides the flux is the physical flux of the upwind state, whatever the density and tangential
  velocity jumps (FluxFcnHLLC.h:159-245);
* supersonic pairs: HLLC's wave-speed bounds (FluxFcnHLLC.h:58-154) put both waves on one side, so
  the flux is the upwind physical flux;
* the generalised Roe flux (FluxFcnGenRoe.h:64-221, after Hu, Adams & Iaccarino 2009): for a
  stiffened gas (ideal gas included) dp/drho = (gamma-1) e and Gamma = gamma-1, and the identity
  d(rho e) = rho_hat de + e_hat drho (rho_hat = sqrt(rho_m rho_p), e_hat the sqrt(rho)-weighted mean)
  makes its averaged derivatives collapse to dp/drho_hat = (gamma-1) e_hat and Gamma_hat = gamma-1,
  whatever the weights w_rho, w_e. Its sound speed is then
  c_hat^2 = (gamma-1) (e_hat + mean(p/rho) + (du_n / (sqrt(rho_m) + sqrt(rho_p)))^2 / 2),
  and the flux must equal a Roe-form flux written from the Euler Jacobian's eigenstructure with
  that c_hat, the sqrt(rho)-weighted velocity and enthalpy, and the textbook wave strengths
  (Toro section 11.2). Note that the reference's (p/rho)_hat term lacks the factor rho_hat that
  Roe's own average has ((gamma-1)(H_hat - |u_hat|^2/2)), so the flux is not Roe's for an ideal gas
  and does not satisfy property U; the pin below holds the reference's form.
The physical flux and the Roe-form flux below are written from the Euler equations, not from the
oracle.
"""
import numpy as np
import pytest

from oracle import pyoracle as po


def oracle(mats, flux, **kw):
    cfg = po.make_config(4, 4, 4, (0, 1, 0, 1, 0, 1), [po.BC_SYMMETRY] * 6, mats, flux=flux, **kw)
    return po.Oracle(cfg)


def euler_flux(V, gam, pinf, d):
    rho, u, p = V[0], np.asarray(V[1:4], dtype=float), V[4]
    e = (p + gam * pinf) / ((gam - 1.0) * rho)
    E = rho * (e + 0.5 * u @ u)
    F = np.zeros(5)
    F[0] = rho * u[d]
    F[1:4] = rho * u[d] * u
    F[1 + d] += p
    F[4] = (E + p) * u[d]
    return F


def cons(V, gam, pinf):
    rho, u, p = V[0], np.asarray(V[1:4], dtype=float), V[4]
    e = (p + gam * pinf) / ((gam - 1.0) * rho)
    return np.array([rho, *(rho * u), rho * (e + 0.5 * u @ u)])


def close(F, G, rtol):
    scale = max(np.abs(G).max(), 1e-300)
    return np.abs(F - G).max() <= rtol * scale


MATS = [(1.4, 0.0), (6.12, 3.43e8)]


@pytest.mark.parametrize("gam,pinf", MATS)
@pytest.mark.parametrize("d", [0, 1, 2])
def test_hllc_stationary_contact_is_exact(gam, pinf, d):
    o = oracle([po.sg(gam, pinf)], po.FLUX_HLLC)
    rng = np.random.default_rng(11 + d)
    for _ in range(40):
        p = rng.uniform(0.5, 3.0) * (1.0 if pinf == 0 else 1e6)
        Vm = np.zeros(5)
        Vp = np.zeros(5)
        Vm[0], Vp[0] = rng.uniform(0.1, 5.0), rng.uniform(0.1, 5.0)
        # shear: tangential velocities jump, the normal one is zero on both sides
        Vm[1:4] = rng.uniform(-1, 1, 3)
        Vp[1:4] = rng.uniform(-1, 1, 3)
        Vm[1 + d] = Vp[1 + d] = 0.0
        Vm[4] = Vp[4] = p
        F = o.flux(d, Vm, Vp)
        expect = np.zeros(5)
        expect[1 + d] = p
        # rounding scale: wave speed times the conserved state (the star-state jump S (U* - U) is
        # zero only in exact arithmetic)
        c = max(np.sqrt(gam * (p + pinf) / Vm[0]), np.sqrt(gam * (p + pinf) / Vp[0]))
        scale = p + c * max(np.abs(cons(Vm, gam, pinf)).max(), np.abs(cons(Vp, gam, pinf)).max())
        assert np.abs(F - expect).max() <= 1e-13 * scale, (F, expect)


@pytest.mark.parametrize("gam,pinf", MATS)
def test_hllc_moving_contact_takes_upwind_flux(gam, pinf):
    o = oracle([po.sg(gam, pinf)], po.FLUX_HLLC)
    rng = np.random.default_rng(3)
    for _ in range(60):
        d = int(rng.integers(0, 3))
        un = rng.uniform(-0.8, 0.8)
        p = rng.uniform(0.5, 3.0) * (1.0 if pinf == 0 else 1e6)
        Vm = np.array([rng.uniform(0.1, 5.0), *rng.uniform(-1, 1, 3), p])
        Vp = np.array([rng.uniform(0.1, 5.0), *rng.uniform(-1, 1, 3), p])
        Vm[1 + d] = Vp[1 + d] = un
        up = Vm if un > 0 else Vp
        F = o.flux(d, Vm, Vp)
        G = euler_flux(up, gam, pinf, d)
        c = max(np.sqrt(gam * (p + pinf) / Vm[0]), np.sqrt(gam * (p + pinf) / Vp[0]))
        scale = np.abs(G).max() + c * max(np.abs(cons(Vm, gam, pinf)).max(), np.abs(cons(Vp, gam, pinf)).max())
        assert np.abs(F - G).max() <= 1e-13 * scale, (d, un, F, G)


@pytest.mark.parametrize("gam,pinf", MATS)
def test_hllc_supersonic_pair_takes_upwind_flux(gam, pinf):
    flux = po.FLUX_HLLC
    o = oracle([po.sg(gam, pinf)], flux)
    rng = np.random.default_rng(7)
    for _ in range(60):
        d = int(rng.integers(0, 3))
        Vm = np.array([rng.uniform(0.5, 2.0), *rng.uniform(-0.3, 0.3, 3), rng.uniform(0.5, 2.0)])
        Vp = np.array([rng.uniform(0.5, 2.0), *rng.uniform(-0.3, 0.3, 3), rng.uniform(0.5, 2.0)])
        if pinf:
            Vm[4] *= 1e6
            Vp[4] *= 1e6
            Vm[0] *= 1e3
            Vp[0] *= 1e3
        cm = np.sqrt(gam * (Vm[4] + pinf) / Vm[0])
        cp = np.sqrt(gam * (Vp[4] + pinf) / Vp[0])
        s = 1.0 if rng.uniform() < 0.5 else -1.0
        mach = rng.uniform(1.6, 4.0)
        Vm[1 + d] = s * mach * max(cm, cp)
        Vp[1 + d] = s * (mach + rng.uniform(0.0, 0.5)) * max(cm, cp)
        up = Vm if s > 0 else Vp
        F = o.flux(d, Vm, Vp)
        assert close(F, euler_flux(up, gam, pinf, d), 1e-10), (flux, d, F, euler_flux(up, gam, pinf, d))


def roe_form_flux(Vm, Vp, gam, pinf, d):
    """Roe-form flux (Toro section 11.2 eigenstructure and wave strengths), no entropy fix, with the
    averaged sound speed the stiffened-gas algebra of the generalised average gives."""
    rl, rr = np.sqrt(Vm[0]), np.sqrt(Vp[0])
    w = lambda a, b: (rl * a + rr * b) / (rl + rr)
    em = (Vm[4] + gam * pinf) / ((gam - 1.0) * Vm[0])
    ep = (Vp[4] + gam * pinf) / ((gam - 1.0) * Vp[0])
    Hm = em + Vm[4] / Vm[0] + 0.5 * Vm[1:4] @ Vm[1:4]
    Hp = ep + Vp[4] / Vp[0] + 0.5 * Vp[1:4] @ Vp[1:4]
    u = w(Vm[1:4], Vp[1:4])
    H = w(Hm, Hp)
    rho_hat = np.sqrt(Vm[0] * Vp[0])
    diff = (Vp[1 + d] - Vm[1 + d]) / (rl + rr)
    c2 = (gam - 1.0) * (w(em, ep) + w(Vm[4] / Vm[0], Vp[4] / Vp[0]) + 0.5 * diff * diff)
    c = np.sqrt(c2)
    un = u[d]
    n = np.eye(3)[d]
    t = [a for a in range(3) if a != d]
    dp, drho = Vp[4] - Vm[4], Vp[0] - Vm[0]
    dv = Vp[1:4] - Vm[1:4]
    waves = [(un - c, (dp - rho_hat * c * dv[d]) / (2 * c2), np.array([1.0, *(u - c * n), H - un * c])),
             (un, drho - dp / c2, np.array([1.0, *u, 0.5 * u @ u])),
             (un, rho_hat * dv[t[0]], np.array([0.0, *np.eye(3)[t[0]], u[t[0]]])),
             (un, rho_hat * dv[t[1]], np.array([0.0, *np.eye(3)[t[1]], u[t[1]]])),
             (un + c, (dp + rho_hat * c * dv[d]) / (2 * c2), np.array([1.0, *(u + c * n), H + un * c]))]
    F = 0.5 * (euler_flux(Vm, gam, pinf, d) + euler_flux(Vp, gam, pinf, d))
    for lam, a, r in waves:
        F -= 0.5 * abs(lam) * a * r
    return F


@pytest.mark.parametrize("gam,pinf", MATS)
@pytest.mark.parametrize("d", [0, 1, 2])
def test_generalised_roe_stiffened_gas_form(gam, pinf, d):
    o = oracle([po.sg(gam, pinf)], po.FLUX_ROE, roe_delta=0.0)
    rng = np.random.default_rng(21 + d)
    for _ in range(60):
        Vm = np.array([rng.uniform(0.2, 3.0), *rng.uniform(-0.6, 0.6, 3), rng.uniform(0.2, 3.0)])
        Vp = np.array([rng.uniform(0.2, 3.0), *rng.uniform(-0.6, 0.6, 3), rng.uniform(0.2, 3.0)])
        if pinf:
            Vm[0] *= 1e3
            Vp[0] *= 1e3
            Vm[4] *= 1e6
            Vp[4] *= 1e6
            Vm[1:4] *= 30.0
            Vp[1:4] *= 30.0
        F = o.flux(d, Vm, Vp)
        G = roe_form_flux(Vm, Vp, gam, pinf, d)
        scale = np.abs(G).max() + np.sqrt(gam * (max(Vm[4], Vp[4]) + pinf) / min(Vm[0], Vp[0])) * max(
            np.abs(cons(Vm, gam, pinf)).max(), np.abs(cons(Vp, gam, pinf)).max())
        assert np.abs(F - G).max() <= 1e-12 * scale, (d, F, G)


def test_generalised_roe_is_not_roes_average():
    """The reference's (p/rho)_hat has no rho_hat factor: with a normal-velocity jump and
    rho_m rho_p != 1 the flux differs from Roe's ideal-gas flux by far more than rounding."""
    o = oracle([po.sg(1.4, 0.0)], po.FLUX_ROE, roe_delta=0.0)
    Vm = np.array([4.0, 0.5, 0.0, 0.0, 2.0])
    Vp = np.array([2.0, -0.5, 0.0, 0.0, 1.0])
    F = o.flux(0, Vm, Vp)
    assert np.allclose(F, roe_form_flux(Vm, Vp, 1.4, 0.0, 0), rtol=1e-12, atol=1e-12)
    # Roe's own average: c^2 = (gamma-1)(H_hat - |u_hat|^2/2) includes rho_hat * diff^2 / 2
    rl, rr = 2.0, np.sqrt(2.0)
    assert abs(np.sqrt(8.0) - 1.0) > 1.0   # rho_hat = sqrt(8), the missing factor
    Hm = 3.5 * 2.0 / 4.0 + 0.125
    Hp = 3.5 * 1.0 / 2.0 + 0.125
    H = (rl * Hm + rr * Hp) / (rl + rr)
    u = (rl * 0.5 - rr * 0.5) / (rl + rr)
    c2_roe = 0.4 * (H - 0.5 * u * u)
    em, ep = 2.0 / (0.4 * 4.0), 1.0 / (0.4 * 2.0)
    c2_ref = 0.4 * ((rl * em + rr * ep) / (rl + rr) + (rl * 0.5 + rr * 0.5) / (rl + rr) + 0.5 * (1.0 / (rl + rr)) ** 2)
    assert abs(c2_roe - c2_ref) > 1e-3 * c2_roe
