"""Pins of the oracle's one-sided (wall) Riemann solver (oracle/riemann.c, os_*) and of the wall-cut
face flux against what the mathematics fixes, independently of the solver's own formulas:

* stiffened-gas closed form: the fluid's 1-wave must bring its normal velocity to the wall's,
  f(p*) = u_l - u_wall, with Toro's shock and rarefaction branches for p -> p + p_inf (Newton to
  1e-14). The secant stops at tol_main = 1e-4 of p + rho u^2 / 2 and of the sound speed
  (ExactRiemannSolverBase.cpp:2146-2150), so p* is pinned to that order; the star velocity is the
  wall's exactly; rho* follows the Hugoniot or the isentrope to the same order;
* a wall at rest is the symmetric two-material problem (V, mirrored V): same p* and rho* to the
  two solvers' stopping tolerances;
* Galilean invariance along the normal, the tangential velocity carried unchanged, and rotation:
  the same problem along any unit normal gives the rotated answer;
* the state at the wall: the star state when the wave leaves the wall, id = -1 (nothing written)
  when the wall recedes into the fluid side, and for a transonic rarefaction the sonic state
  u = c of the closed form, to the accuracy of the integrator's linear interpolation;
* vacuum: a wall receding faster than 2c/(gamma-1) gets the reference's fallback (status 1, the
  input moving with the wall);
* GetNormalForOneSidedRiemann's three modes; the wall-face flux reduces to the physical flux when
  the wall moves with the fluid, and with the mesh normal the side -1 face is the mirror image of
  the side +1 face (HLLC, LLF: the generalised Roe flux is not mirror-symmetric in the reference,
  whose entropy-fix threshold starts from the signed u - c, FluxFcnGenRoe.h:186).
A dropped term, a wrong sign or a swapped side gives errors of order one and fails these.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as po

AIR, WATER = (1.4, 0.0), (6.12, 3.43e8)


def oracle_with(mats, flux=None):
    cfg = po.make_config(4, 4, 4, (0, 1, 0, 1, 0, 1), [po.BC_SYMMETRY] * 6, [po.sg(g, pi) for g, pi in mats])
    if flux is not None:
        cfg.flux = flux
    return po.Oracle(cfg)


def wall_closed_form(rho, un, p, uw, g, pinf):
    """(p*, rho*) of the half problem: 1-wave from (rho, un, p) to normal velocity uw."""
    A = 2.0 / ((g + 1.0) * rho)
    B = (g - 1.0) / (g + 1.0) * (p + pinf)
    c = math.sqrt(g * (p + pinf) / rho)

    def f(q):
        if q > p:
            s = math.sqrt(A / (q + pinf + B))
            return (q - p) * s, s * (1.0 - 0.5 * (q - p) / (q + pinf + B))
        e = (g - 1.0) / (2.0 * g)
        r = (q + pinf) / (p + pinf)
        return 2.0 * c / (g - 1.0) * (r ** e - 1.0), 1.0 / (rho * c) * r ** (-(g + 1.0) / (2.0 * g))

    q = p
    for _ in range(200):
        fv, df = f(q)
        qn = q - (fv - (un - uw)) / df
        if qn <= -pinf:
            qn = 0.5 * (q - pinf)
        if abs(qn - q) <= 1e-15 * abs(qn + pinf):
            q = qn
            break
        q = qn
    qs, ps = q + pinf, p + pinf
    if q > p:
        k = (g - 1.0) / (g + 1.0)
        rs = rho * (qs / ps + k) / (k * qs / ps + 1.0)
    else:
        rs = rho * (qs / ps) ** (1.0 / g)
    return q, rs


# (name, (rho, u_normal, p), wall normal velocity, material, tolerances: p* / (p + rho u^2/2), rho*/rho*)
CASES = [
    ("air_piston_shock", (1.0, 0.0, 1.0), -0.5, AIR, (2e-4, 2e-4)),
    ("air_impinging_on_wall", (1.2, 300.0, 1e5), 0.0, AIR, (2e-4, 2e-4)),
    ("air_strong_piston", (1.0, 0.0, 1e5), -1500.0, AIR, (2e-4, 5e-4)),
    ("air_receding_wall", (1.0, 0.0, 1.0), 0.5, AIR, (1e-4, 1e-4)),
    ("air_flow_leaving_wall", (1.2, -100.0, 1e5), 0.0, AIR, (1e-4, 1e-4)),
    ("water_hammer", (1000.0, 20.0, 1e5), 0.0, WATER, (2e-4, 2e-4)),
    ("water_pulled_wall", (1000.0, 0.0, 5e7), 10.0, WATER, (1e-4, 1e-4)),
]


@pytest.mark.parametrize("name,S,uw,mat,tol", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_star_state_matches_stiffened_gas_closed_form(name, S, uw, mat, tol, d):
    o = oracle_with([mat])
    rho, un, p = S
    Vm = np.zeros(5)
    Vm[0], Vm[1 + d], Vm[4] = rho, un, p
    Vm[1 + (d + 1) % 3] = 7.0            # a tangential velocity the answer must carry
    n = np.zeros(3)
    n[d] = 1.0
    Uw = np.zeros(3)
    Uw[d] = uw
    Uw[(d + 2) % 3] = 3.0                 # the wall's tangential motion plays no part
    st, Vs, mid, Vsm = o.riemann_onesided(n, Vm, 0, Uw)
    assert st == 0
    ps, rs = wall_closed_form(rho, un, p, uw, *mat)
    assert abs(Vsm[4] - ps) <= tol[0] * abs(p + 0.5 * rho * un * un)
    assert abs(Vsm[0] - rs) <= tol[1] * rs
    assert Vsm[1 + d] == uw
    assert Vsm[1 + (d + 1) % 3] == 7.0 and Vsm[1 + (d + 2) % 3] == 0.0
    if uw < 0:                            # the wall advanced into the fluid's side of x = 0
        assert mid == -1 and not Vs.any()
    else:
        assert mid == 0
        # a shock or a rarefaction tail moving away from the wall leaves the star state there
        assert np.array_equal(Vs, Vsm)


def test_wall_at_rest_is_the_symmetric_two_material_problem():
    o = oracle_with([AIR, WATER])
    for mid, V in ((0, [1.2, 250.0, 0.0, 0.0, 1e5]), (0, [1.2, -80.0, 0.0, 0.0, 1e5]),
                   (1, [1000.0, 15.0, 0.0, 0.0, 2e6]), (1, [1000.0, -2.0, 0.0, 0.0, 2e6])):
        V = np.array(V)
        M_ = V.copy()
        M_[1] = -V[1]
        st2, _, _, Vsm2, Vsp2 = o.riemann(np.array([1.0, 0, 0]), V, mid, M_, mid)
        st1, Vs1, id1, Vsm1 = o.riemann_onesided(np.array([1.0, 0, 0]), V, mid, np.zeros(3))
        assert st1 == 0 and st2 == 0
        # both secants stop on a velocity mismatch below 1e-4 c, worth rho c^2 1e-4 in pressure
        c2 = o.c2(V[0], V[4], mid)
        scale = max(V[4] + 0.5 * V[0] * V[1] ** 2, V[0] * c2)
        assert abs(Vsm1[4] - Vsm2[4]) <= 2e-4 * scale
        assert abs(Vsm1[0] - Vsm2[0]) <= 2e-4 * Vsm2[0]
        assert Vsm1[1] == 0.0 and abs(Vsm2[1]) <= 1e-4 * math.sqrt(c2)


def test_galilean_invariance_and_rotation():
    o = oracle_with([AIR])
    V = np.array([1.1, 120.0, 0.0, 0.0, 9e4])
    base = o.riemann_onesided(np.array([1.0, 0, 0]), V, 0, np.array([-60.0, 0, 0]))
    # the same relative motion at a different frame velocity
    Vb = V.copy()
    Vb[1] += 40.0
    shifted = o.riemann_onesided(np.array([1.0, 0, 0]), Vb, 0, np.array([-20.0, 0, 0]))
    assert abs(shifted[3][4] - base[3][4]) <= 1e-4 * (V[4] + 0.5 * V[0] * V[1] ** 2)
    assert abs(shifted[3][0] - base[3][0]) <= 1e-4 * base[3][0]
    # the same problem along an oblique unit normal
    n = np.array([2.0, -1.0, 2.0]) / 3.0
    Vr = V.copy()
    Vr[1:4] = 120.0 * n
    rot = o.riemann_onesided(n, Vr, 0, -60.0 * n)
    assert rot[0] == 0
    assert abs(rot[3][4] - base[3][4]) <= 1e-4 * (V[4] + 0.5 * V[0] * V[1] ** 2)
    assert abs(rot[3][0] - base[3][0]) <= 1e-4 * base[3][0]
    assert np.allclose(rot[3][1:4], -60.0 * n, rtol=0, atol=1e-12)


def test_transonic_rarefaction_gives_the_sonic_state():
    g = 1.4
    o = oracle_with([AIR])
    rho, p = 1.0, 1.0
    c = math.sqrt(g * p / rho)
    st, Vs, mid, Vsm = o.riemann_onesided(np.array([1.0, 0, 0]), np.array([rho, 0, 0, 0, p]), 0,
                                          np.array([3.0, 0, 0]))
    assert st == 0 and mid == 0
    # u - c = 0 inside the fan: u = 2c/(g+1), rho and p on the isentrope
    u0 = 2.0 * c / (g + 1.0)
    r0 = rho * (2.0 / (g + 1.0)) ** (2.0 / (g - 1.0))
    p0 = p * (2.0 / (g + 1.0)) ** (2.0 * g / (g - 1.0))
    assert abs(Vs[1] - u0) <= 2e-3 * u0
    assert abs(Vs[0] - r0) <= 2e-3 * r0
    assert abs(Vs[4] - p0) <= 2e-3 * p0
    # the tail state after a 140-fold expansion: the reference's Cash-Karp step places its third node
    # at p_0 + dp/2 (ExactRiemannSolverBase.cpp:1763), which the restatement keeps, so deep
    # expansions are accurate to about 2e-2 (shallow ones, above, to 1e-4)
    ps, rs = wall_closed_form(rho, 0.0, p, 3.0, g, 0.0)
    assert abs(Vsm[4] - ps) <= 3e-2 * ps and abs(Vsm[0] - rs) <= 3e-2 * rs


def test_vacuum_returns_the_input_moving_with_the_wall():
    o = oracle_with([AIR])
    V = np.array([1.0, 0.0, 2.0, 0.0, 1.0])
    c = math.sqrt(1.4)
    uw = 2.0 * c / 0.4 * 1.5
    st, Vs, mid, Vsm = o.riemann_onesided(np.array([1.0, 0, 0]), V, 0, np.array([uw, 0, 0]))
    assert st == 1 and mid == 0
    assert np.array_equal(Vsm, [1.0, uw, 2.0, 0.0, 1.0]) and np.array_equal(Vs, Vsm)
    # receding along -x: the wall velocity -uw is +uw along the normal
    st, Vs, mid, Vsm = o.riemann_onesided(np.array([-1.0, 0, 0]), V, 0, np.array([-uw, 0, 0]))
    assert st == 1 and mid == 0 and Vsm[1] == -uw


def test_trivial_problem_returns_the_input():
    o = oracle_with([AIR])
    V = np.array([1.3, 5.0, -2.0, 1.0, 7e4])
    st, Vs, mid, Vsm = o.riemann_onesided(np.array([1.0, 0, 0]), V, 0, np.array([5.0, 9.0, 9.0]))
    assert st == 0 and mid == 0 and np.array_equal(Vsm, V) and np.array_equal(Vs, V)


def test_onesided_normal_modes():
    o = oracle_with([AIR])
    nw = np.array([-0.6, 0.0, 0.8])
    assert np.array_equal(o.onesided_normal(0, 1, 1, nw), [1.0, 0.0, 0.0])
    assert np.array_equal(o.onesided_normal(2, -1, 1, nw), [0.0, 0.0, -1.0])
    # the wall normal with only the face-axis component turned to point along the side
    assert np.array_equal(o.onesided_normal(0, 1, 0, nw), [0.6, 0.0, 0.8])
    assert np.array_equal(o.onesided_normal(2, 1, 0, nw), [-0.6, 0.0, 0.8])
    assert np.array_equal(o.onesided_normal(2, -1, 0, nw), [-0.6, 0.0, -0.8])
    a = o.onesided_normal(0, 1, 2, nw)
    expect = np.array([1.6, 0.0, 0.8]) / math.hypot(1.6, 0.8)
    assert np.allclose(a, expect, rtol=0, atol=1e-15)


@pytest.mark.parametrize("flux", [po.FLUX_HLLC, po.FLUX_ROE, po.FLUX_LLF])
def test_wall_face_flux(flux):
    o = oracle_with([AIR], flux)
    V = np.array([1.2, 30.0, -4.0, 2.0, 1e5])
    # a wall moving with the fluid: no wave, F(V, V) is the physical flux
    st, F, Vst = o.wall_face_flux(0, 1, 1, np.array([-1.0, 0, 0]), V[1:4], V)
    assert st == 0 and np.array_equal(Vst, V)
    E = V[4] / 0.4 + 0.5 * V[0] * (V[1:4] ** 2).sum()
    phys = [V[0] * V[1], V[0] * V[1] ** 2 + V[4], V[0] * V[1] * V[2], V[0] * V[1] * V[3], V[1] * (E + V[4])]
    assert np.allclose(F, phys, rtol=1e-14, atol=0)
    # mesh normal: the face with the fluid on the plus side is the mirror image of the minus side
    W = np.array([1.2, 50.0, -4.0, 2.0, 1e5])
    Wm = W.copy()
    Wm[1] = -W[1]
    st1, F1, S1 = o.wall_face_flux(0, 1, 1, np.zeros(3), np.array([-10.0, 0, 0]), W)
    st2, F2, S2 = o.wall_face_flux(0, -1, 1, np.zeros(3), np.array([10.0, 0, 0]), Wm)
    assert st1 == 0 and st2 == 0
    assert np.allclose(S2 * [1, -1, 1, 1, 1], S1, rtol=1e-14, atol=0)
    if flux != po.FLUX_ROE:   # the Roe entropy fix's threshold starts from the signed u - c
        assert np.allclose(F2 * [-1, 1, -1, -1, -1], F1, rtol=1e-12, atol=1e-9)   # (FluxFcnGenRoe.h:186)
    # the reflected mass flux is small next to the incoming one (the wall is impermeable up to the
    # numerical flux's dissipation) and the momentum flux carries the wall pressure
    assert abs(F1[0]) < 0.5 * W[0] * W[1] and F1[1] > W[4]
