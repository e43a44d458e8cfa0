"""Pins of the oracle's equations of state (oracle/eos.h) by mathematics, not by retyping them.

Each check is chosen so that a plausible slip in a formula (a dropped term, a wrong exponent,
rho0 where rho belongs, a sign) fails it:
* special cases that reduce one law to another: NASG with b = q = 0 is the stiffened gas
  (VarFcnNASG.h:55-59 vs VarFcnSG.h:64-68), JWL with A1 = A2 = 0 is an ideal gas with
  gamma - 1 = omega (VarFcnJWL.h:49-53), Mie-Grueneisen at rho = rho0 is linear in e
  (VarFcnMG.h:64-81);
* shock physics: the Mie-Grueneisen pressure on its own principal Hugoniot (linear Us-up relation
  Us = c0 + s up, Rankine-Hugoniot energy jump up^2/2) equals rho0 Us up;
* the textbook JWL form in relative volume (Dobratz & Crawford, LLNL Explosives Handbook):
  p = A1 (1 - omega/(R1 v)) e^(-R1 v) + A2 (1 - omega/(R2 v)) e^(-R2 v) + omega E / v, v = rho0/rho,
  E = rho0 e;
* the NASG sound speed of Le Metayer & Saurel (2016): c^2 = gamma v^2 (p + p_inf) / (v - b);
* derivatives: dp/drho at constant e and Gamma = (1/rho) dp/de against central differences of
  p(rho, e); c^2 against the isentropic derivative dp/drho along de = (p/rho^2) drho, also by
  central differences (VarFcnBase.h:215-219);
* inverses: p(rho, e(rho, p)) = p;
* CheckState (VarFcnSG.h:96-108, VarFcnNASG.h:73-85, VarFcnBase.h:113-130) on the boundary of the
  valid region.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as po

WATER_NASG = dict(gam=1.19, pc=6.217e8, b=6.7212e-4, q=-1177788.0)
ALU_MG = dict(rho0=2785.0, c0=5328.0, Gamma0=2.0, s=1.338, e0=0.0)
TNT_JWL = dict(omega=0.28, A1=3.712e11, A2=3.23e9, R1=4.15, R2=0.95, rho0=1630.0)


def one(m):
    cfg = po.make_config(2, 2, 2, (0, 1, 0, 1, 0, 1), [po.BC_SYMMETRY] * 6, [m])
    return po.Oracle(cfg)


def rel(a, b):
    return abs(a - b) / max(abs(a), abs(b), 1e-300)


# (material, states (rho, e)) chosen inside each law's valid region
CASES = {
    "sg_water": (po.sg(6.12, 3.43e8), [(1000.0, 4.2e5), (1200.0, 5.0e5), (980.0, 4.25e5)]),
    "nasg": (po.eos(po.EOS_NASG, **WATER_NASG), [(1000.0, 3.0e5), (1100.0, 6.0e5), (950.0, 1.0e5)]),
    "mg": (po.eos(po.EOS_MG, **ALU_MG), [(2785.0, 1.0e5), (3000.0, 4.0e5), (3500.0, 2.0e6), (2600.0, 5.0e4)]),
    "jwl": (po.eos(po.EOS_JWL, **TNT_JWL), [(1630.0, 4.3e6), (1000.0, 3.0e6), (300.0, 1.0e6), (2000.0, 6.0e6)]),
}


def test_nasg_with_b_q_zero_is_stiffened_gas():
    nasg = one(po.eos(po.EOS_NASG, gam=4.4, pc=6e8, b=0.0, q=0.0))
    sg = one(po.sg(4.4, 6e8))
    for rho, e in [(1000.0, 2.0e5), (1100.0, 7.5e5), (900.0, 1.4e5)]:
        a = nasg.eos_point(rho, e, 1e7)
        b = sg.eos_point(rho, e, 1e7)
        # (gamma-1)(e-q)/(1/rho - b) and (gamma-1) rho e round differently: a few ulp
        for t in range(5):
            assert rel(a[t], b[t]) < 4e-15, (t, a[t], b[t])
        assert a[5] == b[5]


def test_jwl_without_exponentials_is_ideal_gas():
    jwl = one(po.eos(po.EOS_JWL, omega=0.4, A1=0.0, A2=0.0, R1=4.15, R2=0.95, rho0=1630.0))
    ig = one(po.sg(1.4, 0.0))
    for rho, e in [(1.2, 2.5e5), (10.0, 1.0e6), (0.3, 3.0e4)]:
        a = jwl.eos_point(rho, e, 8e4)
        b = ig.eos_point(rho, e, 8e4)
        assert a[0] == 0.4 * rho * e
        for t in range(5):
            assert rel(a[t], b[t]) < 4e-15, (t, a[t], b[t])


def test_mie_grueneisen_at_reference_density_is_linear_in_e():
    o = one(po.eos(po.EOS_MG, **dict(ALU_MG, e0=1.5e4)))
    rho0, G0 = ALU_MG["rho0"], ALU_MG["Gamma0"]
    for e in [0.0, 1.5e4, 2.0e5, -3.0e4]:
        p = o.eos_point(rho0, e, 0.0)[0]
        assert p == G0 * rho0 * (e - 1.5e4)


def test_mie_grueneisen_principal_hugoniot():
    o = one(po.eos(po.EOS_MG, **ALU_MG))
    rho0, c0, s = ALU_MG["rho0"], ALU_MG["c0"], ALU_MG["s"]
    for up in [50.0, 300.0, 1000.0, 2500.0]:
        Us = c0 + s * up
        rho = rho0 * Us / (Us - up)
        pH = rho0 * Us * up
        eH = 0.5 * up * up          # e - e0 = p_H (1/rho0 - 1/rho) / 2 from rest with p0 = 0
        assert rel(o.eos_point(rho, eH, 0.0)[0], pH) < 1e-12


def test_jwl_textbook_form():
    o = one(po.eos(po.EOS_JWL, **TNT_JWL))
    A, B, R1, R2, w, rho0 = (TNT_JWL[k] for k in ("A1", "A2", "R1", "R2", "omega", "rho0"))
    for rho, e in CASES["jwl"][1]:
        v = rho0 / rho
        E = rho0 * e
        p = A * (1 - w / (R1 * v)) * math.exp(-R1 * v) + B * (1 - w / (R2 * v)) * math.exp(-R2 * v) + w * E / v
        assert rel(o.eos_point(rho, e, 0.0)[0], p) < 1e-13


def test_nasg_sound_speed_closed_form():
    o = one(po.eos(po.EOS_NASG, **WATER_NASG))
    g, pinf, b = WATER_NASG["gam"], WATER_NASG["pc"], WATER_NASG["b"]
    for rho, e in CASES["nasg"][1]:
        p = o.eos_point(rho, e, 0.0)[0]
        v = 1.0 / rho
        c2 = g * v * v * (p + pinf) / (v - b)
        assert rel(o.eos_point(rho, e, 0.0)[4], c2) < 1e-12


@pytest.mark.parametrize("name", list(CASES))
def test_derivatives_match_finite_differences(name):
    m, states = CASES[name]
    o = one(m)

    def P(r, e):
        return o.eos_point(r, e, 0.0)[0]

    for rho, e in states:
        out = o.eos_point(rho, e, 0.0)
        p, dpdrho, G, c2 = out[0], out[2], out[3], out[4]
        hr, he = 1e-5 * rho, 1e-5 * abs(e)
        fd_rho = (P(rho + hr, e) - P(rho - hr, e)) / (2 * hr)
        fd_e = (P(rho, e + he) - P(rho, e - he)) / (2 * he)
        # scale of the pressure terms that the differences cancel
        scale = max(abs(p), abs(P(rho, 0.0)), 1.0)
        assert abs(dpdrho - fd_rho) <= 1e-7 * max(abs(dpdrho), scale / rho), (dpdrho, fd_rho)
        assert abs(G - fd_e / rho) <= 1e-7 * abs(G), (G, fd_e / rho)
        # isentropic derivative: de = (p / rho^2) drho
        s = p / (rho * rho)
        fd_s = (P(rho + hr, e + s * hr) - P(rho - hr, e - s * hr)) / (2 * hr)
        assert abs(c2 - fd_s) <= 1e-7 * max(abs(c2), scale / rho), (c2, fd_s)
        assert c2 > 0


@pytest.mark.parametrize("name", list(CASES))
def test_energy_inverts_pressure(name):
    m, states = CASES[name]
    o = one(m)
    for rho, e in states:
        p = o.eos_point(rho, e, 0.0)[0]
        e2 = o.eos_point(rho, 0.0, p)[1]
        assert rel(e2, e) < 1e-9 * max(1.0, abs(p) / (rho * abs(e)))
        assert rel(o.eos_point(rho, e2, 0.0)[0], p) < 1e-12 * max(1.0, abs(p - o.eos_point(rho, 0.0, 0.0)[0]) / abs(p)) + 1e-12


def test_check_state_boundaries():
    sg = one(po.sg(6.12, 3.43e8))
    assert sg.eos_point(1000.0, 0.0, -3.43e8)[5] == 1.0       # p + p_inf = 0
    assert sg.eos_point(1000.0, 0.0, -3.42e8)[5] == 0.0
    assert sg.eos_point(0.0, 0.0, 1e5)[5] == 1.0
    assert sg.eos_point(1000.0, 0.0, float("nan"))[5] == 1.0
    nasg = one(po.eos(po.EOS_NASG, **WATER_NASG))
    assert nasg.eos_point(1000.0, 0.0, -6.217e8)[5] == 1.0
    assert nasg.eos_point(1000.0, 0.0, -6.2e8)[5] == 0.0
    # generic rule: invalid where c^2 <= 0; Mie-Grueneisen in strong tension has c^2 < 0
    mg = one(po.eos(po.EOS_MG, **ALU_MG))
    assert mg.eos_point(2785.0, 0.0, 1e5)[5] == 0.0
    rho_t = 1200.0
    e_t = mg.eos_point(rho_t, 0.0, -4e9)[1]
    assert (mg.eos_point(rho_t, e_t, 0.0)[4] <= 0) == (mg.eos_point(rho_t, 0.0, -4e9)[5] == 1.0)
