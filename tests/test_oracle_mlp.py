"""CPU checks of the learned-solver oracle (oracle/mlp_ref.py).

The network gradient is pinned against central finite differences, bf16 rounding against torch's
independent conversion, the closed-form stiffened-gas solver against Toro's published star states
(Table 4.2) and against the oracle's restatement of the reference's exact two-material solver, and
the xi = 0 sampling against the sonic-point invariant of a transonic fan.
"""
import numpy as np
import pytest

from oracle import mlp_ref as R
from oracle import pyoracle as po


def rand_params(seed):
    rng = np.random.default_rng(seed)
    p = np.empty(R.NPARAM)
    for (wo, bo), (no, ni) in zip(R.OFFS, R.DIMS):
        a = 1.0 / np.sqrt(ni)
        p[wo:bo + no] = rng.uniform(-a, a, size=bo + no - wo)
    return p


def sample_pairs(n, seed):
    """Random stiffened-gas pairs drawn like the GPU sample generator."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        liq = rng.random(2) < 0.5
        gam = [4.0 + 3.5 * rng.random() if q else 1.1 + 0.6 * rng.random() for q in liq]
        pinf = [10.0 ** (8.0 + rng.random()) if q else 0.0 for q in liq]
        pL = 10.0 ** (4.0 + 5.0 * rng.random())
        pR = pL * 10.0 ** (4.0 * (2.0 * rng.random() - 1.0))
        rL = 10.0 ** (-1.0 + 4.0 * rng.random())
        rR = rL * 10.0 ** (3.0 * (2.0 * rng.random() - 1.0))
        L = R.Side(rL, 0.0, pL, gam[0], pinf[0])
        Rr = R.Side(rR, 0.0, pR, gam[1], pinf[1])
        cs = L.c + Rr.c
        uL = cs * (rng.random() - 0.5)
        uR = uL + cs * (2.0 * rng.random() - 1.0)
        out.append((R.Side(rL, uL, pL, gam[0], pinf[0]), R.Side(rR, uR, pR, gam[1], pinf[1])))
    return out


def test_parameter_layout():
    assert R.NPARAM == 200963
    W, b = R.unpack(np.arange(R.NPARAM, dtype=np.float64))
    assert [w.shape for w in W] == [(256, 10), (256, 256), (256, 256), (256, 256), (3, 256)]
    assert np.array_equal(R.pack(W, b), np.arange(R.NPARAM))


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.normal(size=4000).astype(np.float32) * 10.0 ** rng.integers(-30, 30, 4000),
                        np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, -1.0 - 2.0 ** -8, 0.0, -0.0, 1e-40,
                                  3.4e38, -3.4e38], dtype=np.float32)]).astype(np.float32)
    ours = R.bf16(x)
    theirs = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy().astype(np.float64)
    assert np.array_equal(ours, theirs)


def test_gradient_matches_finite_differences():
    p = rand_params(11)
    rng = np.random.default_rng(5)
    X = rng.normal(size=(8, 10))
    T = rng.normal(size=(8, 3))
    T[3] = np.nan   # a sample without a solution contributes nothing
    loss, g = R.loss_grad(p, X, T, emulate=False)
    idx = []
    for (wo, bo), (no, ni) in zip(R.OFFS, R.DIMS):
        idx += list(rng.integers(wo, bo, size=8)) + list(rng.integers(bo, bo + no, size=3))
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (R.loss_grad(pp, X, T)[0] - R.loss_grad(pm, X, T)[0]) / (2 * h)
        assert abs(g[i] - fd) <= 1e-6 * max(1.0, abs(fd)), (i, g[i], fd)


def test_emulated_gradient_close_to_exact():
    p = rand_params(12)
    rng = np.random.default_rng(6)
    X = rng.normal(size=(64, 10))
    T = rng.normal(size=(64, 3))
    l0, g0 = R.loss_grad(p, X, T, emulate=False)
    l1, g1 = R.loss_grad(p, X, T, emulate=True)
    assert abs(l1 - l0) <= 2e-2 * l0
    assert np.linalg.norm(g1 - g0) <= 5e-2 * np.linalg.norm(g0)


# Toro, Table 4.2 (gamma = 1.4): left, right, p*, u*
TORO = [((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 0.30313, 0.92745),
        ((1.0, -2.0, 0.4), (1.0, 2.0, 0.4), 0.00189, 0.0),
        ((1.0, 0.0, 1000.0), (1.0, 0.0, 0.01), 460.894, 19.5975),
        ((1.0, 0.0, 0.01), (1.0, 0.0, 100.0), 46.0950, -6.19633),
        ((5.99924, 19.5975, 460.894), (5.99242, -6.19633, 46.0950), 1691.64, 8.68975)]


@pytest.mark.parametrize("case", range(5))
def test_closed_form_matches_toro_table(case):
    (l, r, p_ref, u_ref) = TORO[case]
    L, Rr = R.Side(*l, 1.4, 0.0), R.Side(*r, 1.4, 0.0)
    ps, us, rl, rr = R.sg_star(L, Rr)
    assert abs(ps - p_ref) <= 5e-5 * p_ref + 1e-5
    assert abs(us - u_ref) <= 5e-5 * max(abs(u_ref), 1.0)
    if case == 0:   # Toro's star densities of test 1
        assert abs(rl - 0.42632) < 1e-5 and abs(rr - 0.26557) < 1e-5


def test_closed_form_matches_exact_solver_restatement():
    """The closed form against the oracle's restatement of ExactRiemannSolverBase (200 random
    stiffened-gas pairs). The exact solver stops at relative changes of tol_main = 1e-4
    (ExactRiemannSolverBase.cpp:300-310) and integrates rarefactions with its adaptive Runge-Kutta
    scheme (:1729-1852, third node at p + 3/10 dp), which departs from the closed-form isentrope by up
    to ~1e-2 -- and by much more near vacuum (P*/P < 1e-2 on a side), left out here."""
    pairs = sample_pairs(200, 21)
    eus = []
    for L, Rr in pairs:
        mats = [po.Eos(), po.Eos()]
        for m, sd in zip(mats, (L, Rr)):
            m.type = po.EOS_SG
            m.rhomin, m.pmin, m.failsafe_density = 0.0, -1.7976931348623157e308, 1e-10
            m.rhomax = m.pmax = 1.7976931348623157e308
            m.gam, m.pstiff = sd.gam, sd.pinf
        o = po.Oracle(po.make_config(4, 1, 1, (0, 4, 0, 1, 0, 1), [5] * 6, mats))
        st, Vs, mid, Vsm, Vsp = o.riemann([1.0, 0.0, 0.0], [L.rho, L.u, 0, 0, L.p], 0, [Rr.rho, Rr.u, 0, 0, Rr.p], 1)
        o.close()
        if st != 0 or min((Vsm[4] + L.pinf) / L.P, (Vsm[4] + Rr.pinf) / Rr.P) < 1e-2:
            continue
        ps, us, rl, rr = R.sg_star(L, Rr)
        pscale = max(abs(L.p + 0.5 * L.rho * L.u ** 2), abs(Rr.p + 0.5 * Rr.rho * Rr.u ** 2))
        assert abs(Vsm[4] - ps) <= 2e-3 * pscale, (L.__dict__, Rr.__dict__, Vsm, ps)
        eu = abs(Vsm[1] - us) / max(L.c, Rr.c)
        assert eu <= 1e-2
        eus.append(eu)
        # star densities from the wave-curve relations at the exact solver's own p*: Hugoniot roots
        # (tol_shock 1e-12) on shock sides, the integrated isentrope on rarefaction sides
        for side, rho in ((L, Vsm[0]), (Rr, Vsp[0])):
            bound = 1e-9 if Vsm[4] > side.p else 2e-2
            assert abs(rho - R.star_rho(side, Vsm[4])) <= bound * rho
    assert len(eus) >= 120
    assert np.median(eus) <= 1e-4


def test_newton_from_acoustic_guess_converges():
    """Pairs with a star state (no cavitation: F(p) < 0 as p + min p_inf -> 0) converge from the
    acoustic estimate; the others have no root and are left to the exact solver."""
    pairs = sample_pairs(300, 22)
    acc = n = 0
    for L, Rr in pairs:
        pf = -min(L.pinf, Rr.pinf) + 1e-12 * min(L.P, Rr.P)
        if R.wave(L, pf)[0] + R.wave(Rr, pf)[0] + Rr.u - L.u >= 0.0:
            continue
        n += 1
        x, sc = R.features(L, Rr)
        p_ac = np.exp(4.0 * float(x[9])) * sc[0] - sc[1]
        good, ps, it = R.newton(L, Rr, p_ac, 60, 1e-10)
        acc += good
        if good:
            assert R.residual(L, Rr, ps) <= 1e-10
    assert n >= 200 and acc >= 0.97 * n


def test_targets_decode_to_star_pressure():
    for L, Rr in sample_pairs(50, 23):
        ps, us, rl, rr = R.sg_star(L, Rr)
        if ps + min(L.pinf, Rr.pinf) <= 0:
            continue
        t = R.targets(L, Rr, ps, us, rl)
        _, sc = R.features(L, Rr)
        assert abs(R.decode_p(sc, t[0]) - ps) <= 1e-5 * (abs(ps) + min(L.pinf, Rr.pinf)) * 10


def test_sample0_sonic_point_of_transonic_fan():
    # left rarefaction whose fan straddles x/t = 0: at the sonic point u = c (Toro 4.56)
    L = R.Side(1.0, 0.75, 1.0, 1.4, 0.0)
    Rr = R.Side(0.125, 0.0, 0.1, 1.4, 0.0)
    ps, us, rl, rr = R.sg_star(L, Rr)
    assert L.u - L.c < 0 < us - L.c * (ps / L.P) ** (0.4 / 2.8)
    rho, u, p = R.sg_sample0(L, Rr, ps, us, rl, rr)
    c = np.sqrt(1.4 * p / rho)
    assert abs(u - c) <= 1e-12 * c
    assert abs(p / rho ** 1.4 - L.p / L.rho ** 1.4) <= 1e-12   # isentropic
    # stiffened gas: the same invariants in effective pressure
    L = R.Side(1000.0, 1200.0, 1e5, 6.12, 3.43e8)
    Rr = R.Side(1000.0, 3000.0, 1e5, 6.12, 3.43e8)
    ps, us, rl, rr = R.sg_star(L, Rr)
    rho, u, p = R.sg_sample0(L, Rr, ps, us, rl, rr)
    if L.u - L.c < 0 < us - L.c * ((ps + L.pinf) / L.P) ** (5.12 / 12.24):
        c = np.sqrt(6.12 * (p + L.pinf) / rho)
        assert abs(u - c) <= 1e-9 * c


def test_sample0_uniform_flow():
    L = R.Side(1.0, 0.3, 1.0, 1.4, 0.0)
    ps, us, rl, rr = R.sg_star(L, L)
    assert abs(ps - 1.0) < 1e-12 and abs(us - 0.3) < 1e-12
    assert np.allclose(R.sg_sample0(L, L, ps, us, rl, rr), (1.0, 0.3, 1.0), rtol=1e-12)
